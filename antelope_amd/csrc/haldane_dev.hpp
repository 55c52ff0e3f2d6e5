// haldane_dev.hpp - device helpers shared by the 2-band (Haldane) kernels:
// wave-level reductions, the Bloch-vector state of one k point and the time
// sample of the currents (antelope.cpp:307-333 with SBEs.h:694-760).
#pragma once

#include "sbe_internal.hpp"

namespace sbe {
namespace dev {

// ------------------------------------------- RK tableaus (SBEs.h:395-416)
template <int RK> struct Coef;
template <> struct Coef<4> {
    static constexpr int S = 4;
    static constexpr double a[4][4] = {{0, 0, 0, 0}, {0.5, 0, 0, 0}, {0, 0.5, 0, 0}, {0, 0, 1.0, 0}};
    static constexpr double b[4] = {1.0 / 6.0, 2.0 / 6.0, 2.0 / 6.0, 1.0 / 6.0};
    static constexpr double c[4] = {0.0, 0.5, 0.5, 1.0};
};
template <> struct Coef<5> {
    static constexpr int S = 6;
    static constexpr double a[6][6] = {
        {0, 0, 0, 0, 0, 0},
        {0.25, 0, 0, 0, 0, 0},
        {3. / 32., 9. / 32., 0, 0, 0, 0},
        {1932. / 2197., -7200. / 2197., 7296. / 2197., 0, 0, 0},
        {439. / 216., -8., 3680. / 513., -845. / 4104., 0, 0},
        {-8. / 27., 2., -3544. / 2565., 1859. / 4104., -11. / 40., 0}};
    static constexpr double b[6] = {16. / 135., 0, 6656. / 12825., 28561. / 56430., -9. / 50., 2. / 55.};
    static constexpr double c[6] = {0.0, 0.25, 3. / 8., 12. / 13., 1.0, 0.5};
};

// Stage bookkeeping of a tableau: the stage at c = 1 (the next step's sample
// time), or -1; for each stage, an earlier stage (>= 1) at the same c, or -1.
template <int RK>
__host__ __device__ constexpr int stage_at_one()
{
    for (int i = 1; i < Coef<RK>::S; ++i)
        if (Coef<RK>::c[i] == 1.0) return i;
    return -1;
}
template <int RK>
__host__ __device__ constexpr int same_time_stage(int i)
{
    for (int j = 1; j < i; ++j)
        if (Coef<RK>::c[j] == Coef<RK>::c[i]) return j;
    return -1;
}

// --------------------------------------------------- wave-level reduction
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}

// a <- rows {0,1} of a and rows {0,1} of b; b <- rows {2,3} of a and b
__device__ __forceinline__ void swap32(double &a, double &b)
{
    const auto lo = __builtin_amdgcn_permlane32_swap(__double2loint(a), __double2loint(b), false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap(__double2hiint(a), __double2hiint(b), false, false);
    a = __hiloint2double(hi[0], lo[0]);
    b = __hiloint2double(hi[1], lo[1]);
}

// a <- even rows of a and b; b <- odd rows of a and b
__device__ __forceinline__ void swap16(double &a, double &b)
{
    const auto lo = __builtin_amdgcn_permlane16_swap(__double2loint(a), __double2loint(b), false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap(__double2hiint(a), __double2hiint(b), false, false);
    a = __hiloint2double(hi[0], lo[0]);
    b = __hiloint2double(hi[1], lo[1]);
}

// Sum each of v[0..7] over the 64 lanes of the wave (reduce-scatter):
// lane l ends with the total of v[(l >> 3) & 7]. Requires all 64 lanes.
__device__ __forceinline__ double wave_reduce8(double v[8])
{
    const int lane = __lane_id();
    // lane bit 5: permlane32 swap + add
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        double a = v[j], b = v[j + 4];
        swap32(a, b);
        v[j] = a + b;  // lanes 0-31: v[j], lanes 32-63: v[j+4]
    }
    // lane bit 4: permlane16 swap + add
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        double a = v[j], b = v[j + 2];
        swap16(a, b);
        v[j] = a + b;  // even rows: v[j], odd rows: v[j+2]
    }
    // lane bit 3: rotate by 8 inside each 16-lane row (= lane ^ 8)
    const bool hi = (lane & 8) != 0;
    const double keep = hi ? v[1] : v[0];
    const double send = hi ? v[0] : v[1];
    double x = keep + dpp_f64<0x128>(send);  // row_ror:8
    // lanes 8q..8q+7 now share value index q: all-reduce over the low 3 bits
    x += dpp_f64<0x141>(x);  // row_half_mirror (i <-> 7-i)
    x += dpp_f64<0x4E>(x);   // quad_perm [2,3,0,1]
    x += dpp_f64<0xB1>(x);   // quad_perm [1,0,3,2]
    return x;
}

// -------------------------------------------------------- model kernels
template <typename R>
struct Vec3 {
    R x, y, z;
};

// d(k+A) from the six phases e_j = e^{i(k+A).v_j} (a_0..a_2 then b_0..b_2)
template <typename R>
__device__ __forceinline__ Vec3<R> dvec(const R er[6], const R ei[6], R t1, R c3, R M0)
{
    Vec3<R> d;
    d.x = t1 * (er[0] + er[1] + er[2]);
    d.y = t1 * (ei[0] + ei[1] + ei[2]);
    d.z = M0 - c3 * (ei[3] + ei[4] + ei[5]);
    return d;
}

template <typename R> struct StateT;
template <> struct StateT<double> {
    using T = double;
    static constexpr int kComp = 0;
};
template <> struct StateT<float> {
    using T = float;
    static constexpr int kComp = 1;
};

// explicit fused multiply-add: the same inputs give the same bits wherever
// it is inlined (no contraction left to the compiler)
__device__ __forceinline__ float fmar(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmar(double a, double b, double c) { return __builtin_fma(a, b, c); }

// 1/x from the hardware reciprocal with Newton steps (two in FP64, one in FP32)
__device__ __forceinline__ double rcp_nr(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    r = fmar(r, fmar(-x, r, 1.0), r);
    return fmar(r, fmar(-x, r, 1.0), r);
}
__device__ __forceinline__ float rcp_nr(float x)
{
    const float r = __builtin_amdgcn_rcpf(x);
    return fmar(r, fmar(-x, r, 1.0f), r);
}
// 1/sqrt(x): the FP32 hardware rsqrt (1 ulp), the library one in FP64
__device__ __forceinline__ double rsqrt_r(double x) { return rsqrt(x); }
__device__ __forceinline__ float rsqrt_r(float x) { return __builtin_amdgcn_rsqf(x); }

// complex product (a + ib)(c + id)
template <typename R>
__device__ __forceinline__ void cmul(R a, R b, R c, R d, R &re, R &im)
{
    re = a * c - b * d;
    im = a * d + b * c;
}

// Per-k registers of one lane.
template <typename R>
struct KPoint {
    R er[6], ei[6];     // e^{i k.v_j} (a_0..a_2, b_0..b_2)
    R n, sx, sy, sz;    // Bloch state
    R cn, cx, cy, cz;   // Kahan compensation (FP32 state only)
    double w;
};

// Everything a k point needs at one time t: the phases e^{i(k+A(t)).v_j},
// d(k+A), |d|^2 and 1/|d|^2 (0 at a gap closing). One function computes it
// wherever it is needed, with explicit FMAs, so equal inputs give equal bits.
template <typename R>
struct Phases {
    R ar[3], ai[3];  // t1 e^{i(k+A).a_j}
    R br[3], bi[3];  // e^{i(k+A).b_j}
    Vec3<R> d;
    R dd, inv;
};

template <typename R>
__device__ __forceinline__ Phases<R> phases_at(const KPoint<R> &p, const R *row, R c3, R M0)
{
    Phases<R> e;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        e.ar[j] = fmar(p.er[j], row[2 * j], -(p.ei[j] * row[2 * j + 1]));
        e.ai[j] = fmar(p.er[j], row[2 * j + 1], p.ei[j] * row[2 * j]);
        e.br[j] = fmar(p.er[j + 3], row[6 + 2 * j], -(p.ei[j + 3] * row[7 + 2 * j]));
        e.bi[j] = fmar(p.er[j + 3], row[7 + 2 * j], p.ei[j + 3] * row[6 + 2 * j]);
    }
    e.d.x = e.ar[0] + e.ar[1] + e.ar[2];
    e.d.y = e.ai[0] + e.ai[1] + e.ai[2];
    e.d.z = fmar(-c3, e.bi[0] + e.bi[1] + e.bi[2], M0);
    e.dd = fmar(e.d.x, e.d.x, fmar(e.d.y, e.d.y, e.d.z * e.d.z));
    e.inv = e.dd > R(0) ? rcp_nr(e.dd) : R(0);
    return e;
}

// Time sample of one k point at t_n (antelope.cpp:307-333 with
// SBEs.h:694-760), from the state nn, s and the phases at t_n: adds
// w x (inter_x, inter_y, intra_x, intra_y, s_x, s_y, s_z, s.dhat) to v (FP64).
// Evaluated in the state precision R; the k sum is always FP64. Branch-free
// so the scheduler can interleave it with the RK chain.
template <typename R>
__device__ __forceinline__ void sample_state(const Phases<R> &e, double w, R nn, R s_x, R s_y, R s_z,
                                             const HaldaneArgs &a, R ii, double v[8])
{
    const R c3 = static_cast<R>(a.c3);
    const Vec3<R> &d = e.d;
    // derivatives d/dk_x, d/dk_y of (d0, dx, dy, dz)
    R xx = 0, xy = 0, yx = 0, yy = 0, zx = 0, zy = 0, ox = 0, oy = 0;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const R ax = static_cast<R>(a.va[j][0]), ay = static_cast<R>(a.va[j][1]);
        const R bx = static_cast<R>(a.vb[j][0]), by = static_cast<R>(a.vb[j][1]);
        xx -= ax * e.ai[j];
        xy -= ay * e.ai[j];
        yx += ax * e.ar[j];
        yy += ay * e.ar[j];
        zx += bx * e.br[j];
        zy += by * e.br[j];
        ox += bx * e.bi[j];
        oy += by * e.bi[j];
    }
    zx *= -c3;
    zy *= -c3;
    const R c0 = static_cast<R>(a.c0);
    ox *= -c0;
    oy *= -c0;
    const R inv = e.inv;
    const R sd = s_x * d.x + s_y * d.y + s_z * d.z;
    const R tx = s_x * xx + s_y * yx + s_z * zx;  // s . d_x d
    const R ty = s_x * xy + s_y * yy + s_z * zy;
    const R px = sd * (d.x * xx + d.y * yx + d.z * zx) * inv;  // (s.dh)(dh.d_x d)
    const R py = sd * (d.x * xy + d.y * yy + d.z * zy) * inv;
    // InterIntra (ii = 1): inter = -(s.d_a d - px), intra = -(n d_a d0 + px)
    // GenCurrent  (ii = 0): inter = -(n d_a d0 + s.d_a d), intra = 0
    const R ni = R(1) - ii;
    v[0] -= static_cast<double>(tx - ii * px + ni * nn * ox) * w;
    v[1] -= static_cast<double>(ty - ii * py + ni * nn * oy) * w;
    v[2] -= static_cast<double>(ii * (nn * ox + px)) * w;
    v[3] -= static_cast<double>(ii * (nn * oy + py)) * w;
    v[4] += static_cast<double>(s_x) * w;
    v[5] += static_cast<double>(s_y) * w;
    v[6] += static_cast<double>(s_z) * w;
    v[7] += static_cast<double>(e.dd > R(0) ? sd * rsqrt_r(e.dd) : -s_z) * w;
}

template <typename R>
__device__ __forceinline__ void time_sample(const KPoint<R> &p, const Phases<R> &e, const HaldaneArgs &a, R ii,
                                            double v[8])
{
    sample_state(e, p.w, p.n - p.cn, p.sx - p.cx, p.sy - p.cy, p.sz - p.cz, a, ii, v);
}

}  // namespace dev
}  // namespace sbe
