// haldane.hip - 2-band (Haldane) SBE propagation kernels for gfx950.
//
// Reference: Materials/Haldane.h:86-170 (H, J), SBEs.h:596-624 (RK step),
// SBEs.h:762-861 (RHS with T2 dephasing), SBEs.h:694-760 (currents) and the
// sampling/accumulation of antelope.cpp:297-338.
//
// Design. A 2x2 Hermitian density matrix is held as its Bloch form
// rho = (n I + s.sigma)/2, and H(k+A) = d0 I + d.sigma. Then
//   -i[H, rho]                   ->  ds/dt = 2 d x s
//   -U (Gamma o (U^dag rho U)) U^dag ->  ds/dt -= g2 (s - s_par) + g1 s_par,
//                                        dn/dt = -g1 n,  s_par = (s.d) d/|d|^2
//   Re Tr(rho J_a) = -(n d_a d0 + s.d_a d),  intra = -(n d_a d0 + (s.dh)(dh.d_a d))
// which is the reference arithmetic restated without eigenvectors (the
// oracle's generic N x N path with a Jacobi eigensolver checks it).
//
// One thread owns one k point for a whole chunk of time steps: the state
// never leaves registers inside a launch, so HBM sees the state once per
// chunk. H(k+A) uses e^{i(k+A).v} = e^{ik.v} e^{iA.v}: the six e^{ik.v} are
// computed once per launch, the k-independent e^{iA(t).v} come from a
// per-stage table read with uniform (scalar) loads. Every time sample is
// reduced over the 64 lanes of a wave with gfx950 permlane swaps and DPP,
// and one row of 8 FP64 partial sums per wave and step goes to HBM; a second
// kernel (series.hip) sums the rows in a fixed order.
#include "sbe_internal.hpp"

namespace sbe {
namespace {

// ----------------------------------------------------------- RK tableaus
template <int RK> struct Coef;
template <> struct Coef<4> {
    static constexpr int S = 4;
    static constexpr double a[4][4] = {{0, 0, 0, 0}, {0.5, 0, 0, 0}, {0, 0.5, 0, 0}, {0, 0, 1.0, 0}};
    static constexpr double b[4] = {1.0 / 6.0, 2.0 / 6.0, 2.0 / 6.0, 1.0 / 6.0};
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
};

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

// Bloch RHS: ds = 2 d x s - g2 (s - s_par) - g1 s_par;  dn = -g1 n
template <typename R, bool T1ON>
__device__ __forceinline__ void rhs(const Vec3<R> &d, R n, const Vec3<R> &s, R g1, R g2, R &dn,
                                    Vec3<R> &ds)
{
    ds.x = R(2) * (d.y * s.z - d.z * s.y);
    ds.y = R(2) * (d.z * s.x - d.x * s.z);
    ds.z = R(2) * (d.x * s.y - d.y * s.x);
    const R dd = d.x * d.x + d.y * d.y + d.z * d.z;
    const R inv = dd > R(0) ? R(1) / dd : R(0);
    const R sp = (s.x * d.x + s.y * d.y + s.z * d.z) * inv;  // s_par = sp * d
    // g2 (s - sp d) + g1 sp d = g2 s + (g1 - g2) sp d
    const R cpar = (g1 - g2) * sp;
    ds.x -= g2 * s.x + cpar * d.x;
    ds.y -= g2 * s.y + cpar * d.y;
    ds.z -= g2 * s.z + cpar * d.z;
    dn = T1ON ? -g1 * n : R(0);
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

// complex product (a + ib)(c + id)
template <typename R>
__device__ __forceinline__ void cmul(R a, R b, R c, R d, R &re, R &im)
{
    re = a * c - b * d;
    im = a * d + b * c;
}

// ------------------------------------------------------------ main kernel
// R: arithmetic of the propagated state (double, or float with a Kahan
// compensation term per component). The time samples are always FP64.
template <typename R, int RK, bool T1ON>
__global__ __launch_bounds__(kBlock) void haldane_steps_kernel(HaldaneArgs a, const R *phase_tab_r)
{
    using C = Coef<RK>;
    constexpr int S = C::S;
    constexpr bool kComp = StateT<R>::kComp != 0;
    const int64_t k = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;  // < nk_stride
    const int64_t ks = a.nk_stride;
    const int wave = static_cast<int>(k >> 6);

    // per-k constants
    const double kx = a.kpts[k], ky = a.kpts[ks + k];
    const double wk = a.weight[k];
    double ekr[6], eki[6];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        sincos(kx * a.va[j][0] + ky * a.va[j][1], &eki[j], &ekr[j]);
        sincos(kx * a.vb[j][0] + ky * a.vb[j][1], &eki[j + 3], &ekr[j + 3]);
    }
    R ekr_r[6], eki_r[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        ekr_r[j] = static_cast<R>(ekr[j]);
        eki_r[j] = static_cast<R>(eki[j]);
    }
    // derivative weights of the measurement (x, y components of a_j, b_j)
    const R t1 = static_cast<R>(a.t1), c3 = static_cast<R>(a.c3), M0 = static_cast<R>(a.M0);
    const R g1 = static_cast<R>(a.gam1), g2 = static_cast<R>(a.gam2);
    const R dt = static_cast<R>(a.dt);

    // state (SoA): n, sx, sy, sz [, compensation terms]
    R *st = static_cast<R *>(a.state);
    R n = st[k], sx = st[ks + k], sy = st[2 * ks + k], sz = st[3 * ks + k];
    R cn = 0, cx = 0, cy = 0, cz = 0;
    if (kComp) {
        cn = st[4 * ks + k];
        cx = st[5 * ks + k];
        cy = st[6 * ks + k];
        cz = st[7 * ks + k];
    }

    for (int step = 0; step < a.nsteps; ++step) {
        const int64_t gstep = a.step0 + step;
        const R *ptab = phase_tab_r + gstep * (S * 12);
        const double *ptab64 = a.phase_tab + gstep * (S * 12);

        // ---- time sample at t_n from the state before the step (FP64)
        {
            double er[6], ei[6];
#pragma unroll
            for (int j = 0; j < 6; ++j) cmul(ekr[j], eki[j], ptab64[2 * j], ptab64[2 * j + 1], er[j], ei[j]);
            const Vec3<double> d = dvec<double>(er, ei, a.t1, a.c3, a.M0);
            // d/dk_a of (d0, d) for a = x, y
            double d0x = 0, d0y = 0, dxx = 0, dxy = 0, dyx = 0, dyy = 0, dzx = 0, dzy = 0;
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                dxx -= a.va[j][0] * ei[j];
                dxy -= a.va[j][1] * ei[j];
                dyx += a.va[j][0] * er[j];
                dyy += a.va[j][1] * er[j];
                dzx += a.vb[j][0] * er[j + 3];
                dzy += a.vb[j][1] * er[j + 3];
                d0x -= a.vb[j][0] * ei[j + 3];
                d0y -= a.vb[j][1] * ei[j + 3];
            }
            dxx *= a.t1; dxy *= a.t1; dyx *= a.t1; dyy *= a.t1;
            dzx *= -a.c3; dzy *= -a.c3; d0x *= a.c0; d0y *= a.c0;
            const double nn = static_cast<double>(n) - static_cast<double>(cn);
            const double s_x = static_cast<double>(sx) - static_cast<double>(cx);
            const double s_y = static_cast<double>(sy) - static_cast<double>(cy);
            const double s_z = static_cast<double>(sz) - static_cast<double>(cz);
            const double dd = d.x * d.x + d.y * d.y + d.z * d.z;
            const double inv = dd > 0.0 ? 1.0 / dd : 0.0;
            const double sd = s_x * d.x + s_y * d.y + s_z * d.z;
            const double tx = s_x * dxx + s_y * dyx + s_z * dzx;  // s . d_x d
            const double ty = s_x * dxy + s_y * dyy + s_z * dzy;
            const double px = sd * (d.x * dxx + d.y * dyx + d.z * dzx) * inv;  // (s.dh)(dh.d_x d)
            const double py = sd * (d.x * dxy + d.y * dyy + d.z * dzy) * inv;
            double v[8];
            if (a.inter_intra) {
                v[0] = -(tx - px) * wk;
                v[1] = -(ty - py) * wk;
                v[2] = -(nn * d0x + px) * wk;
                v[3] = -(nn * d0y + py) * wk;
            } else {  // GenCurrent: the whole Re Tr(rho J) goes to "inter"
                v[0] = -(nn * d0x + tx) * wk;
                v[1] = -(nn * d0y + ty) * wk;
                v[2] = 0.0;
                v[3] = 0.0;
            }
            v[4] = s_x * wk;
            v[5] = s_y * wk;
            v[6] = s_z * wk;
            v[7] = (dd > 0.0 ? sd * rsqrt(dd) : 0.0) * wk;
            const double r = wave_reduce8(v);
            if ((__lane_id() & 7) == 0)
                a.partials[(static_cast<int64_t>(step) * a.nwaves + wave) * kHaldaneNV + (__lane_id() >> 3)] = r;
        }

        // ---- one RK step (SBEs.h:596-624)
        R kn[S];
        Vec3<R> ks_[S];
#pragma unroll
        for (int i = 0; i < S; ++i) {
            R yn = n;
            Vec3<R> y{sx, sy, sz};
#pragma unroll
            for (int j = 0; j < i; ++j) {
                if (C::a[i][j] != 0.0) {
                    const R aij = static_cast<R>(C::a[i][j]) * dt;
                    yn += aij * kn[j];
                    y.x += aij * ks_[j].x;
                    y.y += aij * ks_[j].y;
                    y.z += aij * ks_[j].z;
                }
            }
            R er[6], ei[6];
#pragma unroll
            for (int j = 0; j < 6; ++j) cmul(ekr_r[j], eki_r[j], ptab[i * 12 + 2 * j], ptab[i * 12 + 2 * j + 1], er[j], ei[j]);
            const Vec3<R> d = dvec<R>(er, ei, t1, c3, M0);
            rhs<R, T1ON>(d, yn, y, g1, g2, kn[i], ks_[i]);
        }
        R dn = 0;
        Vec3<R> ds{0, 0, 0};
#pragma unroll
        for (int i = 0; i < S; ++i) {
            if (C::b[i] != 0.0) {
                const R bi = static_cast<R>(C::b[i]);
                dn += bi * kn[i];
                ds.x += bi * ks_[i].x;
                ds.y += bi * ks_[i].y;
                ds.z += bi * ks_[i].z;
            }
        }
        if (kComp) {  // Kahan-compensated state update
            auto kahan = [](R &sum, R &c, R inc) {
                const R y = inc - c;
                const R t = sum + y;
                c = (t - sum) - y;
                sum = t;
            };
            if (T1ON) kahan(n, cn, dt * dn);
            kahan(sx, cx, dt * ds.x);
            kahan(sy, cy, dt * ds.y);
            kahan(sz, cz, dt * ds.z);
        } else {
            if (T1ON) n += dt * dn;
            sx += dt * ds.x;
            sy += dt * ds.y;
            sz += dt * ds.z;
        }
    }
    st[k] = n;
    st[ks + k] = sx;
    st[2 * ks + k] = sy;
    st[3 * ks + k] = sz;
    if (kComp) {
        st[4 * ks + k] = cn;
        st[5 * ks + k] = cx;
        st[6 * ks + k] = cy;
        st[7 * ks + k] = cz;
    }
}

// Initial state (SBEs.h:284-364): rho_H = diag(occupations) at k, rho = U rho_H U^dag
// -> n = f0 + f1, s = (f1 - f0) dh(k). Padding points get a zero state.
template <typename R>
__global__ __launch_bounds__(kBlock) void haldane_init_kernel(HaldaneArgs a, int nval, int fermi, double ef)
{
    const int64_t k = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    const int64_t ks = a.nk_stride;
    R *st = static_cast<R *>(a.state);
    double n = 0, sx = 0, sy = 0, sz = 0;
    if (k < a.nk) {
        const double kx = a.kpts[k], ky = a.kpts[ks + k];
        double er[6], ei[6];
        for (int j = 0; j < 3; ++j) {
            sincos(kx * a.va[j][0] + ky * a.va[j][1], &ei[j], &er[j]);
            sincos(kx * a.vb[j][0] + ky * a.vb[j][1], &ei[j + 3], &er[j + 3]);
        }
        const Vec3<double> d = dvec<double>(er, ei, a.t1, a.c3, a.M0);
        const double d0 = a.c0 * (er[3] + er[4] + er[5]);
        const double dn = sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
        double hx = 0, hy = 0, hz = -1;  // degenerate point: eigenbasis = orbital basis
        if (dn > 0) { hx = d.x / dn; hy = d.y / dn; hz = d.z / dn; }
        double f0, f1;
        if (fermi) {
            const double kT = 300.0 / 3.15775024804e5;
            f0 = 1.0 / (exp((d0 - dn - ef) / kT) + 1.0);
            f1 = 1.0 / (exp((d0 + dn - ef) / kT) + 1.0);
            const double sum = f0 + f1;
            f0 /= sum;
            f1 /= sum;
        } else {
            f0 = nval > 0 ? 1.0 : 0.0;
            f1 = nval > 1 ? 1.0 : 0.0;
        }
        n = f0 + f1;
        sx = (f1 - f0) * hx;
        sy = (f1 - f0) * hy;
        sz = (f1 - f0) * hz;
    }
    st[k] = static_cast<R>(n);
    st[ks + k] = static_cast<R>(sx);
    st[2 * ks + k] = static_cast<R>(sy);
    st[3 * ks + k] = static_cast<R>(sz);
    if (StateT<R>::kComp) {
        // compensation = -(low part) so that value = hi - c is exact in FP64
        st[4 * ks + k] = static_cast<R>(static_cast<double>(static_cast<R>(n)) - n);
        st[5 * ks + k] = static_cast<R>(static_cast<double>(static_cast<R>(sx)) - sx);
        st[6 * ks + k] = static_cast<R>(static_cast<double>(static_cast<R>(sy)) - sy);
        st[7 * ks + k] = static_cast<R>(static_cast<double>(static_cast<R>(sz)) - sz);
    }
}

// Orbital-basis AoS export: rho[k][2][2][re,im] (the reference dmatrix).
template <typename R>
__global__ __launch_bounds__(kBlock) void haldane_export_kernel(HaldaneArgs a, double *out)
{
    const int64_t k = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (k >= a.nk) return;
    const int64_t ks = a.nk_stride;
    const R *st = static_cast<const R *>(a.state);
    double v[4];
    for (int c = 0; c < 4; ++c) {
        v[c] = static_cast<double>(st[c * ks + k]);
        if (StateT<R>::kComp) v[c] -= static_cast<double>(st[(4 + c) * ks + k]);
    }
    const double n = v[0], sx = v[1], sy = v[2], sz = v[3];
    double *o = out + k * 8;
    o[0] = 0.5 * (n + sz); o[1] = 0.0;
    o[2] = 0.5 * sx;       o[3] = -0.5 * sy;
    o[4] = 0.5 * sx;       o[5] = 0.5 * sy;
    o[6] = 0.5 * (n - sz); o[7] = 0.0;
}

}  // namespace

int haldane_waves(int64_t nk) { return static_cast<int>((nk + 63) / 64); }

hipError_t launch_haldane_init(const HaldaneArgs &a, int init_type, double ef, hipStream_t s)
{
    const dim3 grid(static_cast<unsigned>(a.nk_stride / kBlock));
    const int fermi = init_type == SBE_INIT_FERMI_DIRAC;
    if (a.precision == SBE_PREC_FP32) haldane_init_kernel<float><<<grid, kBlock, 0, s>>>(a, a.nval, fermi, ef);
    else haldane_init_kernel<double><<<grid, kBlock, 0, s>>>(a, a.nval, fermi, ef);
    return hipGetLastError();
}

template <typename R, int RK>
static void launch_steps(const HaldaneArgs &a, const R *tab, bool t1_on, hipStream_t s)
{
    const dim3 grid(static_cast<unsigned>(a.nk_stride / kBlock));
    if (t1_on) haldane_steps_kernel<R, RK, true><<<grid, kBlock, 0, s>>>(a, tab);
    else haldane_steps_kernel<R, RK, false><<<grid, kBlock, 0, s>>>(a, tab);
}

hipError_t launch_haldane_steps(const HaldaneArgs &a, int rk_order, bool t1_on, hipStream_t s)
{
    const bool f32 = a.precision == SBE_PREC_FP32;
    if (rk_order == SBE_RK4) {
        if (f32) launch_steps<float, 4>(a, a.phase_tab32, t1_on, s);
        else launch_steps<double, 4>(a, a.phase_tab, t1_on, s);
    } else if (rk_order == SBE_RK5_FEHLBERG) {
        if (f32) launch_steps<float, 5>(a, a.phase_tab32, t1_on, s);
        else launch_steps<double, 5>(a, a.phase_tab, t1_on, s);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_haldane_export(const HaldaneArgs &a, double *out, hipStream_t s)
{
    const dim3 grid(static_cast<unsigned>(a.nk_stride / kBlock));
    if (a.precision == SBE_PREC_FP32) haldane_export_kernel<float><<<grid, kBlock, 0, s>>>(a, out);
    else haldane_export_kernel<double><<<grid, kBlock, 0, s>>>(a, out);
    return hipGetLastError();
}

}  // namespace sbe
