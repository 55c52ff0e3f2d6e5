// haldane_adaptive.hip - adaptive Cash-Karp 5(4) propagation of the 2-band
// (Haldane) SBE, single solves and batches of solves.
//
// The reference integrates with a fixed step (SBEs.h:596-624); this stepper
// is an addition with no counterpart there. It keeps the reference's model
// (Materials/Haldane.h:86-170), right-hand side (SBEs.h:762-861) and time
// samples (antelope.cpp:297-338: the currents at t_n = atmin + n dt, from the
// state at t_n), and replaces the step sequence:
//
//  * Step control per wave. The 64 k points of one wave share a step size,
//    chosen by the standard controller h <- h min(5, max(0.2, 0.9 err^-1/5))
//    from err = max over the wave's k points and state components of
//    |y5 - y4| / (atol + rtol max(|y|, |y5|)); a step with err > 1 is
//    rejected and retried. The k points are independent ODEs that share only
//    the laser, so this meets the tolerance for every k point of the solve,
//    as an instance-wide controller would, with no synchronisation between
//    waves. The error estimate and the controller run in FP64 whatever the
//    state precision.
//  * Laser at the stage times. A(t) is evaluated in FP64 on the device from
//    the same pulse formulas the host uses (laser_core.hpp), for all six
//    stage times of a step at once across the lanes of the wave (see
//    step_rows). Sample times use the host-built table row.
//  * Dense output. Cash-Karp's six stages plus f(t+h, y5) (which is the next
//    step's first stage, so it costs nothing extra) admit a C1 continuous
//    extension of order 4, y(t + th h) = y + h sum_i b_i(th) f_i with the
//    rational b_i(th) below (the order-4 conditions sum b_i(th) Phi_i(tree) =
//    th^r(tree)/gamma(tree) hold for every th, b(1) = the 5th-order weights,
//    b'(0) = e_1, b'(1) = e_7). Every sample time inside an accepted step is
//    evaluated from it.
//  * A launch covers the samples [out0, out1) and ends with the state exactly
//    at t_out1 (the last step is shortened to land there), so consecutive
//    launches continue seamlessly and the partial-sum buffer stays bounded.
#include "haldane_dev.hpp"

namespace sbe {
namespace {

using namespace dev;

// Cash & Karp, ACM TOMS 16 (1990) 201, Table 1 (the host Tableau holds the
// same pair).
struct CashKarp {
    static constexpr double c[6] = {0.0, 0.2, 0.3, 0.6, 1.0, 0.875};
    static constexpr double a[6][5] = {
        {0, 0, 0, 0, 0},
        {0.2, 0, 0, 0, 0},
        {3. / 40., 9. / 40., 0, 0, 0},
        {0.3, -0.9, 1.2, 0, 0},
        {-11. / 54., 2.5, -70. / 27., 35. / 27., 0},
        {1631. / 55296., 175. / 512., 575. / 13824., 44275. / 110592., 253. / 4096.}};
    static constexpr double b[6] = {37. / 378., 0, 250. / 621., 125. / 594., 0, 512. / 1771.};
    // b - bhat (5th- minus 4th-order weights)
    static constexpr double e[6] = {37. / 378. - 2825. / 27648., 0, 250. / 621. - 18575. / 48384.,
                                    125. / 594. - 13525. / 55296., -277. / 14336., 512. / 1771. - 0.25};
    // dense output b_i(th) = sum_m beta[i][m] th^(m+1), stages 1..6 then f(t+h, y5)
    static constexpr double beta[7][4] = {
        {1.0, -65. / 21., 677. / 189., -25. / 18.},
        {0, 0, 0, 0},
        {0, 2500. / 483., -38000. / 4347., 250. / 63.},
        {0, -125. / 44., 3875. / 594., -125. / 36.},
        {0, -45. / 28., 45. / 14., -45. / 28.},
        {0, 1536. / 1771., -1024. / 1771., 0},
        {0, 1.5, -4.0, 2.5}};
};

__device__ __forceinline__ double uniform(double x)
{
    const int lo = __builtin_amdgcn_readfirstlane(__double2loint(x));
    const int hi = __builtin_amdgcn_readfirstlane(__double2hiint(x));
    return __hiloint2double(hi, lo);
}

// max that propagates NaN (fmax would drop it and accept a non-finite step)
__device__ __forceinline__ double nanmax(double a, double b) { return (a != a || a > b) ? a : b; }

__device__ __forceinline__ double wave_max(double x)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x = nanmax(x, __shfl_xor(x, off, 64));
    return uniform(x);
}

// Phase rows of the six stage times of one step (stages 2..6 of Cash-Karp and
// t + h), in the layout of the sample table: t1 e^{iA.a_j} (j = 0..2) then
// e^{iA.b_j}, as (re, im) pairs. All six are evaluated at once across the
// wave: lane s*np + p evaluates pulse p at time s, then lane 6s + j the
// phase j of time s (36 lanes); the rows go to this wave's LDS slice, from
// which every lane reads them as broadcasts.
template <typename R>
__device__ __forceinline__ void step_rows(const AdaptiveArgs &A, double t, double hh, double tn, R *rows /*[6][12]*/)
{
    const int lane = __lane_id();
    const int np = A.npulses;
    auto time_of = [&](int s) {
        double tau = tn;
#pragma unroll
        for (int i = 0; i < 5; ++i)
            if (s == i) tau = t + CashKarp::c[i + 1] * hh;
        return tau;
    };
    double ax = 0.0, ay = 0.0;
    if (lane < 6 * np) {
        double pa[3], pe[3];
        pulse_field(A.pulses[lane % np], time_of(lane / np), false, pa, pe);
        ax = pa[0];
        ay = pa[1];
    }
    const int s = lane / 6, j = lane - 6 * s;  // phase j of time s (lanes < 36)
    const int base = (s < 6 ? s : 0) * np;
    double sx = 0.0, sy = 0.0;  // summed in pulse order, then the value at atmin removed (as the host)
    for (int q = 0; q < np; ++q) {
        sx += __shfl(ax, base + q, 64);
        sy += __shfl(ay, base + q, 64);
    }
    const double Ax = sx - A.A_at_min[0], Ay = sy - A.A_at_min[1];
    const HaldaneArgs &a = A.h;
    double vx = 0.0, vy = 0.0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        if (j == q) vx = a.va[q][0], vy = a.va[q][1];
        if (j == q + 3) vx = a.vb[q][0], vy = a.vb[q][1];
    }
    double sn, cs;
    sincos(Ax * vx + Ay * vy, &sn, &cs);
    const double scale = j < 3 ? a.t1 : 1.0;
    if (lane < 36) {
        rows[s * 12 + 2 * j] = static_cast<R>(cs * scale);
        rows[s * 12 + 2 * j + 1] = static_cast<R>(sn * scale);
    }
    // the writes of other lanes of this wave are read next: LDS operations of
    // one wave complete in order; keep the compiler from moving reads above
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// d(k + A) from the k phases and a stage row (same arithmetic as the fixed-step kernel)
template <typename R>
__device__ __forceinline__ Vec3<R> stage_d(const KPoint<R> &p, const R *row, R c3, R M0)
{
    R dx = 0, dy = 0, sb = 0;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        dx += p.er[j] * row[2 * j] - p.ei[j] * row[2 * j + 1];
        dy += p.er[j] * row[2 * j + 1] + p.ei[j] * row[2 * j];
        sb += p.er[j + 3] * row[7 + 2 * j] + p.ei[j + 3] * row[6 + 2 * j];
    }
    return {dx, dy, M0 - c3 * sb};
}

// f/2 = d x y - (g2/2) y + ((g2 - g1)/2) (y.d / |d|^2) d  (T1 off: g1 = 0)
template <typename R, bool DEPH>
__device__ __forceinline__ Vec3<R> half_f(const Vec3<R> &d, const Vec3<R> &y, R gh2, R gh12)
{
    Vec3<R> f;
    f.x = d.y * y.z - d.z * y.y;
    f.y = d.z * y.x - d.x * y.z;
    f.z = d.x * y.y - d.y * y.x;
    if (DEPH) {
        const R dd = d.x * d.x + d.y * d.y + d.z * d.z;
        const R inv = dd > R(0) ? R(1) / dd : R(0);
        const R c = -gh12 * (y.x * d.x + y.y * d.y + y.z * d.z) * inv;
        f.x += c * d.x - gh2 * y.x;
        f.y += c * d.y - gh2 * y.y;
        f.z += c * d.z - gh2 * y.z;
    }
    return f;
}

template <typename R>
__device__ __forceinline__ void kahan(R &sum, R &c, R inc)
{
    const R y = inc - c;
    const R t = sum + y;
    c = (t - sum) - y;
    sum = t;
}

template <typename R, bool DEPH>
__device__ void adaptive_body(const AdaptiveArgs &A, int64_t blk)
{
    using CK = CashKarp;
    constexpr bool kComp = StateT<R>::kComp != 0;
    const HaldaneArgs &a = A.h;
    const int64_t ks = a.nk_stride;
    const int64_t k = blk * kBlock + threadIdx.x;
    const int wave = static_cast<int>(k >> 6);
    const int lane = __lane_id();
    R *st = static_cast<R *>(a.state);
    const R *samp = StateT<R>::kComp ? reinterpret_cast<const R *>(a.phase_tab32)
                                     : reinterpret_cast<const R *>(a.phase_tab);

    KPoint<R> p;
    {
        const double kx = a.kpts[k], ky = a.kpts[ks + k];
        p.w = a.weight[k];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double sa, ca, sb, cb;
            sincos(kx * a.va[j][0] + ky * a.va[j][1], &sa, &ca);
            sincos(kx * a.vb[j][0] + ky * a.vb[j][1], &sb, &cb);
            p.er[j] = static_cast<R>(ca);
            p.ei[j] = static_cast<R>(sa);
            p.er[j + 3] = static_cast<R>(cb);
            p.ei[j + 3] = static_cast<R>(sb);
        }
        p.n = st[k];
        p.sx = st[ks + k];
        p.sy = st[2 * ks + k];
        p.sz = st[3 * ks + k];
        p.cn = p.cx = p.cy = p.cz = R(0);
        if (kComp) {
            p.cn = st[4 * ks + k];
            p.cx = st[5 * ks + k];
            p.cy = st[6 * ks + k];
            p.cz = st[7 * ks + k];
        }
    }
    const R c3 = static_cast<R>(a.c3), M0 = static_cast<R>(a.M0);
    // T1 is off on this path (the host rejects it): g1 = 0
    const R gh2 = static_cast<R>(0.5 * a.gam2), gh12 = static_cast<R>(-0.5 * a.gam2);
    const R ii = a.inter_intra ? R(1) : R(0);
    const int64_t nw = a.nwaves;
    auto emit = [&](int64_t m, double v[8]) {
        const double r = wave_reduce8(v);
        a.partials[((m - A.out0) * nw + wave) * kHaldaneNV + (lane >> 3)] = r;
    };

    int64_t m = A.out0;
    double t = A.atmin + A.dt_out * static_cast<double>(m);
    const double t_end = A.atmin + A.dt_out * static_cast<double>(A.out1);
    double h = uniform(A.hstep[wave]);
    int64_t nacc = 0, nrej = 0;
    bool failed = false;

    // sample at t_out0 from the state itself; its d(k+A) gives the first stage
    Vec3<R> k1;
    {
        double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        const Vec3<R> d = time_sample<R>(p, samp + m * 12, a, ii, v);
        emit(m, v);
        ++m;
        k1 = half_f<R, DEPH>(d, Vec3<R>{p.sx, p.sy, p.sz}, gh2, gh12);
    }

    __shared__ R lds_rows[kBlock / 64][6 * 12];
    R *rows = lds_rows[threadIdx.x >> 6];

    int iter = 0;
    while (t < t_end) {
        if (++iter > A.max_iter || !(h >= A.hmin)) {
            failed = true;
            break;
        }
        double hh = fmin(h, A.hmax);
        const bool last = t + 1.0001 * hh >= t_end;
        if (last) hh = t_end - t;
        const double tn = last ? t_end : t + hh;
        step_rows<R>(A, t, hh, tn, rows);

        // stages 2..6 (stage 1 is k1)
        Vec3<R> kv[6];
        kv[0] = k1;
        const Vec3<R> y0{p.sx, p.sy, p.sz};
#pragma unroll
        for (int i = 1; i < 6; ++i) {
            Vec3<R> y = y0;
#pragma unroll
            for (int j = 0; j < i; ++j) {
                if (CK::a[i][j] != 0.0) {
                    const R aij = static_cast<R>(2.0 * CK::a[i][j] * hh);
                    y.x += aij * kv[j].x;
                    y.y += aij * kv[j].y;
                    y.z += aij * kv[j].z;
                }
            }
            kv[i] = half_f<R, DEPH>(stage_d(p, rows + (i - 1) * 12, c3, M0), y, gh2, gh12);
        }
        // 5th-order increment (state precision) and the error estimate (FP64)
        Vec3<R> inc{0, 0, 0};
        double ex = 0, ey = 0, ez = 0;
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            if (CK::b[i] != 0.0) {
                const R bi = static_cast<R>(2.0 * CK::b[i] * hh);
                inc.x += bi * kv[i].x;
                inc.y += bi * kv[i].y;
                inc.z += bi * kv[i].z;
            }
            if (CK::e[i] != 0.0) {
                const double ei = 2.0 * CK::e[i] * hh;
                ex += ei * static_cast<double>(kv[i].x);
                ey += ei * static_cast<double>(kv[i].y);
                ez += ei * static_cast<double>(kv[i].z);
            }
        }
        const Vec3<R> y5{y0.x + inc.x, y0.y + inc.y, y0.z + inc.z};
        auto ratio = [&](double e, R y, R z) {  // 0 when the component and its error vanish (atol = 0)
            const double sc = A.atol + A.rtol * fmax(fabs(static_cast<double>(y)), fabs(static_cast<double>(z)));
            return e == 0.0 ? 0.0 : fabs(e) / sc;
        };
        const double err_k = nanmax(ratio(ex, y0.x, y5.x), nanmax(ratio(ey, y0.y, y5.y), ratio(ez, y0.z, y5.z)));
        const double err = wave_max(err_k);

        if (err <= 1.0) {
            const Vec3<R> k7 = half_f<R, DEPH>(stage_d(p, rows + 5 * 12, c3, M0), y5, gh2, gh12);
            // samples inside (t, tn]
            const R nn = p.n - p.cn;
            const Vec3<R> yt{p.sx - p.cx, p.sy - p.cy, p.sz - p.cz};
            while (m < A.out1) {
                const double tm = A.atmin + A.dt_out * static_cast<double>(m);
                if (tm > tn) break;
                const double th = (tm - t) / hh;
                Vec3<R> yo = yt;
#pragma unroll
                for (int i = 0; i < 7; ++i) {
                    if (i == 1) continue;  // b_2(th) = 0
                    const double bt = th * (CK::beta[i][0] + th * (CK::beta[i][1] + th * (CK::beta[i][2] + th * CK::beta[i][3])));
                    const Vec3<R> &kk = i < 6 ? kv[i] : k7;
                    const R w = static_cast<R>(2.0 * hh * bt);
                    yo.x += w * kk.x;
                    yo.y += w * kk.y;
                    yo.z += w * kk.z;
                }
                double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
                (void)sample_state<R>(p, nn, yo.x, yo.y, yo.z, samp + m * 12, a, ii, v);
                emit(m, v);
                ++m;
            }
            if (kComp) {
                kahan(p.sx, p.cx, inc.x);
                kahan(p.sy, p.cy, inc.y);
                kahan(p.sz, p.cz, inc.z);
            } else {
                p.sx += inc.x;
                p.sy += inc.y;
                p.sz += inc.z;
            }
            k1 = k7;
            t = tn;
            ++nacc;
            const double fac = err > 0.0 ? fmin(5.0, fmax(0.2, 0.9 * pow(err, -0.2))) : 5.0;
            if (!last) h = hh * fac;  // a shortened last step keeps the proposal it was cut from
        } else {
            h = hh * fmax(0.2, 0.9 * pow(err, -0.2));  // NaN err -> 0.2: shrinks to hmin and fails
            ++nrej;
        }
    }

    st[k] = p.n;
    st[ks + k] = p.sx;
    st[2 * ks + k] = p.sy;
    st[3 * ks + k] = p.sz;
    if (kComp) {
        st[4 * ks + k] = p.cn;
        st[5 * ks + k] = p.cx;
        st[6 * ks + k] = p.cy;
        st[7 * ks + k] = p.cz;
    }
    if (lane == 0) {
        A.hstep[wave] = h;
        A.counts[2 * wave] += nacc;
        A.counts[2 * wave + 1] += nrej;
        if (failed) *A.fail = 1;
    }
}

// Many solves in one launch: block b works for solve block_inst[b] on that
// solve's block b - block_first[solve]; solves with no samples left leave at
// once.
template <typename R, bool DEPH>
__global__ __launch_bounds__(kBlock) void haldane_adaptive_kernel(const AdaptiveArgs *__restrict__ insts,
                                                                  const int32_t *__restrict__ block_inst,
                                                                  const int32_t *__restrict__ block_first)
{
    const int inst = block_inst[blockIdx.x];
    const AdaptiveArgs &A = insts[inst];
    if (A.out1 <= A.out0) return;  // whole block
    adaptive_body<R, DEPH>(A, static_cast<int64_t>(blockIdx.x) - block_first[inst]);
}

// One solve, arguments by value (no upload per launch).
template <typename R, bool DEPH>
__global__ __launch_bounds__(kBlock) void haldane_adaptive_single_kernel(const AdaptiveArgs A)
{
    adaptive_body<R, DEPH>(A, blockIdx.x);
}

template <typename R>
void launch_adaptive(const AdaptiveArgs *d, const int32_t *bi, const int32_t *bf, int nblocks, bool deph, hipStream_t s)
{
    if (deph) haldane_adaptive_kernel<R, true><<<nblocks, kBlock, 0, s>>>(d, bi, bf);
    else haldane_adaptive_kernel<R, false><<<nblocks, kBlock, 0, s>>>(d, bi, bf);
}

}  // namespace

hipError_t launch_haldane_adaptive(const AdaptiveArgs &a, hipStream_t s)
{
    const dim3 grid(static_cast<unsigned>(a.h.nk_stride / kBlock));
    const bool deph = a.h.gam2 != 0.0;
    if (a.h.precision == SBE_PREC_FP32) {
        if (deph) haldane_adaptive_single_kernel<float, true><<<grid, kBlock, 0, s>>>(a);
        else haldane_adaptive_single_kernel<float, false><<<grid, kBlock, 0, s>>>(a);
    } else {
        if (deph) haldane_adaptive_single_kernel<double, true><<<grid, kBlock, 0, s>>>(a);
        else haldane_adaptive_single_kernel<double, false><<<grid, kBlock, 0, s>>>(a);
    }
    return hipGetLastError();
}

hipError_t launch_haldane_adaptive_batch(const AdaptiveArgs *d_insts, const int32_t *d_block_inst,
                                         const int32_t *d_block_first, int nblocks, int precision, bool deph,
                                         hipStream_t s)
{
    if (nblocks <= 0) return hipSuccess;
    if (precision == SBE_PREC_FP32) launch_adaptive<float>(d_insts, d_block_inst, d_block_first, nblocks, deph, s);
    else launch_adaptive<double>(d_insts, d_block_inst, d_block_first, nblocks, deph, s);
    return hipGetLastError();
}

}  // namespace sbe
