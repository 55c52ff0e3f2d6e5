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
#include "haldane_dev.hpp"

namespace sbe {
namespace {

using namespace dev;

// ------------------------------------------------------------ main kernel
constexpr int kTile = 32;  // time steps per LDS tile of the phase table

// Step kernel. R: arithmetic of the propagated state and of the per-k
// samples (double, or float with a Kahan compensation term per state
// component); the k sums are FP64 in both. DEPH: any of T1/T2 on. KPT: k
// points per lane (independent chains give the scheduler instruction-level
// parallelism when waves are few).
//
// The stage right-hand side is evaluated as half of the reference RHS,
//   f/2 = d x s - (g2/2)(s - s_par) - (g1/2) s_par,  dn/2 = -(g1/2) n,
// with the factor 2 folded into the RK coefficients (dt a_ij, dt b_i).
// The NN part of the phase table carries the factor t1.
//
// d(k+A) is formed once per distinct stage time: stages at an equal c share
// it (RK4: c = 1/2 twice), and the stage at c = 1 is the next step's sample
// time. The table row of that stage equals the next step's first row bit for
// bit (the host evaluates both at t_{n+1}), so with kCarry its phases (and
// d, 1/|d|^2) serve the next sample and stage 0, and a run split into any
// launches gives the same bits as one launch.
// blk: block index within the k block of this solve.
template <typename R, int RK, bool T1ON, bool DEPH, int KPT>
__device__ __forceinline__ void steps_body(const HaldaneArgs &a, const R *__restrict__ tabr, R *__restrict__ st,
                                           double *__restrict__ partials, int64_t blk)
{
    using C = Coef<RK>;
    constexpr int S = C::S;
    constexpr bool kComp = StateT<R>::kComp != 0;
    constexpr int kOne = stage_at_one<RK>();
    constexpr bool kCarry = kOne > 0;
    const int64_t ks = a.nk_stride;
    const int64_t lane0 = blk * (kBlock * KPT) + threadIdx.x;
    const int wave = static_cast<int>((blk * kBlock + threadIdx.x) >> 6);

    KPoint<R> p[KPT];
#pragma unroll
    for (int q = 0; q < KPT; ++q) {
        const int64_t k = lane0 + q * kBlock;
        const double kx = a.kpts[k], ky = a.kpts[ks + k];
        p[q].w = a.weight[k];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double sa, ca, sb, cb;
            sincos(kx * a.va[j][0] + ky * a.va[j][1], &sa, &ca);
            sincos(kx * a.vb[j][0] + ky * a.vb[j][1], &sb, &cb);
            p[q].er[j] = static_cast<R>(ca);
            p[q].ei[j] = static_cast<R>(sa);
            p[q].er[j + 3] = static_cast<R>(cb);
            p[q].ei[j + 3] = static_cast<R>(sb);
        }
        p[q].n = st[k];
        p[q].sx = st[ks + k];
        p[q].sy = st[2 * ks + k];
        p[q].sz = st[3 * ks + k];
        p[q].cn = p[q].cx = p[q].cy = p[q].cz = R(0);
        if (kComp) {
            p[q].cn = st[4 * ks + k];
            p[q].cx = st[5 * ks + k];
            p[q].cy = st[6 * ks + k];
            p[q].cz = st[7 * ks + k];
        }
    }
    const R c3 = static_cast<R>(a.c3), M0 = static_cast<R>(a.M0);
    const R gh1 = static_cast<R>(0.5 * a.gam1), gh2 = static_cast<R>(0.5 * a.gam2);
    const R gh12 = static_cast<R>(0.5 * (a.gam1 - a.gam2));
    const R ii = a.inter_intra ? R(1) : R(0);
    const int lane = __lane_id();

    // The k-independent phase rows of kTile steps are staged in LDS (double
    // buffered): every lane reads the same row, so LDS broadcasts replace a
    // per-step scalar-cache miss that a lone wave per SIMD cannot hide.
    constexpr int ROW = S * 12;
    constexpr int TILE_VALS = kTile * ROW;
    constexpr int PER_THREAD = (TILE_VALS + kBlock - 1) / kBlock;
    __shared__ R lds_tab[2][TILE_VALS];
    const R *gtab = tabr + a.step0 * ROW;
    const int64_t tab_vals = static_cast<int64_t>(a.nsteps) * ROW;
    const int ntiles = (a.nsteps + kTile - 1) / kTile;
    R pre[PER_THREAD];
    auto fetch = [&](int tile) {
#pragma unroll
        for (int r = 0; r < PER_THREAD; ++r) {
            const int i = r * kBlock + threadIdx.x;
            const int64_t g = static_cast<int64_t>(tile) * TILE_VALS + i;
            pre[r] = (i < TILE_VALS && g < tab_vals) ? gtab[g] : R(0);
        }
    };
    auto stash = [&](int buf) {
#pragma unroll
        for (int r = 0; r < PER_THREAD; ++r) {
            const int i = r * kBlock + threadIdx.x;
            if (i < TILE_VALS) lds_tab[buf][i] = pre[r];
        }
    };
    fetch(0);
    stash(0);
    __syncthreads();

    Phases<R> ph[KPT];  // at this step's sample time t_n
    if (kCarry) {
#pragma unroll
        for (int q = 0; q < KPT; ++q) ph[q] = phases_at(p[q], &lds_tab[0][0], c3, M0);
    }

    for (int tile = 0; tile < ntiles; ++tile) {
    if (tile + 1 < ntiles) fetch(tile + 1);  // in flight during this tile
    const int s_end = min(kTile, a.nsteps - tile * kTile);
    for (int st_ = 0; st_ < s_end; ++st_) {
        const int step = tile * kTile + st_;
        const R *tr = &lds_tab[tile & 1][st_ * ROW];
        if (!kCarry) {
#pragma unroll
            for (int q = 0; q < KPT; ++q) ph[q] = phases_at(p[q], tr, c3, M0);
        }

        // ---- time sample at t_n from the state before the step
        double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int q = 0; q < KPT; ++q) time_sample<R>(p[q], ph[q], a, ii, v);
        {
            const double r = wave_reduce8(v);
            // lanes 8q..8q+7 all hold value q: duplicate stores of equal data keep the step branch-free
            partials[(static_cast<int64_t>(step) * a.nwaves + wave) * kHaldaneNV + (lane >> 3)] = r;
        }

        // ---- one RK step (SBEs.h:596-624), stage RHS halved (see above)
#pragma unroll
        for (int q = 0; q < KPT; ++q) {
            R kn[S];
            Vec3<R> kv[S];
            Vec3<R> ds_[S];  // d at each stage time
            R inv_[S];       // 1/|d|^2 at each stage time
#pragma unroll
            for (int i = 0; i < S; ++i) {
                R yn = p[q].n;
                Vec3<R> y{p[q].sx, p[q].sy, p[q].sz};
#pragma unroll
                for (int j = 0; j < i; ++j) {
                    if (C::a[i][j] != 0.0) {
                        const R aij = static_cast<R>(2.0 * C::a[i][j] * a.dt);
                        if (T1ON) yn += aij * kn[j];
                        y.x += aij * kv[j].x;
                        y.y += aij * kv[j].y;
                        y.z += aij * kv[j].z;
                    }
                }
                Vec3<R> d;
                R inv;
                const int same = same_time_stage<RK>(i);
                if (i == 0) {
                    d = ph[q].d;
                    inv = ph[q].inv;
                } else if (same > 0) {
                    d = ds_[same];
                    inv = inv_[same];
                } else if (kCarry && i == kOne) {
                    ph[q] = phases_at(p[q], tr + i * 12, c3, M0);  // the next step's t_n
                    d = ph[q].d;
                    inv = ph[q].inv;
                } else {
                    const R *row = tr + i * 12;
                    R dx = 0, dy = 0, sb = 0;
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        dx += p[q].er[j] * row[2 * j] - p[q].ei[j] * row[2 * j + 1];
                        dy += p[q].er[j] * row[2 * j + 1] + p[q].ei[j] * row[2 * j];
                        sb += p[q].er[j + 3] * row[7 + 2 * j] + p[q].ei[j + 3] * row[6 + 2 * j];
                    }
                    d = {dx, dy, M0 - c3 * sb};
                    const R dd = d.x * d.x + d.y * d.y + d.z * d.z;
                    inv = dd > R(0) ? rcp_nr(dd) : R(0);
                }
                ds_[i] = d;
                inv_[i] = inv;
                // f/2 = d x y - (g2/2) y + ((g2-g1)/2) (y.d / |d|^2) d
                Vec3<R> f;
                f.x = d.y * y.z - d.z * y.y;
                f.y = d.z * y.x - d.x * y.z;
                f.z = d.x * y.y - d.y * y.x;
                if (DEPH) {
                    const R c = -gh12 * (y.x * d.x + y.y * d.y + y.z * d.z) * inv;
                    f.x += c * d.x - gh2 * y.x;
                    f.y += c * d.y - gh2 * y.y;
                    f.z += c * d.z - gh2 * y.z;
                }
                kv[i] = f;
                kn[i] = T1ON ? -gh1 * yn : R(0);
            }
            R dn = 0;
            Vec3<R> ds{0, 0, 0};
#pragma unroll
            for (int i = 0; i < S; ++i) {
                if (C::b[i] != 0.0) {
                    const R bi = static_cast<R>(2.0 * C::b[i] * a.dt);
                    if (T1ON) dn += bi * kn[i];
                    ds.x += bi * kv[i].x;
                    ds.y += bi * kv[i].y;
                    ds.z += bi * kv[i].z;
                }
            }
            if (kComp) {  // Kahan-compensated state update
                auto kahan = [](R &sum, R &c, R inc) {
                    const R y = inc - c;
                    const R t = sum + y;
                    c = (t - sum) - y;
                    sum = t;
                };
                if (T1ON) kahan(p[q].n, p[q].cn, dn);
                kahan(p[q].sx, p[q].cx, ds.x);
                kahan(p[q].sy, p[q].cy, ds.y);
                kahan(p[q].sz, p[q].cz, ds.z);
            } else {
                if (T1ON) p[q].n += dn;
                p[q].sx += ds.x;
                p[q].sy += ds.y;
                p[q].sz += ds.z;
            }
        }
    }
    // the other buffer was last read in the previous tile (all lanes passed
    // the barrier since), so it can take the next tile now
    if (tile + 1 < ntiles) stash((tile + 1) & 1);
    __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < KPT; ++q) {
        const int64_t k = lane0 + q * kBlock;
        st[k] = p[q].n;
        st[ks + k] = p[q].sx;
        st[2 * ks + k] = p[q].sy;
        st[3 * ks + k] = p[q].sz;
        if (kComp) {
            st[4 * ks + k] = p[q].cn;
            st[5 * ks + k] = p[q].cx;
            st[6 * ks + k] = p[q].cy;
            st[7 * ks + k] = p[q].cz;
        }
    }
}

template <typename R, int RK, bool T1ON, bool DEPH, int KPT>
__global__ __launch_bounds__(kBlock) void haldane_steps_kernel(const HaldaneArgs a, const R *__restrict__ tabr,
                                                               R *__restrict__ st, double *__restrict__ partials)
{
    steps_body<R, RK, T1ON, DEPH, KPT>(a, tabr, st, partials, blockIdx.x);
}

// Many independent solves in one launch: block b works for solve
// block_inst[b] on that solve's block b - block_first[solve]. Every solve
// carries its own model, grid, laser table, state and step count; solves
// with nothing left to do this launch leave at once.
template <typename R, int RK, bool T1ON, bool DEPH>
__global__ __launch_bounds__(kBlock) void haldane_batch_kernel(const HaldaneArgs *__restrict__ insts,
                                                               const int32_t *__restrict__ block_inst,
                                                               const int32_t *__restrict__ block_first)
{
    const int inst = block_inst[blockIdx.x];
    const HaldaneArgs a = insts[inst];
    if (a.nsteps <= 0) return;  // whole block
    const R *tab = StateT<R>::kComp ? reinterpret_cast<const R *>(a.phase_tab32) : reinterpret_cast<const R *>(a.phase_tab);
    steps_body<R, RK, T1ON, DEPH, 1>(a, tab, static_cast<R *>(a.state), a.partials,
                                     static_cast<int64_t>(blockIdx.x) - block_first[inst]);
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

template <typename R, int RK, int KPT>
static void launch_steps(const HaldaneArgs &a, const R *tab, bool t1_on, hipStream_t s)
{
    const dim3 grid(static_cast<unsigned>(a.nk_stride / (kBlock * KPT)));
    R *st = static_cast<R *>(a.state);
    const bool deph = a.gam1 != 0.0 || a.gam2 != 0.0;
    if (t1_on) haldane_steps_kernel<R, RK, true, true, KPT><<<grid, kBlock, 0, s>>>(a, tab, st, a.partials);
    else if (deph) haldane_steps_kernel<R, RK, false, true, KPT><<<grid, kBlock, 0, s>>>(a, tab, st, a.partials);
    else haldane_steps_kernel<R, RK, false, false, KPT><<<grid, kBlock, 0, s>>>(a, tab, st, a.partials);
}

template <typename R, int RK>
static void launch_steps_kpt(const HaldaneArgs &a, const R *tab, bool t1_on, hipStream_t s)
{
    if (a.kpt == 2) launch_steps<R, RK, 2>(a, tab, t1_on, s);
    else launch_steps<R, RK, 1>(a, tab, t1_on, s);
}

hipError_t launch_haldane_steps(const HaldaneArgs &a, int rk_order, bool t1_on, hipStream_t s)
{
    const bool f32 = a.precision == SBE_PREC_FP32;
    if (rk_order == SBE_RK4) {
        if (f32) launch_steps_kpt<float, 4>(a, a.phase_tab32, t1_on, s);
        else launch_steps_kpt<double, 4>(a, a.phase_tab, t1_on, s);
    } else if (rk_order == SBE_RK5_FEHLBERG) {
        if (f32) launch_steps_kpt<float, 5>(a, a.phase_tab32, t1_on, s);
        else launch_steps_kpt<double, 5>(a, a.phase_tab, t1_on, s);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename R, int RK>
static void launch_batch_rk(const HaldaneArgs *insts, const int32_t *bi, const int32_t *bf, int nblocks, bool t1, bool deph,
                            hipStream_t s)
{
    if (t1) haldane_batch_kernel<R, RK, true, true><<<nblocks, kBlock, 0, s>>>(insts, bi, bf);
    else if (deph) haldane_batch_kernel<R, RK, false, true><<<nblocks, kBlock, 0, s>>>(insts, bi, bf);
    else haldane_batch_kernel<R, RK, false, false><<<nblocks, kBlock, 0, s>>>(insts, bi, bf);
}

hipError_t launch_haldane_batch(const HaldaneArgs *d_insts, const int32_t *d_block_inst, const int32_t *d_block_first,
                                int nblocks, int rk_order, int precision, bool t1_on, bool deph, hipStream_t s)
{
    if (nblocks <= 0) return hipSuccess;
    const bool f32 = precision == SBE_PREC_FP32;
    if (rk_order == SBE_RK4) {
        if (f32) launch_batch_rk<float, 4>(d_insts, d_block_inst, d_block_first, nblocks, t1_on, deph, s);
        else launch_batch_rk<double, 4>(d_insts, d_block_inst, d_block_first, nblocks, t1_on, deph, s);
    } else if (rk_order == SBE_RK5_FEHLBERG) {
        if (f32) launch_batch_rk<float, 5>(d_insts, d_block_inst, d_block_first, nblocks, t1_on, deph, s);
        else launch_batch_rk<double, 5>(d_insts, d_block_inst, d_block_first, nblocks, t1_on, deph, s);
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
