// haldane_velocity.hip - 2-band (Haldane) SBE propagation in the
// VelocityHamiltonian gauge for gfx950.
//
// Reference: SBEs.h:284-373 (per-k E(k), U(k), pMatrix = U^dag J U, rho0),
// SBEs.h:787-805 + 814-829 + 862-871 (RHS), SBEs.h:626-760 (currents at k,
// plus Nval A(t)), antelope.cpp:297-338 (sampling before each step).
//
// Design. The state is the band-basis Bloch vector of rho_H(k) =
// (n I + s.sigma)/2 in the eigenbasis of H(k), whose phases follow
// band_frame.hpp. With J_a = -(d_a d0 I + d_a d.sigma) and the frame f_i:
//   pMatrix_a = -(d_a d0) I - sum_i g_ai sigma_i,   g_ai = d_a d . f_i
//   diag(E) + A.P = const I + h.sigma,  h = (-A.g_x, -A.g_y, -|d| - A.g_z)
//   -i[H, rho]          -> ds/dt = 2 h x s
//   -(rho - rho0) o G   -> ds_x,y/dt -= g2 s_x,y;  ds_z/dt -= g1 (s_z - s0_z)
//                          (n stays n0: both diagonals relax to rho0)
//   Re Tr(rho P_a)      -> intra = -(n d_a d0 + s_z g_az), inter = -(s_x g_ax + s_y g_ay)
// The per-k constants g (6), d_a d0 (2), |d| and s0_z are computed once per
// launch from k and stay in registers with the state for the whole chunk of
// steps; the only per-step input is A(t) at the stage times (a k-independent
// table staged through LDS). The Nval A(t) term of the current is the same
// for every k and is added by the host (times the block's weight sum).
#include "band_frame.hpp"
#include "haldane_dev.hpp"

namespace sbe {
namespace {

using namespace dev;

template <typename R>
struct VPoint {
    R g[2][3];        // d_a d . f_i
    R o[2];           // d_a d0
    R r, s0z;         // |d(k)|, Bloch z of rho0
    R n, sx, sy, sz;  // band-basis state
    R cn, cx, cy, cz; // Kahan compensation (FP32 state only)
    double w;
};

// d(k), its k derivatives and the band frame at k (Haldane.h:86-127)
struct KGeom {
    double d[3], d0, dd[2][3], od[2];
};

__device__ __forceinline__ KGeom haldane_geom(const HaldaneArgs &a, double kx, double ky)
{
    KGeom G;
    double er[6], ei[6];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        sincos(kx * a.va[j][0] + ky * a.va[j][1], &ei[j], &er[j]);
        sincos(kx * a.vb[j][0] + ky * a.vb[j][1], &ei[j + 3], &er[j + 3]);
    }
    G.d[0] = a.t1 * (er[0] + er[1] + er[2]);
    G.d[1] = a.t1 * (ei[0] + ei[1] + ei[2]);
    G.d[2] = a.M0 - a.c3 * (ei[3] + ei[4] + ei[5]);
    G.d0 = a.c0 * (er[3] + er[4] + er[5]);
#pragma unroll
    for (int ax = 0; ax < 2; ++ax) {
        double x = 0, y = 0, z = 0, o = 0;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            x -= a.va[j][ax] * ei[j];
            y += a.va[j][ax] * er[j];
            z += a.vb[j][ax] * er[j + 3];
            o += a.vb[j][ax] * ei[j + 3];
        }
        G.dd[ax][0] = a.t1 * x;
        G.dd[ax][1] = a.t1 * y;
        G.dd[ax][2] = -a.c3 * z;
        G.od[ax] = -a.c0 * o;
    }
    return G;
}

// occupations of the lower / upper band (SBEs.h:318-340)
__device__ __forceinline__ void occupations(const HaldaneArgs &a, double e0, double e1, double &f0, double &f1)
{
    if (a.fermi) {
        const double kT = 300.0 / 3.15775024804e5;
        f0 = 1.0 / (exp((e0 - a.fermi_energy) / kT) + 1.0);
        f1 = 1.0 / (exp((e1 - a.fermi_energy) / kT) + 1.0);
        const double sum = f0 + f1;
        f0 /= sum;
        f1 /= sum;
    } else {
        f0 = a.nval > 0 ? 1.0 : 0.0;
        f1 = a.nval > 1 ? 1.0 : 0.0;
    }
}

constexpr int kTile = 64;  // time steps per LDS tile of the A(t) table

// Step kernel body; R, RK, T1ON, DEPH, KPT as in the length-gauge kernel.
// The stage right-hand side is evaluated as half of the reference RHS with
// the factor 2 folded into the RK coefficients:
//   f/2 = h x y - (g2/2)(y_x, y_y, 0) - (g1/2)(0, 0, y_z - s0_z).
template <typename R, int RK, bool T1ON, bool DEPH, int KPT>
__device__ __forceinline__ void vel_body(const HaldaneArgs &a, const R *__restrict__ tabr, R *__restrict__ st,
                                         double *__restrict__ partials, int64_t blk)
{
    using C = Coef<RK>;
    constexpr int S = C::S;
    constexpr bool kComp = StateT<R>::kComp != 0;
    const int64_t ks = a.nk_stride;
    const int64_t lane0 = blk * (kBlock * KPT) + threadIdx.x;
    const int wave = static_cast<int>((blk * kBlock + threadIdx.x) >> 6);

    VPoint<R> p[KPT];
#pragma unroll
    for (int q = 0; q < KPT; ++q) {
        const int64_t k = lane0 + q * kBlock;
        const KGeom G = haldane_geom(a, a.kpts[k], a.kpts[ks + k]);
        const BandFrame F = band_frame(G.d[0], G.d[1], G.d[2]);
        const double r = sqrt(G.d[0] * G.d[0] + G.d[1] * G.d[1] + G.d[2] * G.d[2]);
#pragma unroll
        for (int ax = 0; ax < 2; ++ax) {
            const double *v = G.dd[ax];
            p[q].g[ax][0] = static_cast<R>(v[0] * F.fx[0] + v[1] * F.fx[1] + v[2] * F.fx[2]);
            p[q].g[ax][1] = static_cast<R>(v[0] * F.fy[0] + v[1] * F.fy[1] + v[2] * F.fy[2]);
            p[q].g[ax][2] = static_cast<R>(v[0] * F.fz[0] + v[1] * F.fz[1] + v[2] * F.fz[2]);
            p[q].o[ax] = static_cast<R>(G.od[ax]);
        }
        p[q].r = static_cast<R>(r);
        double f0 = 0, f1 = 0;
        if (T1ON) occupations(a, G.d0 - r, G.d0 + r, f0, f1);
        p[q].s0z = static_cast<R>(f0 - f1);
        p[q].w = a.weight[k];
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
    const R gh1 = static_cast<R>(0.5 * a.gam1), gh2 = static_cast<R>(0.5 * a.gam2);
    const R ii = a.inter_intra ? R(1) : R(0);
    const R ni = R(1) - ii;
    const int lane = __lane_id();

    constexpr int ROW = S * 2;  // (A_x, A_y) at the S stage times
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

    for (int tile = 0; tile < ntiles; ++tile) {
    if (tile + 1 < ntiles) fetch(tile + 1);
    const int s_end = min(kTile, a.nsteps - tile * kTile);
    for (int st_ = 0; st_ < s_end; ++st_) {
        const int step = tile * kTile + st_;
        const R *tr = &lds_tab[tile & 1][st_ * ROW];

        // ---- time sample at t_n from the state before the step
        double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int q = 0; q < KPT; ++q) {
            const R nn = p[q].n - p[q].cn, x = p[q].sx - p[q].cx, y = p[q].sy - p[q].cy, z = p[q].sz - p[q].cz;
            const double w = p[q].w;
#pragma unroll
            for (int ax = 0; ax < 2; ++ax) {
                const R pin = x * p[q].g[ax][0] + y * p[q].g[ax][1];  // inter part
                const R pia = nn * p[q].o[ax] + z * p[q].g[ax][2];    // intra part
                v[ax] -= static_cast<double>(pin + ni * pia) * w;
                v[2 + ax] -= static_cast<double>(ii * pia) * w;
            }
            v[4] += static_cast<double>(x) * w;
            v[5] += static_cast<double>(y) * w;
            v[6] += static_cast<double>(z) * w;
            v[7] -= static_cast<double>(z) * w;  // band populations (n -+ v7)/2 as in the length gauge
        }
        {
            const double r = wave_reduce8(v);
            partials[(static_cast<int64_t>(step) * a.nwaves + wave) * kHaldaneNV + (lane >> 3)] = r;
        }

        // ---- one RK step (SBEs.h:596-624)
#pragma unroll
        for (int q = 0; q < KPT; ++q) {
            Vec3<R> kv[S];
            Vec3<R> hs[S];  // h at each stage time (stages at an equal c share it)
#pragma unroll
            for (int i = 0; i < S; ++i) {
                Vec3<R> y{p[q].sx, p[q].sy, p[q].sz};
#pragma unroll
                for (int j = 0; j < i; ++j) {
                    if (C::a[i][j] != 0.0) {
                        const R aij = static_cast<R>(2.0 * C::a[i][j] * a.dt);
                        y.x += aij * kv[j].x;
                        y.y += aij * kv[j].y;
                        y.z += aij * kv[j].z;
                    }
                }
                const int same = same_time_stage<RK>(i);
                Vec3<R> h;
                if (same > 0) {
                    h = hs[same];
                } else {
                    const R Ax = tr[2 * i], Ay = tr[2 * i + 1];
                    h = {-(Ax * p[q].g[0][0] + Ay * p[q].g[1][0]), -(Ax * p[q].g[0][1] + Ay * p[q].g[1][1]),
                         -p[q].r - (Ax * p[q].g[0][2] + Ay * p[q].g[1][2])};
                }
                hs[i] = h;
                Vec3<R> f;
                f.x = h.y * y.z - h.z * y.y;
                f.y = h.z * y.x - h.x * y.z;
                f.z = h.x * y.y - h.y * y.x;
                if (DEPH) {
                    f.x -= gh2 * y.x;
                    f.y -= gh2 * y.y;
                    if (T1ON) f.z -= gh1 * (y.z - p[q].s0z);
                }
                kv[i] = f;
            }
            Vec3<R> ds{0, 0, 0};
#pragma unroll
            for (int i = 0; i < S; ++i) {
                if (C::b[i] != 0.0) {
                    const R bi = static_cast<R>(2.0 * C::b[i] * a.dt);
                    ds.x += bi * kv[i].x;
                    ds.y += bi * kv[i].y;
                    ds.z += bi * kv[i].z;
                }
            }
            if (kComp) {
                auto kahan = [](R &sum, R &c, R inc) {
                    const R y = inc - c;
                    const R t = sum + y;
                    c = (t - sum) - y;
                    sum = t;
                };
                kahan(p[q].sx, p[q].cx, ds.x);
                kahan(p[q].sy, p[q].cy, ds.y);
                kahan(p[q].sz, p[q].cz, ds.z);
            } else {
                p[q].sx += ds.x;
                p[q].sy += ds.y;
                p[q].sz += ds.z;
            }
        }
    }
    if (tile + 1 < ntiles) stash((tile + 1) & 1);
    __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < KPT; ++q) {
        const int64_t k = lane0 + q * kBlock;
        st[ks + k] = p[q].sx;
        st[2 * ks + k] = p[q].sy;
        st[3 * ks + k] = p[q].sz;
        if (kComp) {
            st[5 * ks + k] = p[q].cx;
            st[6 * ks + k] = p[q].cy;
            st[7 * ks + k] = p[q].cz;
        }
    }
}

template <typename R, int RK, bool T1ON, bool DEPH, int KPT>
__global__ __launch_bounds__(kBlock) void haldane_vel_kernel(const HaldaneArgs a, const R *__restrict__ tabr,
                                                             R *__restrict__ st, double *__restrict__ partials)
{
    vel_body<R, RK, T1ON, DEPH, KPT>(a, tabr, st, partials, blockIdx.x);
}

// Initial state: rho_H = diag(f0, f1) -> n = f0 + f1, s = (0, 0, f0 - f1).
// Padding points get a zero state.
template <typename R>
__global__ __launch_bounds__(kBlock) void haldane_vel_init_kernel(HaldaneArgs a)
{
    const int64_t k = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    const int64_t ks = a.nk_stride;
    R *st = static_cast<R *>(a.state);
    double n = 0, sz = 0;
    if (k < a.nk) {
        const KGeom G = haldane_geom(a, a.kpts[k], a.kpts[ks + k]);
        const double r = sqrt(G.d[0] * G.d[0] + G.d[1] * G.d[1] + G.d[2] * G.d[2]);
        double f0, f1;
        occupations(a, G.d0 - r, G.d0 + r, f0, f1);
        n = f0 + f1;
        sz = f0 - f1;
    }
    st[k] = static_cast<R>(n);
    st[ks + k] = R(0);
    st[2 * ks + k] = R(0);
    st[3 * ks + k] = static_cast<R>(sz);
    if (StateT<R>::kComp) {
        st[4 * ks + k] = static_cast<R>(static_cast<double>(static_cast<R>(n)) - n);
        st[5 * ks + k] = R(0);
        st[6 * ks + k] = R(0);
        st[7 * ks + k] = static_cast<R>(static_cast<double>(static_cast<R>(sz)) - sz);
    }
}

template <typename R, int RK, int KPT>
void launch_vel(const HaldaneArgs &a, const R *tab, bool t1_on, hipStream_t s)
{
    const dim3 grid(static_cast<unsigned>(a.nk_stride / (kBlock * KPT)));
    R *st = static_cast<R *>(a.state);
    const bool deph = a.gam1 != 0.0 || a.gam2 != 0.0;
    if (t1_on) haldane_vel_kernel<R, RK, true, true, KPT><<<grid, kBlock, 0, s>>>(a, tab, st, a.partials);
    else if (deph) haldane_vel_kernel<R, RK, false, true, KPT><<<grid, kBlock, 0, s>>>(a, tab, st, a.partials);
    else haldane_vel_kernel<R, RK, false, false, KPT><<<grid, kBlock, 0, s>>>(a, tab, st, a.partials);
}

template <typename R, int RK>
void launch_vel_kpt(const HaldaneArgs &a, const R *tab, bool t1_on, hipStream_t s)
{
    if (a.kpt == 2) launch_vel<R, RK, 2>(a, tab, t1_on, s);
    else launch_vel<R, RK, 1>(a, tab, t1_on, s);
}

}  // namespace

hipError_t launch_haldane_vel_init(const HaldaneArgs &a, hipStream_t s)
{
    const dim3 grid(static_cast<unsigned>(a.nk_stride / kBlock));
    if (a.precision == SBE_PREC_FP32) haldane_vel_init_kernel<float><<<grid, kBlock, 0, s>>>(a);
    else haldane_vel_init_kernel<double><<<grid, kBlock, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_haldane_vel_steps(const HaldaneArgs &a, int rk_order, bool t1_on, hipStream_t s)
{
    const bool f32 = a.precision == SBE_PREC_FP32;
    if (rk_order == SBE_RK4) {
        if (f32) launch_vel_kpt<float, 4>(a, a.phase_tab32, t1_on, s);
        else launch_vel_kpt<double, 4>(a, a.phase_tab, t1_on, s);
    } else if (rk_order == SBE_RK5_FEHLBERG) {
        if (f32) launch_vel_kpt<float, 5>(a, a.phase_tab32, t1_on, s);
        else launch_vel_kpt<double, 5>(a, a.phase_tab, t1_on, s);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace sbe
