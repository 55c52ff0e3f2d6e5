// nband16.hip - N-band (Wannier) stage and sample kernels for N <= 16 with
// every matrix in MFMA registers.
//
// Same mathematics as nb_rhs_kernel / nb_sample_kernel (SBEs.h:762-861 RHS,
// SBEs.h:694-760 currents), restructured for one 16x16 FP64 MFMA tile per
// matrix and one wave per k point:
//  * A 16x16 complex matrix is a Tile: lane l = c + 16 g holds the elements
//    (g + 4r, c), r = 0..3 - the D layout of v_mfma_f64_16x16x4f64. Bands
//    beyond N are zero padding (identity in the eigenvectors).
//  * C = A B runs as four real MFMAs per K block of 4. The B operand of
//    block kb is register kb of B's tile; the A operand is register kb of
//    A^T's tile. For Hermitian A (H, H~, rho) and for A = X^dag that is
//    the conjugate of A's own tile, so products of that form need no data
//    movement; a transpose (U^T, T^dag) goes through a padded 16x17 LDS
//    buffer of the wave, conflict-free.
//  * Eigenvectors U(k + A(t)) are refined from the stored ones (warm start)
//    by simultaneous Jacobi rotations written as matrix products: with
//    W = U^dag H U nearly diagonal, K_ij = W_ij / (D/2 + s sqrt(D^2/4 +
//    |W_ij|^2)), D = W_jj - W_ii (the 2x2 Jacobi angle; W_ij / D for small
//    W_ij, bounded for near-degenerate pairs), K anti-Hermitian;
//    U <- U (I + K + K^2/2), W <- V^dag W V until max|K| < 1e-8, then a last
//    U (I + K) and one Newton-Schulz step U (3I - U^dag U)/2 to keep U
//    unitary to rounding. After a round with max|K| > 1e-3 (V unitary only to
//    |K|^4 / 4) U is made unitary again by Newton-Schulz steps until
//    |U^dag U - I| < 1e-8 and W is rebuilt from H; rounds with max|K| > 0.5
//    are damped to 0.5 (a cold-like start converges in ~10 rounds).
//    Convergence is quadratic; between RK stages two or three rounds reach
//    machine precision. The columns keep the order of
//    the warm start (bands are followed continuously); the sample kernel
//    ranks them by eigenvalue, as the reference's ascending zheevr order.
//  * The dephasing U (Gamma o (U^dag y U)) U^dag is g2 y + (g1 - g2)
//    U diag(U^dag y U) U^dag (Gamma = g1 on the diagonal, g2 off it): one
//    product y U, a column reduction and one product U diag(d) U^dag.
//  * Currents: Tr(rho J_a) elementwise; the intraband part sum_i
//    (U^dag rho U)_ii (U^dag J_a U)_ii from the products rho U and J_a U.
//    J_a (and pMatrix_a below) is read transposed and used as a general
//    matrix: some model classes' current matrices are not Hermitian.
//  * VelocityHamiltonian gauge (band basis): one product H y per stage with
//    H = sum_a A_a P_a + diag(E) formed elementwise, the stage derivatives
//    kept in registers.
#include <cmath>
#include <map>
#include <mutex>

#include "nband.hpp"
#include "tile16.hpp"

namespace sbe {
namespace {

using namespace tile;


// U <- U (3I - U^dag U) / 2 (Newton-Schulz: unitarity error e -> ~e^2);
// returns e = max |U^dag U - I| before the step (over the wave)
__device__ __forceinline__ double orthonormalize(Tile &U, Tile &Ut, cd *buf)
{
    const int c = lane_c(), g = lane_g();
    const Tile G = cmm(U, true, U);
    Tile S;
    double e = 0.0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const double dr = G.re[r] - (g + 4 * r == c ? 1.0 : 0.0);
        e = fmax(e, fmax(fabs(dr), fabs(G.im[r])));
        S.re[r] = (g + 4 * r == c ? 1.5 : 0.0) - 0.5 * G.re[r];
        S.im[r] = -0.5 * G.im[r];
    }
    U = cmm(Ut, false, S);
    Ut = transpose(U, buf);
    return wave_max(e);
}

// Warm-started eigenvectors of the Hermitian part of H (as zheevr, which
// reads one triangle; H from FP16-split sums is Hermitian only to their
// error): U (and its transpose Ut) in, refined U, Ut out.
// small: max|K| below which the last round is U (I + K) (its neglected K^2/2 and the
// next round's O(K^2) residual are the eigenvectors' error); ortho: end with a
// Newton-Schulz step (U unitary to rounding rather than to O(small^2))
// herm: Hin is Hermitian to rounding already (FP64 sums), no symmetrization; lam_out:
// this lane's column of diag(U^dag H U) before the last rotation (the band energies to
// O(max|K|^2), for ranking)
__device__ void eig_refine(const Tile &Hin, Tile &U, Tile &Ut, int N, cd *buf, double small_tol = 1e-8,
                           bool ortho = true, bool herm = false, double *lam_out = nullptr)
{
    const int c = lane_c(), g = lane_g();
    Tile H = Hin;
    if (!herm) {
        const Tile T = transpose(Hin, buf);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            H.re[r] = 0.5 * (Hin.re[r] + T.re[r]);
            H.im[r] = 0.5 * (Hin.im[r] - T.im[r]);
        }
    }
    Tile W = cmm(U, true, cmm(H, true, U));  // U^dag (H U)
    double lam_last = 0.0;
    for (int it = 0; it < 30; ++it) {
        double lam_r[4], lam_c;
        diag_bcast(W, lam_r, lam_c);
        lam_last = lam_c;
        Tile K;
        double kmax = 0.0;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = g + 4 * r;
            K.re[r] = 0.0;
            K.im[r] = 0.0;
            const double w2 = W.re[r] * W.re[r] + W.im[r] * W.im[r];
            if (i != c && i < N && c < N && w2 > 0.0) {
                const double D = lam_c - lam_r[r];
                const double s = (D > 0.0 || (D == 0.0 && i < c)) ? 1.0 : -1.0;
                const double inv = 1.0 / (0.5 * D + s * sqrt(0.25 * D * D + w2));
                K.re[r] = W.re[r] * inv;
                K.im[r] = W.im[r] * inv;
                kmax = fmax(kmax, K.re[r] * K.re[r] + K.im[r] * K.im[r]);  // |K|^2 here, the root once below
            }
        }
        kmax = sqrt(wave_max(kmax));
        if (kmax == 0.0) break;
        if (kmax > 0.5) {  // far from converged (a cold-like start): damped rotations
            const double f = 0.5 / kmax;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                K.re[r] *= f;
                K.im[r] *= f;
            }
        }
        Tile V;
        const bool small = kmax < small_tol;
        if (small) {  // I + K: the neglected K^2 / 2 is below 1e-16
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                V.re[r] = K.re[r] + (g + 4 * r == c ? 1.0 : 0.0);
                V.im[r] = K.im[r];
            }
        } else {  // I + K + K^2/2, unitary to |K|^4 / 4; K^T's tile is -conj(K)
            Tile Kn;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                Kn.re[r] = -K.re[r];
                Kn.im[r] = -K.im[r];
            }
            const Tile K2 = cmm(Kn, true, K);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                V.re[r] = K.re[r] + 0.5 * K2.re[r] + (g + 4 * r == c ? 1.0 : 0.0);
                V.im[r] = K.im[r] + 0.5 * K2.im[r];
            }
        }
        U = cmm(Ut, false, V);  // U V
        Ut = transpose(U, buf);
        if (small) break;
        if (kmax > 1e-3) {  // V far from unitary: restore U fully, then W from H itself
            for (int q = 0; q < 8; ++q)
                if (orthonormalize(U, Ut, buf) < 1e-8) break;
            W = cmm(U, true, cmm(H, true, U));
        } else {
            W = cmm(V, true, cmm(W, true, V));  // V^dag (W V)
        }
    }
    if (ortho) orthonormalize(U, Ut, buf);
    if (lam_out) *lam_out = lam_last;
}

// ------------------------------------------------------------- RK stage --
// The band-diagonal part of y, U diag(U^dag y U) U^dag, and this lane's column
// of its diagonal d = (U^dag y U)_cc (two products)
__device__ __forceinline__ Tile band_diag(const Tile &Y, const Tile &U, const Tile &Ut, double &d)
{
    const int g = lane_g();
    const Tile G = cmm(Y, true, U);  // y U
    d = col_dot(U, G);               // (U^dag y U)_cc
    Tile Ud;  // (U diag(d))^T's tile: U^T[4kb + g][c] d_{4kb+g}
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const double dr = __shfl(d, 4 * r + g, 64);
        Ud.re[r] = Ut.re[r] * dr;
        Ud.im[r] = Ut.im[r] * dr;
    }
    return cmm(Ud, false, Ut, true);  // U diag(d) U^dag (B = U^dag: tile conj(U^T))
}

// The stage derivative at y given H~^T's tile (HERM: H~'s own tile, H~ Hermitian, so
// that H~^T's tile is its conjugate) and (dephasing on) y's band-diagonal part M:
// -i(H~ y - (H~ y)^dag) - g2 y - (g1 - g2) M.
template <bool HERM = false>
__device__ __forceinline__ Tile stage_derivative_m(const NbArgs &a, const Tile &Y, const Tile &HtT, const Tile &M,
                                                   cd *buf)
{
    const Tile T = cmm(HtT, HERM, Y);  // H~ y
    const Tile Tt = transpose(T, buf);
    Tile X;  // -i (T - T^dag)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const double dr = T.re[r] - Tt.re[r], di = T.im[r] + Tt.im[r];
        X.re[r] = di;
        X.im[r] = -dr;
    }
    if (a.deph) {
        const double g2 = a.gam2, g12 = a.gam1 - a.gam2;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            X.re[r] -= g2 * Y.re[r] + g12 * M.re[r];
            X.im[r] -= g2 * Y.im[r] + g12 * M.im[r];
        }
    }
    return X;
}

// The same with U, U^T given (M formed here when dephasing)
__device__ __forceinline__ Tile stage_derivative(const NbArgs &a, const Tile &Y, const Tile &HtT, const Tile &U,
                                                 const Tile &Ut, cd *buf)
{
    Tile M{};
    if (a.deph) {
        double d;
        M = band_diag(Y, U, Ut, d);
    }
    return stage_derivative_m(a, Y, HtT, M, buf);
}

// Stage i: y = rho + sum_j dt a_ij k_j; k_i = f(y); after the last stage
// rho += sum_i dt b_i k_i. PAIR: stage sg and the next stage sg2 at the same
// time (RK4's c = 1/2, 1/2) in one pass, sharing H~, U and rho in registers;
// sg2's input takes sg's derivative from registers.
template <bool PAIR>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(PAIR ? 2 : 3))) void nb16_rhs_kernel(
    const NbArgs a, const NbStage sg, const NbStage sg2, const cd *__restrict__ mats, cd *__restrict__ rho,
    cd *__restrict__ kbuf, cd *__restrict__ Ust)
{
    __shared__ cd buf[16 * kTpad];
    const int64_t k = blockIdx.x;
    if (k >= a.nk) return;
    const int N = a.N, nn = N * N;
    const int64_t kstride = a.nk_stride * nn;
    cd *rk = rho + k * nn;
    // rho is read again where needed rather than held across the stages (registers)
    Tile Y = load_tile(rk, N);
    for (int j = 0; j < sg.stage; ++j) {
        if (sg.a[j] == 0.0) continue;
        const Tile Kj = load_tile(kbuf + j * kstride + k * nn, N);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            Y.re[r] += sg.a[j] * Kj.re[r];
            Y.im[r] += sg.a[j] * Kj.im[r];
        }
    }
    const Tile HtT = load_tile_t(mats + k * nn, N);  // mat 0: H~ = H + E.D, transposed (A operand)
    Tile U, Ut;
    if (a.deph) {
        U = load_tile(Ust + k * nn, N, true);
        Ut = transpose(U, buf);
        if (!sg.ureuse) {  // eigenvectors at this stage time, warm-started
            const Tile H = load_tile(mats + a.nk_stride * nn + k * nn, N);  // mat 1: H
            eig_refine(H, U, Ut, N, buf);
            store_tile(Ust + k * nn, U, N);
        }
    }
    Tile X = stage_derivative(a, Y, HtT, U, Ut, buf);
    store_tile(kbuf + sg.stage * kstride + k * nn, X, N);
    const NbStage &fin = PAIR ? sg2 : sg;  // the stage whose derivative X holds at the end
    if (PAIR) {
        Tile Y2 = load_tile(rk, N);
        for (int j = 0; j < sg2.stage; ++j) {
            if (sg2.a[j] == 0.0) continue;
            const Tile Kj = j == sg.stage ? X : load_tile(kbuf + j * kstride + k * nn, N);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                Y2.re[r] += sg2.a[j] * Kj.re[r];
                Y2.im[r] += sg2.a[j] * Kj.im[r];
            }
        }
        X = stage_derivative(a, Y2, HtT, U, Ut, buf);
        store_tile(kbuf + sg2.stage * kstride + k * nn, X, N);
    }
    if (fin.last) {
        Tile Rn = load_tile(rk, N);
        for (int j = 0; j < fin.nstage; ++j) {
            if (fin.b[j] == 0.0) continue;
            const Tile Kj = j == fin.stage ? X : load_tile(kbuf + j * kstride + k * nn, N);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                Rn.re[r] += fin.b[j] * Kj.re[r];
                Rn.im[r] += fin.b[j] * Kj.im[r];
            }
        }
        store_tile(rk, Rn, N);
    }
}

// ------------------------------------------------------------ sample --
// Observables at t_n (mats slots 1..4: H, J_x, J_y, J_z at k + A(t_n)); each wave walks
// k = wave, wave + G, ... and writes one row of partial sums
// [inter(3) intra(3) pop(N) rho(2 N^2)], as nb_sample_kernel.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void nb16_sample_kernel(const NbArgs a, const cd *__restrict__ mats,
                                                         const cd *__restrict__ rho, double *__restrict__ partials,
                                                         int64_t step_local, cd *__restrict__ Ust, int ureuse)
{
    __shared__ cd buf[16 * kTpad];
    const int N = a.N, nn = N * N;
    const int c = lane_c(), g = lane_g();
    const int lane = threadIdx.x;
    double racc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double cur[6] = {0, 0, 0, 0, 0, 0};
    double pop = 0.0;  // lane m < N: population of the m-th lowest band
    for (int64_t k = blockIdx.x; k < a.nk; k += gridDim.x) {
        const double w = a.weight[k];
        const Tile R = load_tile(rho + k * nn, N);
        const Tile H = load_tile(mats + a.nk_stride * nn + k * nn, N);  // mat 1: H
        Tile U = load_tile(Ust + k * nn, N, true);
        if (!ureuse) {  // eigenvectors at k + A(t_n), warm-started from the last ones
            Tile Ut = transpose(U, buf);
            eig_refine(H, U, Ut, N, buf);
            store_tile(Ust + k * nn, U, N);
        }
        // band energies (for the ascending order) and populations of U's columns
        const double lam = col_dot(U, cmm(H, true, U));
        const double p = col_dot(U, cmm(R, true, U));
        int rank = 0;  // rank of column c among the N energies, ties by index
        for (int j = 0; j < N; ++j) {
            const double lj = __shfl(lam, j, 64);
            rank += (lj < lam) || (lj == lam && j < c);
        }
        for (int j = 0; j < N; ++j) {
            const int rj = __shfl(rank, j, 64);
            const double pj = __shfl(p, j, 64);
            if (rj == lane) pop += pj * w;
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            racc[2 * r] += R.re[r] * w;
            racc[2 * r + 1] += R.im[r] * w;
        }
        for (int ax = 0; ax < 3; ++ax) {
            // mats 2..4: J_a, read transposed: some model classes' J is not Hermitian
            const Tile Jt = load_tile_t(mats + (2 + ax) * a.nk_stride * nn + k * nn, N);
            double tot = 0.0;  // Re Tr(rho J) = sum_ij Re(rho_ij J_ji)
#pragma unroll
            for (int r = 0; r < 4; ++r) tot += R.re[r] * Jt.re[r] - R.im[r] * Jt.im[r];
            tot = wave_sum(tot);
            double intra = 0.0;
            if (a.inter_intra) {
                const double jd = col_dot(U, cmm(Jt, false, U));  // Re (U^dag J U)_cc
                intra = wave_sum(lane < 16 && c < N ? p * jd : 0.0);
            }
            cur[ax] += (tot - intra) * w;
            cur[3 + ax] += intra * w;
        }
    }
    double *row = partials + (step_local * gridDim.x + blockIdx.x) * static_cast<int64_t>(a.nv);
    if (lane < 6) {
        double v = 0.0;
#pragma unroll
        for (int q = 0; q < 6; ++q)
            if (lane == q) v = cur[q];
        row[lane] = v;
    }
    if (lane < N) row[6 + lane] = pop;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = g + 4 * r;
        if (i < N && c < N) {
            row[6 + N + 2 * (i * N + c)] = racc[2 * r];
            row[6 + N + 2 * (i * N + c) + 1] = racc[2 * r + 1];
        }
    }
}

// ------------------------------------------------- velocity gauge ----
// One time step of the VelocityHamiltonian gauge (SBEs.h:762-861 with the
// band-basis H = sum_a A_a(t) P_a + diag(E)), as nb_vel_step_kernel: the
// sample at t_n, S stages with the derivatives in registers, the update.
// Each wave walks k = wave, wave + G, ... and writes one row of partial sums.
template <int S>
__global__ __launch_bounds__(64) void nb16_vel_step_kernel(const NbArgs a, const NbVelStep v,
                                                           const cd *__restrict__ P, const double *__restrict__ Eocc,
                                                           cd *__restrict__ rho, double *__restrict__ partials,
                                                           int64_t step_local)
{
    __shared__ cd buf[16 * kTpad];
    const int N = a.N, nn = N * N;
    const int c = lane_c(), g = lane_g();
    const int lane = threadIdx.x;
    const int64_t ks = a.nk_stride * nn;
    double racc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double cur[6] = {0, 0, 0, 0, 0, 0};
    double pop = 0.0;
    for (int64_t k = blockIdx.x; k < a.nk; k += gridDim.x) {
        const double w = a.weight[k];
        const Tile R = load_tile(rho + k * nn, N);
        // pMatrix_a read transposed: not Hermitian for the model classes whose J is not
        const Tile Px = load_tile_t(P + k * nn, N), Py = load_tile_t(P + ks + k * nn, N),
                   Pz = load_tile_t(P + 2 * ks + k * nn, N);
        const double Ec = c < N ? Eocc[k * 2 * N + c] : 0.0;       // E of band c
        const double oc = c < N ? Eocc[k * 2 * N + N + c] : 0.0;   // occupation of band c
        // ---- sample at t_n: band populations, currents Tr(rho P_a) split
        // into the diagonal (intraband) and off-diagonal parts, rho sums
        double lr[4], lc;
        diag_bcast(R, lr, lc);
        if (lane < N) pop += lc * w;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            racc[2 * r] += R.re[r] * w;
            racc[2 * r + 1] += R.im[r] * w;
        }
        for (int ax = 0; ax < 3; ++ax) {
            const Tile &Pa = ax == 0 ? Px : (ax == 1 ? Py : Pz);
            double inter = 0.0, intra = 0.0;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const double val = R.re[r] * Pa.re[r] - R.im[r] * Pa.im[r];  // Re(rho_ij P_ji)
                if (g + 4 * r == c && a.inter_intra) intra += val;
                else inter += val;
            }
            cur[ax] += wave_sum(inter) * w;
            cur[3 + ax] += wave_sum(intra) * w;
        }
        // ---- RK stages
        Tile Kd[S];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            Tile Y = R;
#pragma unroll
            for (int j = 0; j < s; ++j) {
                if (v.a[s][j] == 0.0) continue;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    Y.re[r] += v.a[s][j] * Kd[j].re[r];
                    Y.im[r] += v.a[s][j] * Kd[j].im[r];
                }
            }
            const double A0 = v.A[s][0], A1 = v.A[s][1], A2 = v.A[s][2];
            Tile M;  // H^T = sum_a A_a P_a^T + diag(E)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                M.re[r] = A0 * Px.re[r] + A1 * Py.re[r] + A2 * Pz.re[r] + (g + 4 * r == c ? Ec : 0.0);
                M.im[r] = A0 * Px.im[r] + A1 * Py.im[r] + A2 * Pz.im[r];
            }
            const Tile T = cmm(M, false, Y);  // H y
            const Tile Tt = transpose(T, buf);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const double dr = T.re[r] - Tt.re[r], di = T.im[r] + Tt.im[r];
                double xr = di, xi = -dr;  // -i (H y - (H y)^dag)
                if (a.deph) {
                    const bool dg = g + 4 * r == c;
                    const double gm = dg ? a.gam1 : a.gam2;
                    xr -= gm * (Y.re[r] - (dg ? oc : 0.0));
                    xi -= gm * Y.im[r];
                }
                Kd[s].re[r] = xr;
                Kd[s].im[r] = xi;
            }
        }
        Tile Rn = R;
#pragma unroll
        for (int j = 0; j < S; ++j) {
            if (v.b[j] == 0.0) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                Rn.re[r] += v.b[j] * Kd[j].re[r];
                Rn.im[r] += v.b[j] * Kd[j].im[r];
            }
        }
        store_tile(rho + k * nn, Rn, N);
    }
    double *row = partials + (step_local * gridDim.x + blockIdx.x) * static_cast<int64_t>(a.nv);
    if (lane < 6) {
        double vv = 0.0;
#pragma unroll
        for (int q = 0; q < 6; ++q)
            if (lane == q) vv = cur[q];
        row[lane] = vv;
    }
    if (lane < N) row[6 + lane] = pop;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = g + 4 * r;
        if (i < N && c < N) {
            row[6 + N + 2 * (i * N + c)] = racc[2 * r];
            row[6 + N + 2 * (i * N + c) + 1] = racc[2 * r + 1];
        }
    }
}


// ------------------------------------------------- fused RK4 step ----
// One RK4 step per launch (SBEs.h:596-624 with the stage right-hand side of
// SBEs.h:762-861 and the sample of SBEs.h:694-760), for a product k grid with
// separable Fourier sums. Block = W k points that share (k2, k3) and differ
// along a1, one wave each (W = 8; W = 4 where the grid has fewer blocks than
// CUs, so each wave has a SIMD of its own). The sums' last pass (along a1),
//   X(k1, k2, k3) = sum_m E1[k1][m] T2[m][k2][k3],
// runs here: the block stages T2's M1 rows of its column for two matrices at a
// time in LDS and every wave forms its own tiles from them, so H~, H and J of
// the three times of a step are read once per column instead of written and
// read back per k point. Per wave and step: the sample at t_n (U, rho and the
// band energies from the last step), stage 1 at t_n, stages 2 and 3 at
// t_n + dt/2 (eigenvectors refined there), stage 4 at t_n + dt (eigenvectors
// refined for the next sample); rho, U and the energies are the only per-k
// HBM traffic. The W waves' sample sums are added in LDS in wave order into
// one partial row per block.
// dynamic LDS of a block: the waves' transpose buffers, then two buffers of two
// staging slots of kNbFusedMaxM1 x 256 complex each (row m of the second-pass output
// at m * 256, so the m offsets are immediates). While a block computes from one
// buffer, the other fills straight from global memory (global_load_lds); a buffer
// also holds the block's sample sums (W nv doubles) once read.
constexpr int kFusedPitch = 256;
constexpr int64_t kFusedSlot = static_cast<int64_t>(kNbFusedMaxM1) * kFusedPitch;  // complex per slot
template <int W>
constexpr int64_t fused_tb() { return static_cast<int64_t>(W) * 16 * kTpad; }  // complex of the tb

template <int W>
__device__ __forceinline__ cd *fused_slot(cd *lds, int buf, int slot) { return lds + fused_tb<W>() + (2 * buf + slot) * kFusedSlot; }

// This wave's share of filling buffer buf with matrix j0 of src0 and j1 of src1 (src1
// may be null), rows m < M1 of column col: chunks of 64 complex go straight to LDS,
// each lane one element; their completion is counted by vmcnt (fused_wait).
template <int W>
__device__ __forceinline__ void stage_async(const NbArgs &a, cd *lds, int buf, const cd *src0, int j0, const cd *src1,
                                            int j1, int64_t col, int w, int lane)
{
    const int nn = a.N * a.N, M1 = a.boxM[0];
    const int Q = (nn + 63) >> 6;  // chunks per row
    const int64_t mstride = static_cast<int64_t>(a.boxK[1]) * a.boxK[2] * a.bcols;
    const int64_t cbase = col * a.bcols;
    const int per = M1 * Q;
    const int total = src1 ? 2 * per : per;
    for (int ch = w; ch < total; ch += W) {  // uniform over the wave
        const int slot = ch >= per ? 1 : 0;
        const int rem = ch - slot * per;
        const int m = rem / Q, q = rem - m * Q;
        const cd *src = slot ? src1 : src0;
        const int j = slot ? j1 : j0;
        const int ij = q * 64 + lane;
        cd *dst = lds + fused_tb<W>() + (2 * buf + slot) * kFusedSlot + m * kFusedPitch + q * 64;
        if (ij < nn) __builtin_amdgcn_global_load_lds(src + m * mstride + cbase + j * nn + ij, dst, 16, 0, 0);
    }
}

// every wave's staging loads have landed and every wave is past its last reads
__device__ __forceinline__ void fused_wait()
{
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
}

// this lane's elements of X = sum_m E1[m] Y[m] from a staging slot, in the order of
// operations of the a1 pass of the separable sums
__device__ __forceinline__ Tile form_tile(const cd *ys, const cd *__restrict__ e1, int M1, int N)
{
    const int c = lane_c(), g = lane_g();
    int off[4];
    bool ok[4];
    double re[4], im[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = g + 4 * r;
        ok[r] = i < N && c < N;
        off[r] = ok[r] ? i * N + c : 0;
        re[r] = im[r] = 0.0;
    }
#pragma unroll
    for (int m = 0; m < kNbFusedMaxM1; ++m) {
        if (m < M1) {
            const cd w = e1[m];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const cd x = ys[m * kFusedPitch + off[r]];
                re[r] = fma(w.x, x.x, fma(-w.y, x.y, re[r]));
                im[r] = fma(w.x, x.y, fma(w.y, x.x, im[r]));
            }
        }
    }
    Tile T;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        T.re[r] = ok[r] ? re[r] : 0.0;
        T.im[r] = ok[r] ? im[r] : 0.0;
    }
    return T;
}

// x of lane j (j uniform over the wave)
__device__ __forceinline__ double lane_value(double x, int j)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(x), j);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(x), j);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ void axpy_tile(Tile &Y, const Tile &X, double s, const Tile &B)
{  // Y = B + s X
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        Y.re[r] = fma(s, X.re[r], B.re[r]);
        Y.im[r] = fma(s, X.im[r], B.im[r]);
    }
}

// the refinements' last-round threshold: eigenvector error O(1e-12), one rotation
// round fewer than 1e-8 where the first round turns by ~1e-4
constexpr double kFusedSmallK = 1e-6;

template <int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(2))) void nb16_step_kernel(
    const NbArgs a, const NbFused f, cd *__restrict__ rho, cd *__restrict__ Ust, double *__restrict__ partials)
{
    extern __shared__ cd step_lds[];
    const int N = a.N, nn = N * N;
    const int c = lane_c(), g = lane_g(), lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
    const int K1 = a.boxK[0], K2 = a.boxK[1], K3 = a.boxK[2], M1 = a.boxM[0];
    // blocks of one column adjacent in dispatch order on one XCD (blocks go round robin
    // over the eight XCDs), so the column's T2 rows are shared through that XCD's L2
    const int nb = static_cast<int>(gridDim.x), b = static_cast<int>(blockIdx.x);
    const int lin = nb % 8 == 0 ? (b % 8) * (nb / 8) + b / 8 : b;
    const int col = lin / f.g1, grp = lin - col * f.g1;
    const int k2 = col / K3, k3 = col - k2 * K3;
    const int k1 = grp * W + w;
    // a k block of the grid leaves holes in its box (k map -1): those waves only stage
    const int64_t kmap = k1 < K1 ? a.box_kinv[(static_cast<int64_t>(k1) * K2 + k2) * K3 + k3] : -1;
    const bool active = kmap >= 0;
    const int64_t k = active ? kmap : 0;
    cd *buf = step_lds + w * 16 * kTpad;
    const double dt = f.dt;
    // this wave's a1 factors E1[k1][m] in LDS (read as broadcasts by every tile formed)
    cd *e1 = step_lds + fused_tb<W>() + 4 * kFusedSlot + w * kNbFusedMaxM1;
    if (lane < M1) e1[lane] = a.boxE[0][static_cast<int64_t>(active ? k1 : 0) * M1 + lane];

    Tile R, U, Ut, Y, Acc, Ht, X;
    double lam = 0.0, p = 0.0;
    double cur[6] = {0, 0, 0, 0, 0, 0};
    double pop = 0.0;
    double racc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const double wk = active ? a.weight[k] : 0.0;
    const bool need_h = !f.u_valid || !f.lam_valid;
    // staging order: [H(t_n) if needed] [J_x J_y] [J_z H~(t_n)] [H~ H at t + dt/2] [H~ H at t + dt],
    // buffers alternating, each filled while the previous one is read
    if (need_h) stage_async<W>(a, step_lds, 1, f.t2_prev, 1, nullptr, 0, col, w, lane);
    stage_async<W>(a, step_lds, 0, f.t2_j, 0, f.t2_j, 1, col, w, lane);
    if (active) {
        R = load_tile(rho + k * nn, N);
        U = load_tile(Ust + k * nn, N, true);
        Ut = transpose(U, buf);
    }
    // ---- sample at t_n
    if (need_h) {  // eigenvectors / energies at t_n from H(t_n)
        fused_wait();
        if (active) {
            const Tile H = form_tile(fused_slot<W>(step_lds, 1, 0), e1, M1, N);
            if (!f.u_valid) eig_refine(H, U, Ut, N, buf);
            lam = col_dot(U, cmm(H, true, U));
        }
    } else if (active) {
        lam = f.lam[k * 16 + c];
    }
    // rho's band-diagonal part: the populations, the intraband currents Tr(M J) and
    // stage 1's dephasing
    Tile M;
    if (active) {
        M = band_diag(R, U, Ut, p);  // p: populations of U's columns
        int rank = 0;                // rank of column c among the N energies, ties by index
        for (int j = 0; j < N; ++j) {
            const double lj = lane_value(lam, j);
            rank += (lj < lam) || (lj == lam && j < c);
        }
        // lane rank(c) receives column c's population (forward permute; lanes c < N)
        {
            const int dst = (lane < N ? rank : lane) * 4;  // byte address of the receiving lane
            const int lo = __builtin_amdgcn_ds_permute(dst, __double2loint(p));
            const int hi = __builtin_amdgcn_ds_permute(dst, __double2hiint(p));
            if (lane < N) pop += __hiloint2double(hi, lo) * wk;
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            racc[2 * r] = R.re[r] * wk;
            racc[2 * r + 1] = R.im[r] * wk;
        }
    }
    // Re Tr(rho J_a) = Re sum_ij conj(rho_ij) J_ij (rho Hermitian; J need not be) and
    // Re Tr(M J_a) = sum_c p_c Re (U^dag J_a U)_cc the same way, elementwise on J's tile
    auto current = [&](int ax, const Tile &J) {
        double tot = 0.0, intra = 0.0;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            tot += R.re[r] * J.re[r] + R.im[r] * J.im[r];
            intra += M.re[r] * J.re[r] + M.im[r] * J.im[r];
        }
        tot = wave_sum(tot);
        intra = a.inter_intra ? wave_sum(intra) : 0.0;
        cur[ax] += (tot - intra) * wk;
        cur[3 + ax] += intra * wk;
    };
    fused_wait();  // [J_x J_y] in buffer 0; buffer 1 read (H(t_n))
    stage_async<W>(a, step_lds, 1, f.t2_j, 2, f.t2_prev, 0, col, w, lane);
    if (active) {
        current(0, form_tile(fused_slot<W>(step_lds, 0, 0), e1, M1, N));
        current(1, form_tile(fused_slot<W>(step_lds, 0, 1), e1, M1, N));
    }
    // ---- stage 1 at t_n (H~ of t_n), with J_z
    fused_wait();  // [J_z H~] in buffer 1; buffer 0 read
    stage_async<W>(a, step_lds, 0, f.t2_half, 0, f.t2_half, 1, col, w, lane);
    if (active) {
        current(2, form_tile(fused_slot<W>(step_lds, 1, 0), e1, M1, N));
        Ht = form_tile(fused_slot<W>(step_lds, 1, 1), e1, M1, N);  // H~ (Hermitian: its conjugate is H~^T's tile)
        X = stage_derivative_m<true>(a, R, Ht, M, buf);
        axpy_tile(Acc, X, dt / 6.0, R);
        axpy_tile(Y, X, 0.5 * dt, R);
    }
    // ---- the block's sample sums in buffer 1 (read): one partial row, waves added in order
    __syncthreads();
    {
        double *red = reinterpret_cast<double *>(fused_slot<W>(step_lds, 1, 0));  // [W][nv]
        double *mine = red + w * a.nv;
        if (lane < 6) {
            double v = 0.0;
#pragma unroll
            for (int q = 0; q < 6; ++q)
                if (lane == q) v = cur[q];
            mine[lane] = v;
        }
        if (lane < N) mine[6 + lane] = pop;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = g + 4 * r;
            if (i < N && c < N) {
                mine[6 + N + 2 * (i * N + c)] = racc[2 * r];
                mine[6 + N + 2 * (i * N + c) + 1] = racc[2 * r + 1];
            }
        }
        __syncthreads();
        double *row = partials + (f.step_local * nb + b) * static_cast<int64_t>(a.nv);
        for (int v = threadIdx.x; v < a.nv; v += blockDim.x) {
            double sum = 0.0;
            for (int q = 0; q < W; ++q) sum += red[q * a.nv + v];
            row[v] = sum;
        }
    }
    // ---- stages 2 and 3 at t_n + dt/2
    fused_wait();  // [H~ H at t + dt/2] in buffer 0; buffer 1 read (sample sums)
    stage_async<W>(a, step_lds, 1, f.t2_end, 0, f.t2_end, 1, col, w, lane);
    if (active) {
        Ht = form_tile(fused_slot<W>(step_lds, 0, 0), e1, M1, N);
        // within the step only: the last round at max|K| < 1e-6 and no final Newton-Schulz step
        if (a.deph)
            eig_refine(form_tile(fused_slot<W>(step_lds, 0, 1), e1, M1, N), U, Ut, N, buf, kFusedSmallK, false, true);
        Tile Mh{};
        double dd;
        if (a.deph) Mh = band_diag(Y, U, Ut, dd);
        X = stage_derivative_m<true>(a, Y, Ht, Mh, buf);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            Acc.re[r] = fma(dt / 3.0, X.re[r], Acc.re[r]);
            Acc.im[r] = fma(dt / 3.0, X.im[r], Acc.im[r]);
        }
        axpy_tile(Y, X, 0.5 * dt, R);
        if (a.deph) Mh = band_diag(Y, U, Ut, dd);
        X = stage_derivative_m<true>(a, Y, Ht, Mh, buf);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            Acc.re[r] = fma(dt / 3.0, X.re[r], Acc.re[r]);
            Acc.im[r] = fma(dt / 3.0, X.im[r], Acc.im[r]);
        }
        axpy_tile(Y, X, dt, R);
    }
    // ---- stage 4 at t_n + dt; eigenvectors and energies for the next sample
    fused_wait();  // [H~ H at t + dt] in buffer 1
    if (active) {
        Ht = form_tile(fused_slot<W>(step_lds, 1, 0), e1, M1, N);
        const Tile H = form_tile(fused_slot<W>(step_lds, 1, 1), e1, M1, N);
        eig_refine(H, U, Ut, N, buf, kFusedSmallK, true, true, &lam);
        Tile Me{};
        double dd;
        if (a.deph) Me = band_diag(Y, U, Ut, dd);
        X = stage_derivative_m<true>(a, Y, Ht, Me, buf);
        axpy_tile(R, X, dt / 6.0, Acc);
        store_tile(rho + k * nn, R, N);
        store_tile(Ust + k * nn, U, N);
        if (lane < 16) f.lam[k * 16 + c] = lam;
    }
}


// ------------------------------------- velocity gauge, time-fused ----
// Steps [s0, s1) of the VelocityHamiltonian gauge (SBEs.h:787-805,862-871: the
// band-basis H = sum_a A_a(t) P_a + diag(E)) with the classical RK4 tableau in one
// launch: a wave owns one k point for every step, with pMatrix, E, the occupations
// and rho in registers, so only the per-step partial rows reach HBM. A block's
// kNbFusedWaves k points add their sample sums in LDS each step into one partial
// row (waves in order). A at the stage times comes from a table of the whole pulse
// ([nt][4][3], the host's values), read with uniform loads.
template <int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(2))) void nb16_vel_fused_kernel(
    const NbArgs a, const NbVelFused f, const cd *__restrict__ P, const double *__restrict__ Eocc,
    cd *__restrict__ rho, double *__restrict__ partials)
{
    extern __shared__ cd vel_lds[];
    const int N = a.N, nn = N * N;
    const int c = lane_c(), g = lane_g(), lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
    const int64_t k = static_cast<int64_t>(blockIdx.x) * W + w;
    const bool active = k < a.nk;
    const int64_t ks = a.nk_stride * nn;
    cd *buf = vel_lds + w * 16 * kTpad;
    double *red = reinterpret_cast<double *>(vel_lds + W * 16 * kTpad);  // [W][nv]
    Tile R, Px, Py, Pz;
    double Ec = 0.0, oc = 0.0, wk = 0.0;
    if (active) {
        R = load_tile(rho + k * nn, N);
        // pMatrix_a read transposed: not Hermitian for the model classes whose J is not
        Px = load_tile_t(P + k * nn, N);
        Py = load_tile_t(P + ks + k * nn, N);
        Pz = load_tile_t(P + 2 * ks + k * nn, N);
        Ec = c < N ? Eocc[k * 2 * N + c] : 0.0;      // E of band c
        oc = c < N ? Eocc[k * 2 * N + N + c] : 0.0;  // occupation of band c
        wk = a.weight[k];
    }
    const double dt = f.dt;
    for (int64_t step = f.s0; step < f.s1; ++step) {
        // ---- sample at t_n: band populations, currents Tr(rho P_a) split into the
        // diagonal (intraband) and off-diagonal parts, rho sums
        double cur[6] = {0, 0, 0, 0, 0, 0};
        double pop = 0.0;
        if (active) {
            double lr[4], lc;
            diag_bcast(R, lr, lc);
            pop = lc * wk;
            for (int ax = 0; ax < 3; ++ax) {
                const Tile &Pa = ax == 0 ? Px : (ax == 1 ? Py : Pz);
                double inter = 0.0, intra = 0.0;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const double val = R.re[r] * Pa.re[r] - R.im[r] * Pa.im[r];  // Re(rho_ij P_ji)
                    if (g + 4 * r == c && a.inter_intra) intra += val;
                    else inter += val;
                }
                cur[ax] = wave_sum(inter) * wk;
                cur[3 + ax] = wave_sum(intra) * wk;
            }
        }
        double *mine = red + w * a.nv;
        if (lane < 6) {
            double v = 0.0;
#pragma unroll
            for (int q = 0; q < 6; ++q)
                if (lane == q) v = cur[q];
            mine[lane] = v;
        }
        if (lane < N) mine[6 + lane] = pop;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = g + 4 * r;
            if (i < N && c < N) {
                mine[6 + N + 2 * (i * N + c)] = active ? R.re[r] * wk : 0.0;
                mine[6 + N + 2 * (i * N + c) + 1] = active ? R.im[r] * wk : 0.0;
            }
        }
        __syncthreads();
        double *row = partials + ((step - f.s0) * f.nblocks + blockIdx.x) * static_cast<int64_t>(a.nv);
        for (int v = threadIdx.x; v < a.nv; v += blockDim.x) {
            double sum = 0.0;
            for (int q = 0; q < W; ++q) sum += red[q * a.nv + v];
            row[v] = sum;
        }
        __syncthreads();  // the rows are read before the next step's sums
        // ---- the four RK4 stages
        if (active) {
            const double *At = f.A + step * 12;
            Tile Acc, Y = R, X;
#pragma unroll
            for (int st = 0; st < 4; ++st) {
                const double A0 = At[3 * st], A1 = At[3 * st + 1], A2 = At[3 * st + 2];
                Tile M;  // H^T = sum_a A_a P_a^T + diag(E)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    M.re[r] = A0 * Px.re[r] + A1 * Py.re[r] + A2 * Pz.re[r] + (g + 4 * r == c ? Ec : 0.0);
                    M.im[r] = A0 * Px.im[r] + A1 * Py.im[r] + A2 * Pz.im[r];
                }
                const Tile T = cmm(M, false, Y);  // H y
                const Tile Tt = transpose(T, buf);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const double dr = T.re[r] - Tt.re[r], di = T.im[r] + Tt.im[r];
                    double xr = di, xi = -dr;  // -i (H y - (H y)^dag)
                    if (a.deph) {
                        const bool dg = g + 4 * r == c;
                        const double gm = dg ? a.gam1 : a.gam2;
                        xr -= gm * (Y.re[r] - (dg ? oc : 0.0));
                        xi -= gm * Y.im[r];
                    }
                    X.re[r] = xr;
                    X.im[r] = xi;
                }
                if (st == 0) {
                    axpy_tile(Acc, X, dt / 6.0, R);
                    axpy_tile(Y, X, 0.5 * dt, R);
                } else if (st < 3) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        Acc.re[r] = fma(dt / 3.0, X.re[r], Acc.re[r]);
                        Acc.im[r] = fma(dt / 3.0, X.im[r], Acc.im[r]);
                    }
                    axpy_tile(Y, X, st == 1 ? 0.5 * dt : dt, R);
                } else {
                    axpy_tile(R, X, dt / 6.0, Acc);
                }
            }
        }
    }
    if (active) store_tile(rho + k * nn, R, N);
}

}  // namespace

hipError_t nb16_launch_vel_step(const NbArgs &a, const NbVelStep &v, const double2 *P, const double *Eocc,
                                double2 *rho, double *partials, int64_t step_local, hipStream_t s)
{
    if (a.N > 16) return hipErrorInvalidValue;
    const dim3 grid(static_cast<unsigned>(a.sample_waves));
    if (v.nstage == 4) nb16_vel_step_kernel<4><<<grid, 64, 0, s>>>(a, v, P, Eocc, rho, partials, step_local);
    else if (v.nstage == 6) nb16_vel_step_kernel<6><<<grid, 64, 0, s>>>(a, v, P, Eocc, rho, partials, step_local);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t nb16_launch_rhs(const NbArgs &a, const NbStage &sg, const double2 *mats, double2 *rho, double2 *kbuf,
                           double2 *U, hipStream_t s)
{
    if (a.N > 16) return hipErrorInvalidValue;
    nb16_rhs_kernel<false><<<dim3(static_cast<unsigned>(a.nk)), 64, 0, s>>>(a, sg, sg, mats, rho, kbuf, U);
    return hipGetLastError();
}

hipError_t nb16_launch_rhs_pair(const NbArgs &a, const NbStage &sg, const NbStage &sg2, const double2 *mats,
                                double2 *rho, double2 *kbuf, double2 *U, hipStream_t s)
{
    if (a.N > 16) return hipErrorInvalidValue;
    nb16_rhs_kernel<true><<<dim3(static_cast<unsigned>(a.nk)), 64, 0, s>>>(a, sg, sg2, mats, rho, kbuf, U);
    return hipGetLastError();
}

hipError_t nb16_launch_sample(const NbArgs &a, const double2 *mats, const double2 *rho, double *partials,
                              int64_t step_local, double2 *U, int ureuse, hipStream_t s)
{
    if (a.N > 16) return hipErrorInvalidValue;
    nb16_sample_kernel<<<dim3(static_cast<unsigned>(a.sample_waves)), 64, 0, s>>>(a, mats, rho, partials, step_local,
                                                                                 U, ureuse);
    return hipGetLastError();
}

// dynamic LDS beyond the default limit: raise the kernel's limit, once per device and kernel
static hipError_t allow_lds(const void *kernel, size_t lds)
{
    if (lds <= (size_t(64) << 10)) return hipSuccess;
    static std::mutex mu;
    static std::map<std::pair<int, const void *>, size_t> raised;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lk(mu);
    size_t &have = raised[{dev, kernel}];
    if (have < lds) {
        e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        if (e != hipSuccess) return e;
        have = lds;
    }
    return hipSuccess;
}

size_t nb16_step_lds_bytes(int N, int M1, int nv, int waves)
{
    (void)N;
    (void)M1;
    // the sample sums must fit in one buffer (two slots)
    if (static_cast<int64_t>(waves) * nv > 2 * kFusedSlot * 2) return ~size_t(0);
    return sizeof(double2) * static_cast<size_t>(static_cast<int64_t>(waves) * 16 * kTpad + 4 * kFusedSlot +
                                                 static_cast<int64_t>(waves) * kNbFusedMaxM1);
}

template <int W>
static hipError_t launch_step_w(const NbArgs &a, const NbFused &f, double2 *rho, double2 *U, double *partials,
                                hipStream_t s)
{
    const size_t lds = nb16_step_lds_bytes(a.N, a.boxM[0], a.nv, W);
    const hipError_t e = allow_lds(reinterpret_cast<const void *>(&nb16_step_kernel<W>), lds);
    if (e != hipSuccess) return e;
    nb16_step_kernel<W><<<dim3(static_cast<unsigned>(f.nblocks)), 64 * W, lds, s>>>(a, f, rho, U, partials);
    return hipGetLastError();
}

hipError_t nb16_launch_step(const NbArgs &a, const NbFused &f, double2 *rho, double2 *U, double *partials,
                            hipStream_t s)
{
    if (a.N > 16 || !a.box || a.boxM[0] > kNbFusedMaxM1 || f.nblocks < 1 || (f.waves != 4 && f.waves != 8))
        return hipErrorInvalidValue;
    return f.waves == 4 ? launch_step_w<4>(a, f, rho, U, partials, s) : launch_step_w<8>(a, f, rho, U, partials, s);
}

size_t nb16_vel_fused_lds_bytes(int nv, int waves)
{
    return sizeof(double2) * static_cast<size_t>(waves) * 16 * kTpad + sizeof(double) * waves * nv;
}

template <int W>
static hipError_t launch_vel_fused_w(const NbArgs &a, const NbVelFused &f, const double2 *P, const double *Eocc,
                                     double2 *rho, double *partials, hipStream_t s)
{
    const size_t lds = nb16_vel_fused_lds_bytes(a.nv, W);
    const hipError_t e = allow_lds(reinterpret_cast<const void *>(&nb16_vel_fused_kernel<W>), lds);
    if (e != hipSuccess) return e;
    nb16_vel_fused_kernel<W><<<dim3(static_cast<unsigned>(f.nblocks)), 64 * W, lds, s>>>(a, f, P, Eocc, rho, partials);
    return hipGetLastError();
}

hipError_t nb16_launch_vel_fused(const NbArgs &a, const NbVelFused &f, const double2 *P, const double *Eocc,
                                 double2 *rho, double *partials, hipStream_t s)
{
    if (a.N > 16 || f.nblocks < 1 || f.s1 < f.s0 || (f.waves != 4 && f.waves != 8)) return hipErrorInvalidValue;
    return f.waves == 4 ? launch_vel_fused_w<4>(a, f, P, Eocc, rho, partials, s)
                        : launch_vel_fused_w<8>(a, f, P, Eocc, rho, partials, s);
}

}  // namespace sbe
