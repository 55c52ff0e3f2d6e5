// nband32.hip - one whole RK4 step of the N-band (Wannier) LengthWannier
// gauge for 16 < N <= 32, every matrix as four 16x16 FP64 MFMA register tiles
// dealt over the four waves of a block.
//
// Same mathematics as the N <= 16 fused step (nband16.hip: SBEs.h:596-624
// RK step, 762-861 right-hand side, 694-760 sample) for one k point per block
// at a time (blocks walk k = block, block + G, ...; one partial row of sample
// sums per block and step). Wave w holds quadrant (p, q) = (w >> 1, w & 1) of
// every 32 x 32 matrix: rows 16p + g + 4r, column 16q + c of lane c + 16g
// (bands beyond N are zero, identity in the eigenvectors). A product
// C_pq = sum_r A_pr B_rq puts both operands into LDS (pitch 33: transposed
// reads are conflict-free), one barrier, and each wave multiplies its two
// quadrant pairs on MFMA (32 MFMAs); the exchange slots alternate between two
// pairs so the next exchange needs no second barrier. Column dot products,
// the diagonal and block maxima go through a small scalar area the same way.
// The matrices (H~, H at t_n + dt/2 and t_n + dt, H~ and J at t_n) come from
// the Fourier-sum outputs in HBM: this path is for the few-hundred-k-point
// solves where the separate per-stage kernels leave the GPU idle.
#include "nband.hpp"
#include "tile16.hpp"

namespace sbe {
namespace {

using namespace tile;

constexpr int kQP = 33;              // LDS pitch (complex) of an exchanged 32 x 32 matrix
constexpr int kQSlot = 32 * kQP;     // complex per exchange slot

// The block's exchange state: four matrix slots used in alternating pairs, two
// scalar areas alternating with them, this wave's quadrant.
struct Q {
    cd *slot;      // [4][kQSlot]
    double *sc;    // [2][128]
    int p, q, ph;  // quadrant; exchange phase (pair 0 or 1)
    int N;

    __device__ cd *s(int i) const { return slot + (2 * ph + i) * kQSlot; }
    __device__ double *scal() const { return sc + ph * 128; }
    __device__ void next() { ph ^= 1; }
};

__device__ __forceinline__ void q_put(cd *S, const Tile &X, int p, int q)
{
    const int c = lane_c(), g = lane_g();
#pragma unroll
    for (int r = 0; r < 4; ++r) S[(16 * p + g + 4 * r) * kQP + 16 * q + c] = make_double2(X.re[r], X.im[r]);
}

// quadrant (p, q) of S as a tile; trans: the tile of its transpose (element (i, c) =
// S[16p + c][16q + i]); conj: conjugated
__device__ __forceinline__ Tile q_get(const cd *S, int p, int q, bool trans, bool conj)
{
    const int c = lane_c(), g = lane_g();
    Tile T;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = g + 4 * r;
        const cd v = trans ? S[(16 * p + c) * kQP + 16 * q + i] : S[(16 * p + i) * kQP + 16 * q + c];
        T.re[r] = v.x;
        T.im[r] = conj ? -v.y : v.y;
    }
    return T;
}

// acc += A B for one 16 x 16 pair (Aop: A^T's tile, B: B's tile)
__device__ __forceinline__ void mma_acc(d4 &re, d4 &im, const Tile &Aop, const Tile &B)
{
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
        const double ar = Aop.re[kb], ai = Aop.im[kb], br = B.re[kb], bi = B.im[kb];
        re = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, re, 0, 0, 0);
        re = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, bi, re, 0, 0, 1);  // 1: A negated
        im = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bi, im, 0, 0, 0);
        im = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, br, im, 0, 0, 0);
    }
}

// this wave's quadrant of op(X) op(Y) (adjX: X^dag, adjY: Y^dag), all four waves
__device__ Tile q_prod(Q &L, const Tile &X, bool adjX, const Tile &Y, bool adjY)
{
    cd *SA = L.s(0), *SB = L.s(1);
    q_put(SA, X, L.p, L.q);
    q_put(SB, Y, L.p, L.q);
    __syncthreads();
    d4 re = d4{0.0, 0.0, 0.0, 0.0}, im = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        // A_pr^T's tile: A = X: transposed read of X_pr; A = X^dag: A_pr^T = conj(X_rp)
        const Tile Aop = adjX ? q_get(SA, r, L.p, false, true) : q_get(SA, L.p, r, true, false);
        // B_rq's tile: B = Y: Y_rq (this wave's own Y for r = p); B = Y^dag: element (i, c) =
        // conj(Y[16q + c][16r + i])
        const Tile Bop = adjY ? q_get(SB, L.q, r, true, true) : (r == L.p ? Y : q_get(SB, r, L.q, false, false));
        mma_acc(re, im, Aop, Bop);
    }
    L.next();
    Tile C;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        C.re[r] = re[r];
        C.im[r] = im[r];
    }
    return C;
}

// this wave's quadrant of X^dag
__device__ Tile q_adjoint(Q &L, const Tile &X)
{
    cd *S = L.s(0);
    q_put(S, X, L.p, L.q);
    __syncthreads();
    const Tile T = q_get(S, L.q, L.p, true, true);
    L.next();
    return T;
}

// sum over the 32 rows of conj(X_ic) Y_ic for this lane's column 16q + c
__device__ double q_col_dot(Q &L, const Tile &X, const Tile &Y)
{
    const double part = col_dot(X, Y);  // this wave's 16 rows
    double *S = L.scal();
    if (lane_g() == 0) S[32 * L.p + 16 * L.q + lane_c()] = part;
    __syncthreads();
    const double d = S[16 * L.q + lane_c()] + S[32 + 16 * L.q + lane_c()];  // rows 0-15, then 16-31
    L.next();
    return d;
}

// the diagonal of a block matrix: this lane's rows (lam_r) and column (lam_c)
__device__ void q_diag(Q &L, const Tile &W, double lam_r[4], double &lam_c)
{
    const int c = lane_c(), g = lane_g();
    double *S = L.scal();
    if (L.p == L.q) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
            if (g + 4 * r == c) S[16 * L.p + c] = W.re[r];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; ++r) lam_r[r] = S[16 * L.p + g + 4 * r];
    lam_c = S[16 * L.q + c];
    L.next();
}

// max over the block (NaN propagates)
__device__ double q_max(Q &L, double v)
{
    v = wave_max(v);
    double *S = L.scal();
    if ((threadIdx.x & 63) == 0) S[64 + (threadIdx.x >> 6)] = v;
    __syncthreads();
    double m = S[64];
    for (int i = 1; i < 4; ++i) m = fmax(m, S[64 + i]);
    L.next();
    return m;
}

// sum over the block's four waves of a wave-uniform value, in wave order
__device__ double q_sum(Q &L, double v)
{
    double *S = L.scal();
    if ((threadIdx.x & 63) == 0) S[64 + (threadIdx.x >> 6)] = v;
    __syncthreads();
    const double t = ((S[64] + S[65]) + S[66]) + S[67];
    L.next();
    return t;
}

__device__ __forceinline__ bool q_valid(const Q &L, int r)
{
    return 16 * L.p + lane_g() + 4 * r < L.N && 16 * L.q + lane_c() < L.N;
}

__device__ __forceinline__ bool q_diagonal(const Q &L, int r) { return L.p == L.q && lane_g() + 4 * r == lane_c(); }

// this wave's quadrant of an N x N row-major matrix (zero padding; pad_one: identity padding)
__device__ Tile q_load(const Q &L, const cd *M, bool pad_one)
{
    const int c = lane_c(), g = lane_g(), N = L.N;
    Tile T;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = 16 * L.p + g + 4 * r, j = 16 * L.q + c;
        cd v = make_double2(pad_one && i == j ? 1.0 : 0.0, 0.0);
        if (i < N && j < N) v = M[i * N + j];
        T.re[r] = v.x;
        T.im[r] = v.y;
    }
    return T;
}

__device__ void q_store(const Q &L, cd *M, const Tile &T)
{
    const int c = lane_c(), g = lane_g(), N = L.N;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = 16 * L.p + g + 4 * r, j = 16 * L.q + c;
        if (i < N && j < N) M[i * N + j] = make_double2(T.re[r], T.im[r]);
    }
}

// U <- U (3I - U^dag U) / 2; returns max |U^dag U - I| before the step
__device__ double q_orthonormalize(Q &L, Tile &U)
{
    const Tile G = q_prod(L, U, true, U, false);
    Tile S;
    double e = 0.0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const bool dg = q_diagonal(L, r);
        const double dr = G.re[r] - (dg ? 1.0 : 0.0);
        e = fmax(e, fmax(fabs(dr), fabs(G.im[r])));
        S.re[r] = (dg ? 1.5 : 0.0) - 0.5 * G.re[r];
        S.im[r] = -0.5 * G.im[r];
    }
    U = q_prod(L, U, false, S, false);
    return q_max(L, e);
}

// Warm-started eigenvectors of H = Hk (Hermitian to rounding, FP64 sums; read again where
// needed rather than held in registers), as the N <= 16
// refinement: simultaneous Jacobi rotations written as products, the last round
// U (I + K) once max|K| < small_tol, then (ortho) a Newton-Schulz step; lam_c: this
// lane's column of diag(U^dag H U) before the last rotation.
__device__ void q_refine(Q &L, const cd *Hk, Tile &U, double small_tol, bool ortho, double &lam_out)
{
    const int N = L.N;
    Tile W = q_prod(L, U, true, q_prod(L, q_load(L, Hk, false), false, U, false), false);  // U^dag (H U)
    double lam_last = 0.0;
    for (int it = 0; it < 30; ++it) {
        double lam_r[4], lam_c;
        q_diag(L, W, lam_r, lam_c);
        lam_last = lam_c;
        Tile K;
        double kmax = 0.0;
        const int cg = 16 * L.q + lane_c();
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = 16 * L.p + lane_g() + 4 * r;
            K.re[r] = 0.0;
            K.im[r] = 0.0;
            const double w2 = W.re[r] * W.re[r] + W.im[r] * W.im[r];
            if (i != cg && i < N && cg < N && w2 > 0.0) {
                const double D = lam_c - lam_r[r];
                const double s = (D > 0.0 || (D == 0.0 && i < cg)) ? 1.0 : -1.0;
                const double inv = 1.0 / (0.5 * D + s * sqrt(0.25 * D * D + w2));
                K.re[r] = W.re[r] * inv;
                K.im[r] = W.im[r] * inv;
                kmax = fmax(kmax, K.re[r] * K.re[r] + K.im[r] * K.im[r]);  // |K|^2 here, the root once below
            }
        }
        kmax = sqrt(q_max(L, kmax));
        if (kmax == 0.0) break;
        if (kmax > 0.5) {  // far from converged: damped rotations
            const double fct = 0.5 / kmax;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                K.re[r] *= fct;
                K.im[r] *= fct;
            }
        }
        Tile V;
        const bool small = kmax < small_tol;
        if (small) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                V.re[r] = K.re[r] + (q_diagonal(L, r) ? 1.0 : 0.0);
                V.im[r] = K.im[r];
            }
        } else {  // I + K + K^2/2
            const Tile K2 = q_prod(L, K, false, K, false);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                V.re[r] = K.re[r] + 0.5 * K2.re[r] + (q_diagonal(L, r) ? 1.0 : 0.0);
                V.im[r] = K.im[r] + 0.5 * K2.im[r];
            }
        }
        U = q_prod(L, U, false, V, false);
        if (small) break;
        if (kmax > 1e-3) {  // V far from unitary: restore U fully, then W from H itself
            for (int n = 0; n < 8; ++n)
                if (q_orthonormalize(L, U) < 1e-8) break;
            W = q_prod(L, U, true, q_prod(L, q_load(L, Hk, false), false, U, false), false);
        } else {
            W = q_prod(L, V, true, q_prod(L, W, false, V, false), false);  // V^dag (W V)
        }
    }
    if (ortho) q_orthonormalize(L, U);
    lam_out = lam_last;
}

// y's band-diagonal part U diag(U^dag y U) U^dag and the diagonal d for this lane's column
__device__ Tile q_band_diag(Q &L, const Tile &Y, const Tile &U, double &d)
{
    const Tile G = q_prod(L, Y, true, U, false);  // y U (y Hermitian: y = y^dag)
    d = q_col_dot(L, U, G);
    Tile Ud;  // U diag(d): column 16q + c scaled by its own d
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        Ud.re[r] = U.re[r] * d;
        Ud.im[r] = U.im[r] * d;
    }
    return q_prod(L, Ud, false, U, true);
}

// -i(H~ y - (H~ y)^dag) - g2 y - (g1 - g2) M (H~ Hermitian)
__device__ Tile q_stage(Q &L, const NbArgs &a, const Tile &Y, const Tile &Ht, const Tile &M)
{
    const Tile T = q_prod(L, Ht, true, Y, false);  // H~ y
    const Tile Td = q_adjoint(L, T);
    Tile X;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const double dr = T.re[r] - Td.re[r], di = T.im[r] - Td.im[r];
        X.re[r] = di;
        X.im[r] = -dr;
        if (a.deph) {
            X.re[r] -= a.gam2 * Y.re[r] + (a.gam1 - a.gam2) * M.re[r];
            X.im[r] -= a.gam2 * Y.im[r] + (a.gam1 - a.gam2) * M.im[r];
        }
    }
    return X;
}

__device__ __forceinline__ void q_axpy(Tile &Y, const Tile &X, double s, const Tile &B)
{
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        Y.re[r] = fma(s, X.re[r], B.re[r]);
        Y.im[r] = fma(s, X.im[r], B.im[r]);
    }
}

constexpr int kQStage = 1024;  // complex per staged matrix (N <= 32)

// the k point's matrices src0 and (if given) src1 into staging buffer b, asynchronously:
// chunks of 64 complex dealt over the four waves, one 16-byte global-to-LDS load per
// lane; their completion is counted by vmcnt (stage_wait)
__device__ __forceinline__ void stage_async(cd *stg, int b, const cd *src0, const cd *src1, int nn, int w, int lane)
{
    const int Q = (nn + 63) >> 6;
    const int total = src1 ? 2 * Q : Q;
    for (int ch = w; ch < total; ch += 4) {  // uniform over the wave
        const int m = ch >= Q ? 1 : 0, q = ch - m * Q;
        const cd *src = m ? src1 : src0;
        const int ij = q * 64 + lane;
        if (ij < nn) __builtin_amdgcn_global_load_lds(src + ij, stg + (2 * b + m) * kQStage + q * 64, 16, 0, 0);
    }
}

// every wave's staging (and other) loads have landed and every wave is past its earlier reads
__device__ __forceinline__ void stage_wait()
{
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1))) void nb32_step_kernel(
    const NbArgs a, const NbStep32 f, cd *__restrict__ rho, cd *__restrict__ Ust, double *__restrict__ partials)
{
    __shared__ cd slots[4 * kQSlot];
    __shared__ cd stg[4 * kQStage];     // staging: buffers A (0, 1) and B (2, 3)
    __shared__ double rsum[2 * kQStage];  // the block's sum_k w rho (each element owned by one lane)
    __shared__ double scal[2 * 128];
    const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
    Q L{slots, scal, w >> 1, w & 1, 0, a.N};
    const int N = a.N, nn = N * N;
    const int c = lane_c(), g = lane_g(), lane = threadIdx.x & 63;
    const int64_t slot = a.nk_stride * nn;
    const double dt = f.dt;
    const bool need_h = !f.u_valid || !f.lam_valid;
    const cd *A0 = stg, *A1 = stg + kQStage, *B0 = stg + 2 * kQStage, *B1 = stg + 3 * kQStage;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int e = (16 * L.p + g + 4 * r) * N + 16 * L.q + c;
        if (q_valid(L, r)) rsum[2 * e] = rsum[2 * e + 1] = 0.0;
    }
    double curl = 0.0;  // wave 0, lane a < 6: current component a (inter x, y, z, intra x, y, z)
    double pop = 0.0;   // wave 0, lane m < N: population of the m-th lowest band
    // staging: A = [J_x, H(t_n) if needed], B = [J_y, J_z], A = [H~(t_n), H(t + dt/2) if
    // dephasing], B = [H~(t + dt/2), H(t + dt)], A = [H~(t + dt)]; each buffer refilled
    // after the wait that follows its last reads, while the other one is in use
    for (int64_t k = blockIdx.x; k < a.nk; k += gridDim.x) {
        const int64_t kn = k * nn;
        const double wk = a.weight[k];
        stage_async(stg, 0, f.j + kn, need_h ? f.h_prev + kn : nullptr, nn, w, lane);
        Tile R = q_load(L, rho + kn, false);
        Tile U = q_load(L, Ust + kn, true);
        stage_wait();
        stage_async(stg, 1, f.j + slot + kn, f.j + 2 * slot + kn, nn, w, lane);
        double lam;  // energy of column 16q + c
        if (need_h) {
            if (!f.u_valid) q_refine(L, A1, U, 1e-8, true, lam);
            const Tile HU = q_prod(L, q_load(L, A1, false), false, U, false);
            lam = q_col_dot(L, U, HU);
        } else {
            lam = f.lam[k * 32 + 16 * L.q + c];
        }
        // rho's band-diagonal part: the populations (p for column 16q + c), the intraband
        // currents Tr(M J) and stage 1's dephasing
        double pcol;
        const Tile M = q_band_diag(L, R, U, pcol);
        // band ranks: all 32 energies and populations through the scalar area
        {
            double *S = L.scal();
            if (g == 0 && L.p == 0) {
                S[16 * L.q + c] = lam;
                S[32 + 16 * L.q + c] = pcol;
            }
            __syncthreads();
            if (w == 0 && lane < N) {
                const double lm = S[lane];
                int rank = 0;
                for (int j = 0; j < N; ++j) {
                    const double lj = S[j];
                    rank += (lj < lm) || (lj == lm && j < lane);
                }
                S[64 + rank] = S[32 + lane];  // population of the band with this rank
            }
            __syncthreads();
            if (w == 0 && lane < N) pop += S[64 + lane] * wk;
            L.next();
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int e = (16 * L.p + g + 4 * r) * N + 16 * L.q + c;
            if (q_valid(L, r)) {
                rsum[2 * e] = fma(R.re[r], wk, rsum[2 * e]);
                rsum[2 * e + 1] = fma(R.im[r], wk, rsum[2 * e + 1]);
            }
        }
        // currents: Re Tr(rho J) = Re sum conj(rho_ij) J_ij, Re Tr(M J) likewise (rho, M
        // Hermitian): the waves' sums of the three axes, then one exchange for all six
        double part[6];
        auto current = [&](int ax, const cd *Js) {
            const Tile J = q_load(L, Js, false);
            double tot = 0.0, intra = 0.0;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                tot += R.re[r] * J.re[r] + R.im[r] * J.im[r];
                intra += M.re[r] * J.re[r] + M.im[r] * J.im[r];
            }
            part[ax] = wave_sum(tot);
            part[3 + ax] = a.inter_intra ? wave_sum(intra) : 0.0;
        };
        current(0, A0);
        stage_wait();  // B = [J_y, J_z]; A free
        stage_async(stg, 0, f.ht_prev + kn, a.deph ? f.h_half + kn : nullptr, nn, w, lane);
        current(1, B0);
        current(2, B1);
        {
            double *S = L.scal();
            if (lane == 0)
#pragma unroll
                for (int j = 0; j < 6; ++j) S[8 * w + j] = part[j];
            __syncthreads();
            if (w == 0 && lane < 3) {  // waves added in order
                const double tot = ((S[lane] + S[8 + lane]) + S[16 + lane]) + S[24 + lane];
                const double intra = ((S[3 + lane] + S[11 + lane]) + S[19 + lane]) + S[27 + lane];
                curl += (tot - intra) * wk;  // lane a: inter a
            }
            if (w == 0 && lane >= 3 && lane < 6) {
                const double intra = ((S[lane] + S[8 + lane]) + S[16 + lane]) + S[24 + lane];
                curl += intra * wk;  // lane 3 + a: intra a
            }
            L.next();
        }
        stage_wait();  // A = [H~(t_n), H(t + dt/2)]; B free
        stage_async(stg, 1, f.ht_half + kn, f.h_end + kn, nn, w, lane);
        // stage 1 at t_n
        Tile X = q_stage(L, a, R, q_load(L, A0, false), M);
        Tile Acc, Y;
        q_axpy(Acc, X, dt / 6.0, R);
        q_axpy(Y, X, 0.5 * dt, R);
        // stages 2 and 3 at t_n + dt/2
        if (a.deph) {
            double dd;
            q_refine(L, A1, U, 1e-6, false, dd);
        }
        stage_wait();  // B = [H~(t + dt/2), H(t + dt)]; A free
        stage_async(stg, 0, f.ht_end + kn, nullptr, nn, w, lane);
        for (int st = 1; st <= 2; ++st) {
            Tile Mh{};
            double dd;
            if (a.deph) Mh = q_band_diag(L, Y, U, dd);
            X = q_stage(L, a, Y, q_load(L, B0, false), Mh);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                Acc.re[r] = fma(dt / 3.0, X.re[r], Acc.re[r]);
                Acc.im[r] = fma(dt / 3.0, X.im[r], Acc.im[r]);
            }
            q_axpy(Y, X, st == 1 ? 0.5 * dt : dt, R);
        }
        // stage 4 at t_n + dt; eigenvectors and energies for the next sample
        q_refine(L, B1, U, 1e-6, true, lam);
        stage_wait();  // A = [H~(t + dt)]
        {
            Tile Me{};
            double dd;
            if (a.deph) Me = q_band_diag(L, Y, U, dd);
            X = q_stage(L, a, Y, q_load(L, A0, false), Me);
        }
        q_axpy(R, X, dt / 6.0, Acc);
        q_store(L, rho + kn, R);
        q_store(L, Ust + kn, U);
        if (g == 0 && L.p == 0) f.lam[k * 32 + 16 * L.q + c] = lam;
        // the next k point's first staging writes A and B: every wave is past its reads of
        // both here (the last exchange of stage 4 has a barrier after them)
    }
    // the block's partial row
    double *row = partials + (f.step_local * gridDim.x + blockIdx.x) * static_cast<int64_t>(a.nv);
    if (w == 0) {
        if (lane < 6) row[lane] = curl;
        if (lane < N) row[6 + lane] = pop;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int e = (16 * L.p + g + 4 * r) * N + 16 * L.q + c;
        if (q_valid(L, r)) {
            row[6 + N + 2 * e] = rsum[2 * e];
            row[6 + N + 2 * e + 1] = rsum[2 * e + 1];
        }
    }
}

// Steps [s0, s1) of the VelocityHamiltonian gauge (SBEs.h:787-805,862-871: the
// band-basis H = sum_a A_a(t) P_a + diag(E)) with the classical RK4 tableau in one launch,
// for 16 < N <= 32: a block of four waves owns one k point for every step, each wave one
// quadrant of rho and of pMatrix in registers (no re-reads); per step the block writes its
// partial row (the currents summed over its waves in order, populations and rho directly
// by the lanes that hold them). A at the stage times from the host-evaluated table.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1))) void nb32_vel_fused_kernel(
    const NbArgs a, const NbVelFused f, const cd *__restrict__ P, const double *__restrict__ Eocc,
    cd *__restrict__ rho, double *__restrict__ partials)
{
    __shared__ cd slots[4 * kQSlot];
    __shared__ double scal[2 * 128];
    const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
    Q L{slots, scal, w >> 1, w & 1, 0, a.N};
    const int N = a.N, nn = N * N;
    const int c = lane_c(), g = lane_g(), lane = threadIdx.x & 63;
    const int64_t k = blockIdx.x, kn = k * nn;
    const int64_t ks = a.nk_stride * nn;
    Tile R = q_load(L, rho + kn, false);
    const Tile Px = q_load(L, P + kn, false), Py = q_load(L, P + ks + kn, false), Pz = q_load(L, P + 2 * ks + kn, false);
    double Er[4], orr[4];  // E and occupation of this lane's rows (used on the diagonal)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = 16 * L.p + g + 4 * r;
        Er[r] = i < N ? Eocc[k * 2 * N + i] : 0.0;
        orr[r] = i < N ? Eocc[k * 2 * N + N + i] : 0.0;
    }
    const double wk = a.weight[k], dt = f.dt;
    for (int64_t step = f.s0; step < f.s1; ++step) {
        // ---- sample at t_n: Re Tr(rho P_a) = sum_ij Re(conj(rho_ij) P_ij) (rho Hermitian),
        // split into the diagonal (intraband) and off-diagonal parts; populations; rho
        double *row = partials + ((step - f.s0) * f.nblocks + blockIdx.x) * static_cast<int64_t>(a.nv);
        double part[6];
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            const Tile &Pa = ax == 0 ? Px : (ax == 1 ? Py : Pz);
            double inter = 0.0, intra = 0.0;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const double val = R.re[r] * Pa.re[r] + R.im[r] * Pa.im[r];
                if (q_diagonal(L, r) && a.inter_intra) intra += val;
                else inter += val;
            }
            part[ax] = wave_sum(inter);
            part[3 + ax] = wave_sum(intra);
        }
        {
            double *S = L.scal();
            if (lane == 0)
#pragma unroll
                for (int j = 0; j < 6; ++j) S[8 * w + j] = part[j];
            __syncthreads();
            if (w == 0 && lane < 6) row[lane] = (((S[lane] + S[8 + lane]) + S[16 + lane]) + S[24 + lane]) * wk;
            L.next();
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            if (!q_valid(L, r)) continue;
            const int i = 16 * L.p + g + 4 * r, j = 16 * L.q + c;
            if (i == j) row[6 + i] = R.re[r] * wk;
            row[6 + N + 2 * (i * N + j)] = R.re[r] * wk;
            row[6 + N + 2 * (i * N + j) + 1] = R.im[r] * wk;
        }
        // ---- the four RK4 stages
        const double *At = f.A + step * 12;
        Tile Acc, Y = R, X;
#pragma unroll 1
        for (int st = 0; st < 4; ++st) {
            const double A0 = At[3 * st], A1 = At[3 * st + 1], A2 = At[3 * st + 2];
            Tile H;  // this quadrant of sum_a A_a P_a + diag(E)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                H.re[r] = A0 * Px.re[r] + A1 * Py.re[r] + A2 * Pz.re[r] + (q_diagonal(L, r) ? Er[r] : 0.0);
                H.im[r] = A0 * Px.im[r] + A1 * Py.im[r] + A2 * Pz.im[r];
            }
            const Tile T = q_prod(L, H, false, Y, false);  // H y
            const Tile Td = q_adjoint(L, T);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const double dr = T.re[r] - Td.re[r], di = T.im[r] - Td.im[r];
                double xr = di, xi = -dr;  // -i (H y - (H y)^dag)
                if (a.deph) {
                    const bool dg = q_diagonal(L, r);
                    const double gm = dg ? a.gam1 : a.gam2;
                    xr -= gm * (Y.re[r] - (dg ? orr[r] : 0.0));
                    xi -= gm * Y.im[r];
                }
                X.re[r] = xr;
                X.im[r] = xi;
            }
            if (st == 0) {
                q_axpy(Acc, X, dt / 6.0, R);
                q_axpy(Y, X, 0.5 * dt, R);
            } else if (st < 3) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    Acc.re[r] = fma(dt / 3.0, X.re[r], Acc.re[r]);
                    Acc.im[r] = fma(dt / 3.0, X.im[r], Acc.im[r]);
                }
                q_axpy(Y, X, st == 1 ? 0.5 * dt : dt, R);
            } else {
                q_axpy(R, X, dt / 6.0, Acc);
            }
        }
    }
    q_store(L, rho + kn, R);
}

}  // namespace

hipError_t nb32_launch_step(const NbArgs &a, const NbStep32 &f, int nblocks, double2 *rho, double2 *U, double *partials,
                            hipStream_t s)
{
    if (a.N <= 16 || a.N > 32 || nblocks < 1) return hipErrorInvalidValue;
    nb32_step_kernel<<<dim3(static_cast<unsigned>(nblocks)), 256, 0, s>>>(a, f, rho, U, partials);
    return hipGetLastError();
}

hipError_t nb32_launch_vel_fused(const NbArgs &a, const NbVelFused &f, const double2 *P, const double *Eocc,
                                 double2 *rho, double *partials, hipStream_t s)
{
    if (a.N <= 16 || a.N > 32 || f.nblocks != a.nk || f.s1 <= f.s0) return hipErrorInvalidValue;
    nb32_vel_fused_kernel<<<dim3(static_cast<unsigned>(f.nblocks)), 256, 0, s>>>(a, f, P, Eocc, rho, partials);
    return hipGetLastError();
}

}  // namespace sbe
