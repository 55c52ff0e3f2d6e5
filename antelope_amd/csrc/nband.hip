// nband.hip - N-band (Wannier tight-binding) SBE kernels for gfx950.
//
// Reference: Materials/Wannier90.h:297-387 (H, D, J as Fourier sums over R,
// with dipoles: isDipoleZero = false, Wannier90.h:124), SBEs.h:762-861 (RHS:
// -i(H~rho - (H~rho)^dag) - U (Gamma o U^dag rho U) U^dag with
// H~ = H + E.D and U the eigenvectors of H), SBEs.h:694-760 (currents),
// SBEs.h:596-624 (RK step), SBEs.h:284-364 (initial state).
//
// Decomposition per RK stage:
//  * nb_fourier_kernel: complex GEMM C[m][k][ij] = sum_R P[k][R] B[R][m,ij]
//    on FP64 MFMA over a tile of 64 k x 64 columns, with P = e^{ik.R} (fixed, built once)
//    and B = w_R e^{iA(t).R} X_R (X = H + E.r, H, or -i R_a H), formed on
//    the fly while staging through LDS. t and E(t) come from a device table
//    indexed by (step, stage), so launch arguments do not carry time.
//  * nb_rhs_kernel / nb_sample_kernel (N > 16; nband16.hip has the N <= 16
//    forms): one wave per k point; matrices live in LDS; the eigenvectors are
//    refined from the stored ones by simultaneous rotations written as MFMA
//    products (weigh_warm). Cold starts (initial state, velocity-gauge set-up,
//    band export) use a parallel-order complex Jacobi method (N/2 disjoint
//    rotations per round).
#include <algorithm>
#include <cmath>

#include "nband.hpp"

namespace sbe {
namespace {

using cd = double2;

__device__ __forceinline__ cd cmk(double r, double i) { return make_double2(r, i); }
__device__ __forceinline__ cd cadd(cd a, cd b) { return cmk(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ cd csub(cd a, cd b) { return cmk(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ cd cmul(cd a, cd b) { return cmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ cd cconj(cd a) { return cmk(a.x, -a.y); }
__device__ __forceinline__ cd cscale(cd a, double s) { return cmk(a.x * s, a.y * s); }
// acc += a * b
__device__ __forceinline__ void cfma(cd &acc, cd a, cd b)
{
    acc.x += a.x * b.x - a.y * b.y;
    acc.y += a.x * b.y + a.y * b.x;
}
// acc += conj(a) * b

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// ------------------------------------------------------------ phase P --
__global__ void nb_phase_kernel(const NbArgs a, cd *__restrict__ P)
{
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t total = a.nk_stride * a.nr;
    if (idx >= total) return;
    const int64_t k = idx / a.nr;
    const int r = static_cast<int>(idx % a.nr);
    if (k >= a.nk) {
        P[idx] = cmk(0.0, 0.0);
        return;
    }
    const double *R = a.rcart + 3 * r;
    const double ph = a.kpts[3 * k] * R[0] + a.kpts[3 * k + 1] * R[1] + a.kpts[3 * k + 2] * R[2];
    double s, c;
    sincos(ph, &s, &c);
    P[idx] = cmk(c, s);
}

// --------------------------------------------------------- Fourier GEMM --
constexpr int GM = 64, GN = 64, GK = 32;

// B element of matrix kind `type` (row r, element ij) at field E
__device__ __forceinline__ cd b_value(const NbArgs &a, const double E[3], int type, int r, int ij, cd phase)
{
    const int nn = a.N * a.N;
    const cd h = a.ham[static_cast<int64_t>(r) * nn + ij];
    cd x;
    if (type == kMatHtilde) {
        x = h;
        const cd *p = a.pos + static_cast<int64_t>(r) * 3 * nn + ij;
        x = cadd(x, cscale(p[0], E[0]));
        x = cadd(x, cscale(p[nn], E[1]));
        x = cadd(x, cscale(p[2 * nn], E[2]));
    } else if (type == kMatH0) {
        x = h;
    } else {  // J_a = -i R_a H (Wannier90.h:373-384)
        const double Ra = a.rcart[3 * r + (type - kMatJx)];
        x = cmk(h.y * Ra, -h.x * Ra);
    }
    return cmul(cscale(phase, a.rweight[r]), x);
}

// B element (row r, column c) of the evaluation: c = mat * N^2 + ij
__device__ __forceinline__ cd b_element(const NbArgs &a, const NbEval &ev, int r, int64_t c, cd phase)
{
    const int nn = a.N * a.N;
    const int mat = static_cast<int>(c / nn);
    return b_value(a, ev.E, ev.mats[mat], r, static_cast<int>(c % nn), phase);
}

typedef double d4 __attribute__((ext_vector_type(4)));

// B of one evaluation, formed once (instead of once per k tile):
// b64[r][c] = w_R e^{iA(t).R} X_R[c] for r < nr, c < ncols; zero padding to
// nr_pad rows and bcols columns.
__global__ void nb_bform_kernel(const NbArgs a, const NbEval ev, cd *__restrict__ b64)
{
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= static_cast<int64_t>(a.nr_pad) * a.bcols) return;
    const int r = static_cast<int>(idx / a.bcols);
    const int64_t c = idx % a.bcols;
    const int64_t ncols = static_cast<int64_t>(ev.nmat) * a.N * a.N;
    cd v = cmk(0.0, 0.0);
    if (r < a.nr && c < ncols) {
        const double *R = a.rcart + 3 * r;
        double sn, cs;
        sincos(ev.A[0] * R[0] + ev.A[1] * R[1] + ev.A[2] * R[2], &sn, &cs);
        v = b_element(a, ev, r, c, cmk(cs, sn));
    }
    b64[idx] = v;
}

// 64 k x 64 column tile per block of 4 waves; wave w owns the 32 x 32
// sub-tile (w / 2, w % 2) as 2 x 2 MFMA tiles of 16 x 16. K (lattice
// vectors) in slices of 32 staged through LDS: e^{ik.R} from P (rows of 32
// consecutive R per k, row pitch padded against bank conflicts) and B from
// the evaluation's b64. The complex product is four real
// v_mfma_f64_16x16x4f64 per tile and K step:
//   re += a_re b_re - a_im b_im,  im += a_re b_im + a_im b_re.
// MFMA operand layout (16x16x4, FP64): A row = lane % 16, k = lane / 16;
// B k = lane / 16, column = lane % 16; D element r of lane l is
// row (l / 16) + 4 r, column l % 16.
__global__ __launch_bounds__(256) void nb_fourier_kernel(const NbArgs a, const NbEval ev, const cd *__restrict__ P,
                                                         const cd *__restrict__ b64, cd *__restrict__ out)
{
    __shared__ cd As[GK][GM + 1];
    __shared__ cd Bs[GK][GN];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int64_t k0 = static_cast<int64_t>(blockIdx.x) * GM;
    const int64_t c0 = static_cast<int64_t>(blockIdx.y) * GN;
    const int nn = a.N * a.N;
    const int64_t ncols = static_cast<int64_t>(ev.nmat) * nn;
    d4 accr[2][2], acci[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            accr[i][j] = d4{0.0, 0.0, 0.0, 0.0};
            acci[i][j] = d4{0.0, 0.0, 0.0, 0.0};
        }

    for (int r0 = 0; r0 < a.nr_pad; r0 += GK) {
        // A tile: P[k0 + m][r0 + kk]
        for (int t = threadIdx.x; t < GM * GK; t += 256) {
            const int m = t / GK, kk = t % GK;
            const int64_t k = k0 + m;
            const int r = r0 + kk;
            As[kk][m] = (k < a.nk_stride && r < a.nr) ? P[k * a.nr + r] : cmk(0.0, 0.0);
        }
        // B tile: b64[r0 + kk][c0 + n]
        for (int t = threadIdx.x; t < GK * GN; t += 256) {
            const int kk = t / GN, n = t % GN;
            Bs[kk][n] = b64[static_cast<int64_t>(r0 + kk) * a.bcols + c0 + n];
        }
        __syncthreads();
#pragma unroll
        for (int kb = 0; kb < GK; kb += 4) {
            const int kk = kb + (lane >> 4);
            cd av[2], bv[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) av[i] = As[kk][32 * wm + 16 * i + (lane & 15)];
#pragma unroll
            for (int j = 0; j < 2; ++j) bv[j] = Bs[kk][32 * wn + 16 * j + (lane & 15)];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    accr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i].x, bv[j].x, accr[i][j], 0, 0, 0);
                    accr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i].y, bv[j].y, accr[i][j], 0, 0, 1);  // 1: A negated
                    acci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i].x, bv[j].y, acci[i][j], 0, 0, 0);
                    acci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i].y, bv[j].x, acci[i][j], 0, 0, 0);
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int64_t k = k0 + 32 * wm + 16 * i + (lane >> 4) + 4 * q;
                const int64_t c = c0 + 32 * wn + 16 * j + (lane & 15);
                if (k >= a.nk_stride || c >= ncols) continue;
                const int mat = static_cast<int>(c / nn), ij = static_cast<int>(c % nn);
                out[(static_cast<int64_t>(mat) * a.nk_stride + k) * nn + ij] = cmk(accr[i][j][q], acci[i][j][q]);
            }
}

// ------------------------------------------ separable Fourier sums (box) --
// On a product k grid the sums of Wannier90.h:297-387 factor per lattice axis:
// with R = n1 a1 + n2 a2 + n3 a3 and theta_i = k.a_i, e^{ik.R} = prod_i
// e^{i n_i theta_i}, so X(k) = sum_R B_R e^{ik.R} is three passes of short sums
// along one axis each, for every column c = mat N^2 + ij:
//   T1[n1][n2][k3] = sum_n3 B[n1][n2][n3] E3[k3][n3]
//   T2[n1][k2][k3] = sum_n2 T1[n1][n2][k3] E2[k2][n2]
//   X[k1][k2][k3]  = sum_n1 T2[n1][k2][k3] E1[k1][n1]
// 2.7 M complex multiply-adds per column for the dense sum at 18,432 k points
// and 147 R become 0.15 M.

// B on the box: bbox[cell][c] = w_R e^{iA.R} X_R[c] for the R at that cell, 0 elsewhere
__global__ void nb_box_bform_kernel(const NbArgs a, const NbEval ev, int64_t ncells, int64_t ncols)
{
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= ncells * ncols) return;
    const int64_t cell = idx / ncols, c = idx % ncols;
    const int r = a.box_r[cell];
    cd v = cmk(0.0, 0.0);
    if (r >= 0) {
        const double *R = a.rcart + 3 * r;
        double sn, cs;
        sincos(ev.A[0] * R[0] + ev.A[1] * R[1] + ev.A[2] * R[2], &sn, &cs);
        v = b_element(a, ev, r, c, cmk(cs, sn));
    }
    a.box_b[cell * a.bcols + c] = v;
}

// One pass: Y[o][k][i][c] = sum_{m < M} E[k][m] X[o][m][i][c] for k < K; block x
// = (o, i), block y = 256 columns, block z = kc consecutive k. The M inputs of a
// thread stay in registers (MM >= M); E is read with uniform loads. SCATTER (the
// last pass, o = 0): the result goes to out[mat][kinv[k][i]][ij] instead of Y.
template <int MM, bool SCATTER>
__device__ __forceinline__ void box_pass_body(const cd *__restrict__ X, cd *__restrict__ Y, const cd *__restrict__ E,
                                              int M, int K, int kc, int64_t inner, int64_t ncols, int64_t pitch,
                                              const int32_t *__restrict__ kinv, cd *__restrict__ out,
                                              int64_t nk_stride, int nn, int kz)
{
    const int64_t oi = blockIdx.x;
    const int64_t o = oi / inner, i = oi - o * inner;
    const int64_t c = static_cast<int64_t>(blockIdx.y) * 256 + threadIdx.x;
    if (c >= ncols) return;  // no barriers below
    cd x[MM];
#pragma unroll
    for (int m = 0; m < MM; ++m) x[m] = m < M ? X[((o * M + m) * inner + i) * pitch + c] : cmk(0.0, 0.0);
    const int mat = static_cast<int>(c / nn), ij = static_cast<int>(c - static_cast<int64_t>(mat) * nn);
    const int k0 = kz * kc, k1 = min(K, k0 + kc);
    for (int k = k0; k < k1; ++k) {
        const cd *e = E + static_cast<int64_t>(k) * M;
        double re = 0.0, im = 0.0;
#pragma unroll
        for (int m = 0; m < MM; ++m) {
            if (m < M) {
                const cd w = e[m];
                re = fma(w.x, x[m].x, fma(-w.y, x[m].y, re));
                im = fma(w.x, x[m].y, fma(w.y, x[m].x, im));
            }
        }
        if constexpr (SCATTER) {
            const int32_t kf = kinv[static_cast<int64_t>(k) * inner + i];
            if (kf >= 0) out[(static_cast<int64_t>(mat) * nk_stride + kf) * nn + ij] = cmk(re, im);
        } else {
            Y[((o * K + k) * inner + i) * pitch + c] = cmk(re, im);
        }
    }
}

template <int MM, bool SCATTER>
__global__ __launch_bounds__(256) void nb_box_pass_kernel(const cd *__restrict__ X, cd *__restrict__ Y,
                                                          const cd *__restrict__ E, int M, int K, int kc, int64_t inner,
                                                          int64_t ncols, int64_t pitch, const int32_t *__restrict__ kinv,
                                                          cd *__restrict__ out, int64_t nk_stride, int nn)
{
    box_pass_body<MM, SCATTER>(X, Y, E, M, K, kc, inner, ncols, pitch, kinv, out, nk_stride, nn,
                               static_cast<int>(blockIdx.z));
}

// Up to four evaluations of the separable sums at once (the fused RK4 steps'
// evaluations of one step): B on the box, then the passes along a3, a2 and
// (with out set) a1, block z = (evaluation, k chunk); each evaluation has its
// own scratch and output.
struct BoxMulti {
    NbEval ev[4];
    int nev;
    cd *b[4];   // B on the box [cells][bcols]
    cd *t1[4];  // first pass [M1][M2][K3][bcols]
    cd *t2[4];  // second pass [M1][K2][K3][bcols]
    cd *out[4];  // matrices [nmat][nk_stride][N][N] (the pass along a1)
};

__global__ void nb_box_bform_multi_kernel(const NbArgs a, const BoxMulti bm, int64_t ncells)
{
    const int e = blockIdx.z;
    const NbEval &ev = bm.ev[e];
    const int64_t ncols = static_cast<int64_t>(ev.nmat) * a.N * a.N;
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= ncells * ncols) return;
    const int64_t cell = idx / ncols, c = idx % ncols;
    const int r = a.box_r[cell];
    cd v = cmk(0.0, 0.0);
    if (r >= 0) {
        const double *R = a.rcart + 3 * r;
        double sn, cs;
        sincos(ev.A[0] * R[0] + ev.A[1] * R[1] + ev.A[2] * R[2], &sn, &cs);
        v = b_element(a, ev, r, c, cmk(cs, sn));
    }
    bm.b[e][cell * a.bcols + c] = v;
}

// pass 0: B -> T1 along a3; pass 1: T1 -> T2 along a2; pass 2: T2 -> the matrices
// along a1 (scattered to the k points)
template <int MM, int PASS>
__global__ __launch_bounds__(256) void nb_box_pass_multi_kernel(const NbArgs a, const BoxMulti bm, int kc, int nz)
{
    const int e = static_cast<int>(blockIdx.z) / nz, kz = static_cast<int>(blockIdx.z) - e * nz;
    const int nn = a.N * a.N;
    const int64_t ncols = static_cast<int64_t>(bm.ev[e].nmat) * nn;
    const int M1 = a.boxM[0], M2 = a.boxM[1], M3 = a.boxM[2], K1 = a.boxK[0], K2 = a.boxK[1], K3 = a.boxK[2];
    if (PASS == 0)
        box_pass_body<MM, false>(bm.b[e], bm.t1[e], a.boxE[2], M3, K3, kc, 1, ncols, a.bcols, nullptr, nullptr, a.nk_stride,
                          nn, kz);
    else
        box_pass_body<MM, false>(bm.t1[e], bm.t2[e], a.boxE[1], M2, K2, kc, K3, ncols, a.bcols, nullptr, nullptr, a.nk_stride,
                          nn, kz);
    (void)M1;
    (void)K1;
}

// evaluation e of a run of RK4 steps (NbBoxSteps): its A, E, matrices and outputs
__device__ __forceinline__ NbEval box_step_eval(const NbBoxSteps &d, int e, cd *&t2)
{
    const int sl = e / 3, typ = e - 3 * sl;
    const int64_t step = d.s0 + sl;
    NbEval ev{};
    const double *lt = d.laser + (step * 3 + typ) * 6;
    for (int i = 0; i < 3; ++i) {
        ev.A[i] = lt[i];
        ev.E[i] = lt[3 + i];
    }
    if (typ == 0) {
        ev.nmat = 3;
        ev.mats[0] = kMatJx;
        ev.mats[1] = kMatJy;
        ev.mats[2] = kMatJz;
        t2 = d.tj + sl * d.t2_stride;
    } else {
        ev.nmat = 2;
        ev.mats[0] = kMatHtilde;
        ev.mats[1] = kMatH0;
        t2 = typ == 1 ? d.th + sl * d.t2_stride : d.ring + (step % d.nring) * d.t2_stride;
    }
    return ev;
}

__global__ void nb_box_steps_bform_kernel(const NbArgs a, const NbBoxSteps d, int64_t ncells)
{
    const int e = blockIdx.z;
    cd *t2;
    const NbEval ev = box_step_eval(d, e, t2);
    const int nn = a.N * a.N;
    const int64_t ncols = static_cast<int64_t>(ev.nmat) * nn;
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= ncells * ncols) return;
    const int64_t cell = idx / ncols, c = idx % ncols;
    const int r = a.box_r[cell];
    cd v = cmk(0.0, 0.0);
    if (r >= 0) {
        const double *R = a.rcart + 3 * r;
        double sn, cs;
        sincos(ev.A[0] * R[0] + ev.A[1] * R[1] + ev.A[2] * R[2], &sn, &cs);
        // the kind without indexing the evaluation's list (J_x, J_y, J_z; H~, H)
        const int mat = static_cast<int>(c / nn);
        const int type = ev.nmat == 3 ? kMatJx + mat : (mat == 0 ? kMatHtilde : kMatH0);
        v = b_value(a, ev.E, type, r, static_cast<int>(c - static_cast<int64_t>(mat) * nn), cmk(cs, sn));
    }
    (d.b + e * d.b_stride)[cell * a.bcols + c] = v;
}

template <int MM, int PASS>
__global__ __launch_bounds__(256) void nb_box_steps_pass_kernel(const NbArgs a, const NbBoxSteps d, int kc, int nz)
{
    const int e = static_cast<int>(blockIdx.z) / nz, kz = static_cast<int>(blockIdx.z) - e * nz;
    cd *t2;
    const NbEval ev = box_step_eval(d, e, t2);
    const int nn = a.N * a.N;
    const int64_t ncols = static_cast<int64_t>(ev.nmat) * nn;
    const int M2 = a.boxM[1], M3 = a.boxM[2], K2 = a.boxK[1], K3 = a.boxK[2];
    cd *t1 = d.t1 + e * d.t1_stride;
    if (PASS == 0)
        box_pass_body<MM, false>(d.b + e * d.b_stride, t1, a.boxE[2], M3, K3, kc, 1, ncols, a.bcols, nullptr, nullptr,
                                 a.nk_stride, nn, kz);
    else
        box_pass_body<MM, false>(t1, t2, a.boxE[1], M2, K2, kc, K3, ncols, a.bcols, nullptr, nullptr, a.nk_stride, nn,
                                 kz);
}

// pass along a1: T2 -> the matrices, scattered to the k points
template <int MM>
__global__ __launch_bounds__(256) void nb_box_a1_multi_kernel(const NbArgs a, const BoxMulti bm, int kc, int nz)
{
    const int e = static_cast<int>(blockIdx.z) / nz, kz = static_cast<int>(blockIdx.z) - e * nz;
    const int nn = a.N * a.N;
    const int64_t ncols = static_cast<int64_t>(bm.ev[e].nmat) * nn;
    box_pass_body<MM, true>(bm.t2[e], nullptr, a.boxE[0], a.boxM[0], a.boxK[0], kc, static_cast<int64_t>(a.boxK[1]) * a.boxK[2],
                      ncols, a.bcols, a.box_kinv, bm.out[e], a.nk_stride, nn, kz);
}

// the padding k points [nk, nk_stride) of every matrix: zero, as the dense sums give them
__global__ void nb_box_pad_kernel(const NbArgs a, int64_t ncols, cd *__restrict__ out)
{
    const int64_t npad = a.nk_stride - a.nk;
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= npad * ncols) return;
    const int64_t k = a.nk + idx / ncols, c = idx % ncols;
    const int nn = a.N * a.N;
    const int64_t mat = c / nn, ij = c % nn;
    out[(mat * a.nk_stride + k) * nn + ij] = cmk(0.0, 0.0);
}

template <int MM>
static void box_pass(const cd *X, cd *Y, const cd *E, int M, int K, int64_t outer, int64_t inner, int64_t ncols,
                     int64_t pitch, const int32_t *kinv, cd *out, int64_t nk_stride, int nn, hipStream_t s)
{
    // k per block: about 2048 blocks in all (each block re-reads its M inputs, so
    // fewer, longer blocks where the (o, i, column) grid alone is large)
    const int64_t xy = outer * inner * ((ncols + 255) / 256);
    const int kc = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(K, K * xy / 2048)));
    const dim3 grid(static_cast<unsigned>(outer * inner), static_cast<unsigned>((ncols + 255) / 256),
                    static_cast<unsigned>((K + kc - 1) / kc));
    if (out)
        nb_box_pass_kernel<MM, true><<<grid, 256, 0, s>>>(X, Y, E, M, K, kc, inner, ncols, pitch, kinv, out, nk_stride, nn);
    else
        nb_box_pass_kernel<MM, false><<<grid, 256, 0, s>>>(X, Y, E, M, K, kc, inner, ncols, pitch, kinv, out, nk_stride, nn);
}

static void box_pass_m(const cd *X, cd *Y, const cd *E, int M, int K, int64_t outer, int64_t inner, int64_t ncols,
                       int64_t pitch, const int32_t *kinv, cd *out, int64_t nk_stride, int nn, hipStream_t s)
{
    if (M <= 4) box_pass<4>(X, Y, E, M, K, outer, inner, ncols, pitch, kinv, out, nk_stride, nn, s);
    else if (M <= 8) box_pass<8>(X, Y, E, M, K, outer, inner, ncols, pitch, kinv, out, nk_stride, nn, s);
    else if (M <= 16) box_pass<16>(X, Y, E, M, K, outer, inner, ncols, pitch, kinv, out, nk_stride, nn, s);
    else box_pass<32>(X, Y, E, M, K, outer, inner, ncols, pitch, kinv, out, nk_stride, nn, s);
}

static hipError_t box_fourier(const NbArgs &a, const NbEval &ev, cd *out, hipStream_t s)
{
    const int nn = a.N * a.N;
    const int64_t ncols = static_cast<int64_t>(ev.nmat) * nn;
    const int M1 = a.boxM[0], M2 = a.boxM[1], M3 = a.boxM[2];
    const int K1 = a.boxK[0], K2 = a.boxK[1], K3 = a.boxK[2];
    const int64_t ncells = static_cast<int64_t>(M1) * M2 * M3;
    nb_box_bform_kernel<<<dim3(static_cast<unsigned>((ncells * ncols + 255) / 256)), 256, 0, s>>>(a, ev, ncells, ncols);
    // axis 3: outer (n1, n2), inner 1
    box_pass_m(a.box_b, a.box_t1, a.boxE[2], M3, K3, static_cast<int64_t>(M1) * M2, 1, ncols, a.bcols, nullptr,
               nullptr, a.nk_stride, nn, s);
    // axis 2: outer n1, inner k3
    box_pass_m(a.box_t1, a.box_t2, a.boxE[1], M2, K2, M1, K3, ncols, a.bcols, nullptr, nullptr, a.nk_stride, nn, s);
    // axis 1: outer 1, inner (k2, k3), scattered to the k points of this block
    box_pass_m(a.box_t2, nullptr, a.boxE[0], M1, K1, 1, static_cast<int64_t>(K2) * K3, ncols, a.bcols, a.box_kinv,
               out, a.nk_stride, nn, s);
    const int64_t npad = (a.nk_stride - a.nk) * ncols;
    if (npad > 0) nb_box_pad_kernel<<<dim3(static_cast<unsigned>((npad + 255) / 256)), 256, 0, s>>>(a, ncols, out);
    return hipGetLastError();
}

// bform and the passes along a3 and a2 of one evaluation, the second pass into t2
// ([M1][K2][K3][bcols]); the fused step kernel does the pass along a1
static hipError_t box_t2(const NbArgs &a, const NbEval &ev, cd *t2, hipStream_t s)
{
    const int nn = a.N * a.N;
    const int64_t ncols = static_cast<int64_t>(ev.nmat) * nn;
    const int M1 = a.boxM[0], M2 = a.boxM[1], M3 = a.boxM[2];
    const int K2 = a.boxK[1], K3 = a.boxK[2];
    const int64_t ncells = static_cast<int64_t>(M1) * M2 * M3;
    nb_box_bform_kernel<<<dim3(static_cast<unsigned>((ncells * ncols + 255) / 256)), 256, 0, s>>>(a, ev, ncells, ncols);
    box_pass_m(a.box_b, a.box_t1, a.boxE[2], M3, K3, static_cast<int64_t>(M1) * M2, 1, ncols, a.bcols, nullptr,
               nullptr, a.nk_stride, nn, s);
    box_pass_m(a.box_t1, t2, a.boxE[1], M2, K2, M1, K3, ncols, a.bcols, nullptr, nullptr, a.nk_stride, nn, s);
    return hipGetLastError();
}

// ---------------------------------------- FP16-split Fourier GEMM (SPLIT16) --
// Each real operand x (|x| <= 1 after a power-of-two scale) is carried as
// hi = fp16(x), lo = fp16(x - hi), about 22 significant bits; a product is
// hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x32_f16 with FP32 accumulation
// (the lo*lo term is below FP32's resolution of the sum). The complex
// product re += ar br - ai bi, im += ar bi + ai br is 12 MFMAs per 16 x 16
// tile and K chunk of 32. Operand layout (probed on gfx950): A row = lane %
// 16, B column = lane % 16, both with the same 8 k values per lane (group
// lane / 16); D element r of lane l is row 4 (l / 16) + r, column l % 16.
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
constexpr int kSplitK = 32;

struct NbScale {
    double s[5];  // power-of-two scale of each matrix's B (|B s| <= 1)
};

__device__ __forceinline__ void split16(double x, uint16_t &hi, uint16_t &lo)
{
    const _Float16 h = static_cast<_Float16>(static_cast<float>(x));
    const _Float16 l = static_cast<_Float16>(static_cast<float>(x - static_cast<double>(h)));
    hi = __builtin_bit_cast(uint16_t, h);
    lo = __builtin_bit_cast(uint16_t, l);
}

__global__ void nb_psplit_kernel(const NbArgs a, const cd *__restrict__ P, uint16_t *__restrict__ p16)
{
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t plane = a.nk_stride * a.nr_pad;
    if (idx >= plane) return;
    const int64_t k = idx / a.nr_pad;
    const int r = static_cast<int>(idx % a.nr_pad);
    const cd v = r < a.nr ? P[k * a.nr + r] : cmk(0.0, 0.0);
    uint16_t h, l;
    split16(v.x, h, l);
    p16[idx] = h;
    p16[plane + idx] = l;
    split16(v.y, h, l);
    p16[2 * plane + idx] = h;
    p16[3 * plane + idx] = l;
}

__global__ void nb_bsplit_kernel(const NbArgs a, const NbEval ev, const NbScale sc, uint16_t *__restrict__ b16)
{
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t plane = a.bcols * a.nr_pad;
    if (idx >= plane) return;
    const int64_t c = idx / a.nr_pad;
    const int r = static_cast<int>(idx % a.nr_pad);
    const int nn = a.N * a.N;
    cd v = cmk(0.0, 0.0);
    if (r < a.nr && c < static_cast<int64_t>(ev.nmat) * nn) {
        const double *R = a.rcart + 3 * r;
        double sn, cs;
        sincos(ev.A[0] * R[0] + ev.A[1] * R[1] + ev.A[2] * R[2], &sn, &cs);
        v = cscale(b_element(a, ev, r, c, cmk(cs, sn)), sc.s[c / nn]);
    }
    uint16_t h, l;
    split16(v.x, h, l);
    b16[idx] = h;
    b16[plane + idx] = l;
    split16(v.y, h, l);
    b16[2 * plane + idx] = h;
    b16[3 * plane + idx] = l;
}

__device__ __forceinline__ h8 neg8(h8 v)
{
    typedef uint32_t u4 __attribute__((ext_vector_type(4)));
    return __builtin_bit_cast(h8, __builtin_bit_cast(u4, v) ^ u4{0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u});
}

// 64 k x 64 column tile per block of 4 waves; wave w owns the 32 x 32
// sub-tile (w / 2, w % 2) as 2 x 2 MFMA tiles of 16 x 16. Operands are read
// straight from the plane arrays (16-byte loads of 8 consecutive R).
__global__ __launch_bounds__(256) void nb_fourier16_kernel(const NbArgs a, const NbEval ev, const NbScale sc,
                                                           const uint16_t *__restrict__ p16,
                                                           const uint16_t *__restrict__ b16, cd *__restrict__ out)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int64_t k0 = static_cast<int64_t>(blockIdx.x) * 64 + 32 * wm;
    const int64_t c0 = static_cast<int64_t>(blockIdx.y) * 64 + 32 * wn;
    const int nn = a.N * a.N;
    const int64_t ncols = static_cast<int64_t>(ev.nmat) * nn;
    const int64_t PS = a.nk_stride * a.nr_pad, BS = a.bcols * a.nr_pad;
    f4 accr[2][2], acci[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            accr[i][j] = f4{0.f, 0.f, 0.f, 0.f};
            acci[i][j] = f4{0.f, 0.f, 0.f, 0.f};
        }
    for (int r0 = 0; r0 < a.nr_pad; r0 += kSplitK) {
        const int kk = r0 + 8 * (lane >> 4);
        h8 A[2][4], B[2][4];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int64_t row = k0 + 16 * i + (lane & 15);
#pragma unroll
            for (int pl = 0; pl < 4; ++pl)
                A[i][pl] = *reinterpret_cast<const h8 *>(p16 + pl * PS + row * a.nr_pad + kk);
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int64_t col = c0 + 16 * j + (lane & 15);
#pragma unroll
            for (int pl = 0; pl < 4; ++pl)
                B[j][pl] = *reinterpret_cast<const h8 *>(b16 + pl * BS + col * a.nr_pad + kk);
        }
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const h8 nih = neg8(A[i][2]), nil = neg8(A[i][3]);
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                f4 re = accr[i][j], im = acci[i][j];
                re = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[i][0], B[j][0], re, 0, 0, 0);
                re = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[i][0], B[j][1], re, 0, 0, 0);
                re = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[i][1], B[j][0], re, 0, 0, 0);
                re = __builtin_amdgcn_mfma_f32_16x16x32_f16(nih, B[j][2], re, 0, 0, 0);
                re = __builtin_amdgcn_mfma_f32_16x16x32_f16(nih, B[j][3], re, 0, 0, 0);
                re = __builtin_amdgcn_mfma_f32_16x16x32_f16(nil, B[j][2], re, 0, 0, 0);
                im = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[i][0], B[j][2], im, 0, 0, 0);
                im = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[i][0], B[j][3], im, 0, 0, 0);
                im = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[i][1], B[j][2], im, 0, 0, 0);
                im = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[i][2], B[j][0], im, 0, 0, 0);
                im = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[i][2], B[j][1], im, 0, 0, 0);
                im = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[i][3], B[j][0], im, 0, 0, 0);
                accr[i][j] = re;
                acci[i][j] = im;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int64_t k = k0 + 16 * i + 4 * (lane >> 4) + q;
                const int64_t c = c0 + 16 * j + (lane & 15);
                if (c >= ncols) continue;
                const int mat = static_cast<int>(c / nn), ij = static_cast<int>(c % nn);
                const double inv = 1.0 / sc.s[mat];
                out[(static_cast<int64_t>(mat) * a.nk_stride + k) * nn + ij] =
                    cmk(static_cast<double>(accr[i][j][q]) * inv, static_cast<double>(acci[i][j][q]) * inv);
            }
}

// ------------------------------------------------- wave matrix helpers --
// All matrices are N x N complex, row-major, in LDS; one wave (64 lanes).
__device__ __forceinline__ void wave_sync() { __syncthreads(); }

// C = op(A) op(B); adjA/adjB select the conjugate transpose. On FP64 MFMA:
// 16 x 16 output tiles (ceil(N/16)^2 of them), K in steps of 4, the complex
// product as four real v_mfma_f64_16x16x4f64 (layout as in the Fourier GEMM:
// A row = lane % 16, k = lane / 16; B k = lane / 16, column = lane % 16; D
// element r of lane l is row l / 16 + 4 r, column l % 16). Operands read
// from LDS once per k step instead of once per output element; entries
// beyond N read as zero. C must not alias A or B.
__device__ void wmm(cd *C, const cd *A, const cd *B, int N, bool adjA, bool adjB)
{
    const int lane = threadIdx.x & 63;
    const int r16 = lane & 15, g = lane >> 4;
    const int T = (N + 15) >> 4, K4 = (N + 3) >> 2;
    // output tiles are dealt to the block's waves
    for (int tt = static_cast<int>(threadIdx.x >> 6); tt < T * T; tt += static_cast<int>(blockDim.x >> 6)) {
        const int ti = tt / T, tj = tt % T;
        {
            d4 re = d4{0.0, 0.0, 0.0, 0.0}, im = d4{0.0, 0.0, 0.0, 0.0};
            const int i = 16 * ti + r16, j = 16 * tj + r16;
            for (int kb = 0; kb < K4; ++kb) {
                const int l = 4 * kb + g;
                cd x = cmk(0.0, 0.0), y = cmk(0.0, 0.0);
                if (i < N && l < N) x = adjA ? cconj(A[l * N + i]) : A[i * N + l];
                if (l < N && j < N) y = adjB ? cconj(B[j * N + l]) : B[l * N + j];
                re = __builtin_amdgcn_mfma_f64_16x16x4f64(x.x, y.x, re, 0, 0, 0);
                re = __builtin_amdgcn_mfma_f64_16x16x4f64(x.y, y.y, re, 0, 0, 1);  // 1: A negated
                im = __builtin_amdgcn_mfma_f64_16x16x4f64(x.x, y.y, im, 0, 0, 0);
                im = __builtin_amdgcn_mfma_f64_16x16x4f64(x.y, y.x, im, 0, 0, 0);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * ti + g + 4 * r, col = 16 * tj + r16;
                if (row < N && col < N) C[row * N + col] = cmk(re[r], im[r]);
            }
        }
    }
    wave_sync();
}

__device__ void wcopy(cd *D, const cd *S, int nn)
{
    for (int e = threadIdx.x; e < nn; e += blockDim.x) D[e] = S[e];
    wave_sync();
}

// sum / max over the block, in a fixed order (each wave's lanes, then the waves)
__device__ __forceinline__ double block_sum(double v)
{
    __shared__ double red[16];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < static_cast<int>(blockDim.x >> 6); ++i) t += red[i];
    __syncthreads();
    return t;
}

__device__ __forceinline__ double block_max(double v)
{
    __shared__ double red[16];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < static_cast<int>(blockDim.x >> 6); ++i) t = fmax(t, red[i]);
    __syncthreads();
    return t;
}

// Hermitian eigensolver: A (destroyed) -> ev ascending, U columns = vectors.
// Parallel-order cyclic Jacobi with complex rotations G = diag(1, e^{-i phi})
// R(theta); rotation pairs follow the round-robin tournament. work: N*N.
// warm: U holds eigenvectors of a nearby matrix on entry; the sweeps then
// start from U^dag A U, which is nearly diagonal, and converge in 1-2 sweeps.
__device__ void weigh(cd *A, cd *U, double *ev, cd *work, int N, double4 *rot, bool warm)
{
    const int lane = threadIdx.x;
    const int nn = N * N;
    const int M = (N + 1) & ~1;  // even number of players (a dummy if N odd)
    if (warm) {
        wmm(work, A, U, N, false, false);
        wmm(A, U, work, N, true, false);
    } else {
        for (int e = lane; e < nn; e += 64) {
            const int i = e / N, j = e % N;
            U[e] = cmk(i == j ? 1.0 : 0.0, 0.0);
        }
    }
    for (int e = lane; e < nn; e += 64) {
        const int i = e / N, j = e % N;
        if (i == j) A[e].y = 0.0;
    }
    wave_sync();
    for (int sweep = 0; sweep < 40; ++sweep) {
        double off = 0.0, dg = 0.0;
        for (int e = lane; e < nn; e += 64) {
            const int i = e / N, j = e % N;
            if (i == j) dg += A[e].x * A[e].x;
            else off += A[e].x * A[e].x + A[e].y * A[e].y;
        }
        off = wave_sum(off);
        dg = wave_sum(dg);
        if (off <= 1e-30 * dg || off < 1e-300) break;
        for (int round = 0; round < M - 1; ++round) {
            if (lane < M / 2) {
                // round-robin: player 0 fixed, players 1..M-1 rotate
                auto pos = [&](int s) { return s == 0 ? 0 : 1 + (s - 1 + round) % (M - 1); };
                int p = pos(lane), q = pos(M - 1 - lane);
                if (p > q) { const int t = p; p = q; q = t; }
                double4 g = make_double4(1.0, 0.0, 1.0, 0.0);  // (c, s, e^{-i phi})
                int pq = -1;
                if (q < N) {
                    const cd b = A[p * N + q];
                    const double ab = sqrt(b.x * b.x + b.y * b.y);
                    if (ab > 1e-300) {
                        const double th = (A[q * N + q].x - A[p * N + p].x) / (2.0 * ab);
                        const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                        const double c = 1.0 / sqrt(t * t + 1.0);
                        g = make_double4(c, t * c, b.x / ab, -b.y / ab);
                        pq = p * 64 + q;
                    }
                }
                rot[lane] = g;
                reinterpret_cast<int *>(rot + 32)[lane] = pq;
            }
            wave_sync();
            const int *pqs = reinterpret_cast<const int *>(rot + 32);
            const int half = M / 2;
            // columns: A <- A G, U <- U G
            for (int t = lane; t < half * N; t += 64) {
                const int j = t / N, r = t % N;
                const int pq = pqs[j];
                if (pq < 0) continue;
                const int p = pq / 64, q = pq % 64;
                const double4 g = rot[j];
                const cd em = cmk(g.z, g.w);  // e^{-i phi}
                const cd gqp = cscale(em, -g.y), gqq = cscale(em, g.x);
                cd ap = A[r * N + p], aq = A[r * N + q];
                A[r * N + p] = cadd(cscale(ap, g.x), cmul(aq, gqp));
                A[r * N + q] = cadd(cscale(ap, g.y), cmul(aq, gqq));
                cd up = U[r * N + p], uq = U[r * N + q];
                U[r * N + p] = cadd(cscale(up, g.x), cmul(uq, gqp));
                U[r * N + q] = cadd(cscale(up, g.y), cmul(uq, gqq));
            }
            wave_sync();
            // rows: A <- G^dag A
            for (int t = lane; t < half * N; t += 64) {
                const int j = t / N, r = t % N;
                const int pq = pqs[j];
                if (pq < 0) continue;
                const int p = pq / 64, q = pq % 64;
                const double4 g = rot[j];
                const cd ep = cmk(g.z, -g.w);  // e^{+i phi}
                const cd ap = A[p * N + r], aq = A[q * N + r];
                A[p * N + r] = csub(cscale(ap, g.x), cmul(cscale(ep, g.y), aq));
                A[q * N + r] = cadd(cscale(ap, g.y), cmul(cscale(ep, g.x), aq));
            }
            wave_sync();
            // exact zeros / real diagonal of the rotated 2x2 blocks
            if (lane < half) {
                const int pq = pqs[lane];
                if (pq >= 0) {
                    const int p = pq / 64, q = pq % 64;
                    A[p * N + q] = cmk(0.0, 0.0);
                    A[q * N + p] = cmk(0.0, 0.0);
                    A[p * N + p].y = 0.0;
                    A[q * N + q].y = 0.0;
                }
            }
            wave_sync();
        }
    }
    // ascending order: rank of each eigenvalue (ties by index), permute U
    if (lane < N) {
        const double e = A[lane * N + lane].x;
        int rank = 0;
        for (int j = 0; j < N; ++j) {
            const double f = A[j * N + j].x;
            rank += (f < e) || (f == e && j < lane);
        }
        ev[rank] = e;
        for (int r = 0; r < N; ++r) work[r * N + rank] = U[r * N + lane];
    }
    wave_sync();
    wcopy(U, work, nn);
}

// Warm-started eigenvectors for the LDS kernels: the simultaneous-rotation
// refinement of nband16.hip's eig_refine on LDS matrices and wmm products.
// In: A = H (overwritten; its Hermitian part is used), U = eigenvectors of a
// nearby H; Hsrc = H in memory (reloaded when W is rebuilt). Out: U refined,
// columns in the warm start's order. W, T, X: scratch. Pointers are swapped
// among the five slots.
__device__ void weigh_warm(cd *&A, cd *&U, cd *&W, cd *&T, cd *&X, int N, double *lam, const cd *Hsrc)
{
    const int lane = threadIdx.x;
    const int nn = N * N;
    // A <- the Hermitian part of H (as zheevr, which reads one triangle)
    auto hermitize = [&]() {
        for (int e = lane; e < nn; e += blockDim.x) {
            const int i = e / N, j = e % N;
            const cd x = A[e], y = A[j * N + i];
            X[e] = cmk(0.5 * (x.x + y.x), 0.5 * (x.y - y.y));
        }
        wave_sync();
        cd *t = A;
        A = X;
        X = t;
    };
    // U <- U (3I - U^dag U) / 2; returns max |U^dag U - I| before the step
    auto newton_schulz = [&]() {
        wmm(T, U, U, N, true, false);
        double e = 0.0;
        for (int q = lane; q < nn; q += blockDim.x) {
            const int i = q / N, j = q % N;
            const double dr = T[q].x - (i == j ? 1.0 : 0.0);
            e = fmax(e, fmax(fabs(dr), fabs(T[q].y)));
            T[q] = cmk((i == j ? 1.5 : 0.0) - 0.5 * T[q].x, -0.5 * T[q].y);
        }
        e = block_max(e);
        wmm(X, U, T, N, false, false);
        cd *t = U;
        U = X;
        X = t;
        return e;
    };
    hermitize();
    wmm(T, A, U, N, false, false);
    wmm(W, U, T, N, true, false);  // W = U^dag H U
    for (int it = 0; it < 30; ++it) {
        if (lane < N) lam[lane] = W[lane * N + lane].x;
        wave_sync();
        double kmax = 0.0;  // K = rotation generators, into A
        for (int e = lane; e < nn; e += blockDim.x) {
            const int i = e / N, j = e % N;
            const cd w = W[e];
            const double w2 = w.x * w.x + w.y * w.y;
            cd kk = cmk(0.0, 0.0);
            if (i != j && w2 > 0.0) {
                const double D = lam[j] - lam[i];
                const double s = (D > 0.0 || (D == 0.0 && i < j)) ? 1.0 : -1.0;
                const double den = 0.5 * D + s * sqrt(0.25 * D * D + w2);
                kk = cmk(w.x / den, w.y / den);
                kmax = fmax(kmax, sqrt(kk.x * kk.x + kk.y * kk.y));
            }
            A[e] = kk;
        }
        kmax = block_max(kmax);
        if (kmax == 0.0) break;
        const bool small = kmax < 1e-8;
        const double damp = kmax > 0.5 ? 0.5 / kmax : 1.0;  // cold-like start: damped rotations
        if (damp != 1.0) {
            for (int e = lane; e < nn; e += blockDim.x) A[e] = cscale(A[e], damp);
            wave_sync();
        }
        if (!small) wmm(T, A, A, N, false, false);  // K^2
        for (int e = lane; e < nn; e += blockDim.x) {  // V = I + K (+ K^2 / 2)
            const int i = e / N, j = e % N;
            cd v = A[e];
            if (!small) v = cadd(v, cscale(T[e], 0.5));
            if (i == j) v.x += 1.0;
            A[e] = v;
        }
        wave_sync();
        wmm(T, U, A, N, false, false);  // U V
        {
            cd *t = U;
            U = T;
            T = t;
        }
        if (small) break;
        if (kmax > 1e-3) {  // V far from unitary: restore U fully, then W from H itself
            for (int q = 0; q < 8; ++q)
                if (newton_schulz() < 1e-8) break;
            for (int e = lane; e < nn; e += blockDim.x) A[e] = Hsrc[e];
            wave_sync();
            hermitize();
            wmm(T, A, U, N, false, false);
            wmm(W, U, T, N, true, false);
        } else {
            wmm(T, W, A, N, false, false);  // W V
            wmm(X, A, T, N, true, false);   // V^dag W V
            cd *t = W;
            W = X;
            X = t;
        }
    }
    newton_schulz();
}

// d[c] = Re (U^dag M)_cc = Re sum_r conj(U_rc) M_rc for c < N (threads c);
// ends with a block barrier, so M may be overwritten next
__device__ __forceinline__ double col_dot_lds(const cd *U, const cd *M, int N)
{
    const int c = threadIdx.x;
    double s = 0.0;
    if (c < N)
        for (int r = 0; r < N; ++r) s += U[r * N + c].x * M[r * N + c].x + U[r * N + c].y * M[r * N + c].y;
    wave_sync();
    return s;
}

// LDS layout of one wave: kMats matrices of N*N complex, N eigenvalues,
// 32 rotation slots + 32 int pair slots.
__device__ __forceinline__ cd *mat_slot(cd *base, int i, int nn) { return base + static_cast<int64_t>(i) * nn; }

// ----------------------------------------------------------- init state --
// rho(k) = U diag(occupation) U^dag with U = eig(H(k)) (SBEs.h:284-364).
__global__ __launch_bounds__(64) void nb_init_kernel(const NbArgs a, const cd *__restrict__ H0, cd *__restrict__ rho,
                                                     cd *__restrict__ Ust)
{
    extern __shared__ double4 smem4[];
    const int N = a.N, nn = N * N;
    cd *base = reinterpret_cast<cd *>(smem4);
    cd *A = mat_slot(base, 0, nn), *U = mat_slot(base, 1, nn), *W = mat_slot(base, 2, nn);
    double *ev = reinterpret_cast<double *>(mat_slot(base, 3, nn));
    double4 *rot = reinterpret_cast<double4 *>(ev + 64);
    const int64_t k = blockIdx.x;
    cd *out = rho + k * nn;
    if (k >= a.nk) {
        for (int e = threadIdx.x; e < nn; e += 64) out[e] = cmk(0.0, 0.0);
        return;
    }
    wcopy(A, H0 + k * nn, nn);
    weigh(A, U, ev, W, N, rot, false);
    for (int e = threadIdx.x; e < nn; e += 64) Ust[k * nn + e] = U[e];
    // occupations: first nval bands, or Fermi-Dirac normalised to 1
    __shared__ double occ[64];
    if (threadIdx.x == 0) {
        double sum = 0.0;
        for (int m = 0; m < N; ++m) {
            double f;
            if (a.fermi) f = 1.0 / (exp((ev[m] - a.fermi_energy) / (300.0 / 3.15775024804e5)) + 1.0);
            else f = m < a.nval ? 1.0 : 0.0;
            occ[m] = f;
            sum += f;
        }
        if (a.fermi)
            for (int m = 0; m < N; ++m) occ[m] /= sum;
    }
    wave_sync();
    for (int e = threadIdx.x; e < nn; e += 64) {
        const int i = e / N, j = e % N;
        cd s = cmk(0.0, 0.0);
        for (int m = 0; m < N; ++m) cfma(s, cscale(U[i * N + m], occ[m]), cconj(U[j * N + m]));
        out[e] = s;
    }
}

// ---------------------------------------------------------- time sample --
// Per k at t_n: U = eig(H(k+A)), rho_H = U^dag rho U, J_H = U^dag J U;
// intra = sum_i Re(rho_H,ii J_H,ii), inter = the rest (or the total
// Re Tr(rho J) when inter_intra == 0); band populations diag(rho_H); and
// w rho for the orbital-basis sum. Each wave walks k = wave, wave + G, ...
// and writes one row of partial sums: [inter(3) intra(3) pop(N) rho(2N^2)].
__global__ __launch_bounds__(256) void nb_sample_kernel(const NbArgs a, const cd *__restrict__ mats,
                                                       const cd *__restrict__ rho, double *__restrict__ partials,
                                                       int64_t step_local, cd *__restrict__ Ust, int ureuse)
{
    extern __shared__ double4 smem4[];
    const int N = a.N, nn = N * N;
    cd *base = reinterpret_cast<cd *>(smem4);
    cd *A = mat_slot(base, 0, nn), *U = mat_slot(base, 1, nn), *W = mat_slot(base, 2, nn);
    cd *R = mat_slot(base, 3, nn), *T = mat_slot(base, 4, nn), *X = mat_slot(base, 5, nn);
    double *ev = reinterpret_cast<double *>(mat_slot(base, 6, nn));
    const int lane = threadIdx.x;
    const int nv = a.nv;
    // per-lane accumulators: rho elements e = lane + 64 t
    constexpr int kMaxPer = 1024 / 64;
    double racc[2 * kMaxPer];
    for (int t = 0; t < 2 * kMaxPer; ++t) racc[t] = 0.0;
    double cur[6] = {0, 0, 0, 0, 0, 0};
    double pop = 0.0;  // lane m < N: population of the m-th lowest band
    const cd *H0 = mats + a.nk_stride * nn;           // mat 1: H0
    const cd *J = mats + 2 * a.nk_stride * nn;        // mats 2..4: J_x, J_y, J_z
    for (int64_t k = blockIdx.x; k < a.nk; k += gridDim.x) {
        const double w = a.weight[k];
        wcopy(R, rho + k * nn, nn);
        wcopy(U, Ust + k * nn, nn);
        wcopy(A, H0 + k * nn, nn);
        if (!ureuse) {  // eigenvectors at k + A(t_n), warm-started from the last ones
            weigh_warm(A, U, W, T, X, N, ev, H0 + k * nn);
            for (int e = lane; e < nn; e += blockDim.x) Ust[k * nn + e] = U[e];
            wcopy(A, H0 + k * nn, nn);
        }
        // band energies (ascending order) and populations of U's columns
        wmm(T, A, U, N, false, false);
        const double lam = col_dot_lds(U, T, N);
        wmm(T, R, U, N, false, false);
        const double p = col_dot_lds(U, T, N);
        int rank = 0;
        for (int j = 0; j < N; ++j) {
            const double lj = __shfl(lam, j, 64);
            rank += (lj < lam) || (lj == lam && j < lane);
        }
        for (int j = 0; j < N; ++j) {
            const int rj = __shfl(rank, j, 64);
            const double pj = __shfl(p, j, 64);
            if (rj == lane) pop += pj * w;
        }
        for (int e = lane, t = 0; e < nn; e += blockDim.x, ++t) {
            racc[2 * t] += R[e].x * w;
            racc[2 * t + 1] += R[e].y * w;
        }
        for (int ax = 0; ax < 3; ++ax) {
            const cd *Ja = J + static_cast<int64_t>(ax) * a.nk_stride * nn + k * nn;
            double tot = 0.0;  // Re Tr(rho J) = sum_ij Re(rho_ij J_ji)
            for (int e = lane; e < nn; e += blockDim.x) {
                const int i = e / N, j = e % N;
                const cd r = R[e], jj = Ja[j * N + i];
                tot += r.x * jj.x - r.y * jj.y;
            }
            tot = block_sum(tot);
            double intra = 0.0;
            if (a.inter_intra) {  // sum_c (U^dag rho U)_cc Re (U^dag J U)_cc
                wcopy(A, Ja, nn);
                wmm(T, A, U, N, false, false);
                const double jd = col_dot_lds(U, T, N);
                intra = block_sum(lane < N ? p * jd : 0.0);
            }
            cur[ax] += (tot - intra) * w;
            cur[3 + ax] += intra * w;
            wave_sync();
        }
    }
    double *row = partials + (step_local * gridDim.x + blockIdx.x) * static_cast<int64_t>(nv);
    if (lane < 6) row[lane] = cur[lane];
    if (lane < N) row[6 + lane] = pop;
    for (int e = lane, t = 0; e < nn; e += blockDim.x, ++t) {
        row[6 + N + 2 * e] = racc[2 * t];
        row[6 + N + 2 * e + 1] = racc[2 * t + 1];
    }
}

// ------------------------------------------------------------- RK stage --
// Stage i of the step: y = rho + dt sum_j a_ij k_j; k_i = f(y) with
// f(y) = -i(H~ y - (H~ y)^dag) - U (Gamma o (U^dag y U)) U^dag; after the
// last stage rho += dt sum_i b_i k_i (SBEs.h:596-624, 762-861). The
// dephasing is formed as g2 y + (g1 - g2) U diag(U^dag y U) U^dag.
__global__ __launch_bounds__(256) void nb_rhs_kernel(const NbArgs a, const NbStage sg, const cd *__restrict__ mats,
                                                    cd *__restrict__ rho, cd *__restrict__ kbuf, cd *__restrict__ Ust)
{
    extern __shared__ double4 smem4[];
    const int N = a.N, nn = N * N;
    cd *base = reinterpret_cast<cd *>(smem4);
    cd *A = mat_slot(base, 0, nn), *U = mat_slot(base, 1, nn), *W = mat_slot(base, 2, nn);
    cd *Y = mat_slot(base, 3, nn), *T = mat_slot(base, 4, nn), *X = mat_slot(base, 5, nn);
    double *ev = reinterpret_cast<double *>(mat_slot(base, 6, nn));
    const int64_t k = blockIdx.x;
    if (k >= a.nk) return;
    const int lane = threadIdx.x;
    const int64_t kstride = a.nk_stride * nn;
    cd *rk = rho + k * nn;
    // stage input
    for (int e = lane; e < nn; e += blockDim.x) {
        cd y = rk[e];
        for (int j = 0; j < sg.stage; ++j)
            if (sg.a[j] != 0.0) y = cadd(y, cscale(kbuf[j * kstride + k * nn + e], sg.a[j]));
        Y[e] = y;
    }
    wave_sync();
    if (a.deph) {
        wcopy(U, Ust + k * nn, nn);
        if (!sg.ureuse) {  // eigenvectors at this stage time, warm-started
            const cd *H0 = mats + a.nk_stride * nn + k * nn;  // mat 1: H
            wcopy(A, H0, nn);
            weigh_warm(A, U, W, T, X, N, ev, H0);
            for (int e = lane; e < nn; e += blockDim.x) Ust[k * nn + e] = U[e];
        }
        wmm(T, Y, U, N, false, false);  // y U
        const double d = col_dot_lds(U, T, N);
        if (lane < N) ev[lane] = d;  // (U^dag y U)_cc
        wave_sync();
        for (int e = lane; e < nn; e += blockDim.x) T[e] = cscale(U[e], ev[e % N]);  // U diag(d)
        wave_sync();
        wmm(A, T, U, N, false, true);  // U diag(d) U^dag
        for (int e = lane; e < nn; e += blockDim.x)
            A[e] = cadd(cscale(Y[e], a.gam2), cscale(A[e], a.gam1 - a.gam2));  // the dephasing term
        wave_sync();
    }
    const cd *Ht = mats + k * nn;                 // mat 0: H~ = H + E.D
    wcopy(W, Ht, nn);
    wmm(T, W, Y, N, false, false);                // T = H~ y
    cd *ko = kbuf + sg.stage * kstride + k * nn;
    // -i (T - T^dag): (-i)(t_ij - conj(t_ji))
    for (int e = lane; e < nn; e += blockDim.x) {
        const int i = e / N, j = e % N;
        const cd d = csub(T[i * N + j], cconj(T[j * N + i]));
        cd x = cmk(d.y, -d.x);
        if (a.deph) x = csub(x, A[e]);
        X[e] = x;
    }
    wave_sync();
    for (int e = lane; e < nn; e += blockDim.x) ko[e] = X[e];
    if (sg.last) {
        for (int e = lane; e < nn; e += blockDim.x) {
            cd r = rk[e];
            for (int j = 0; j < sg.nstage; ++j)
                if (sg.b[j] != 0.0) r = cadd(r, cscale(j == sg.stage ? X[e] : kbuf[j * kstride + k * nn + e], sg.b[j]));
            rk[e] = r;
        }
    }
}

// ------------------------------------------------- velocity gauge ----
// Set-up of the VelocityHamiltonian gauge (SBEs.h:284-373): per k,
// U = eig(H(k)) with the column phases fixed (column j times
// conj(U[p][j]) / |U[p][j]|, p the first row with |U[p][j]|^2 >= 1/(2N)),
// E(k), pMatrix_a = U^dag J_a(k) U, occupations, rho = diag(occupations).
__global__ __launch_bounds__(64) void nb_vel_init_kernel(const NbArgs a, const cd *__restrict__ mats,
                                                         cd *__restrict__ P, double *__restrict__ Eocc,
                                                         cd *__restrict__ Ust, cd *__restrict__ rho)
{
    extern __shared__ double4 smem4[];
    const int N = a.N, nn = N * N;
    cd *base = reinterpret_cast<cd *>(smem4);
    cd *A = mat_slot(base, 0, nn), *U = mat_slot(base, 1, nn), *W = mat_slot(base, 2, nn);
    cd *T = mat_slot(base, 3, nn), *X = mat_slot(base, 4, nn);
    double *ev = reinterpret_cast<double *>(mat_slot(base, 6, nn));
    double4 *rot = reinterpret_cast<double4 *>(ev + 64);
    const int64_t k = blockIdx.x;
    const int lane = threadIdx.x;
    const int64_t ks = a.nk_stride * nn;
    if (k >= a.nk) {  // padding: zero everything the step kernel could read
        for (int e = lane; e < nn; e += 64) {
            rho[k * nn + e] = cmk(0.0, 0.0);
            Ust[k * nn + e] = cmk(0.0, 0.0);
            for (int ax = 0; ax < 3; ++ax) P[ax * ks + k * nn + e] = cmk(0.0, 0.0);
        }
        if (lane < 2 * N) Eocc[k * 2 * N + lane] = 0.0;
        return;
    }
    wcopy(A, mats + k * nn, nn);
    weigh(A, U, ev, W, N, rot, false);
    if (lane < N) {
        const int j = lane;
        int p = 0;
        while (p < N - 1 && (U[p * N + j].x * U[p * N + j].x + U[p * N + j].y * U[p * N + j].y) * (2.0 * N) < 1.0) ++p;
        const cd u = U[p * N + j];
        const double au = sqrt(u.x * u.x + u.y * u.y);
        if (au > 0.0) {
            const cd ph = cmk(u.x / au, -u.y / au);
            for (int r = 0; r < N; ++r) U[r * N + j] = cmul(U[r * N + j], ph);
            U[p * N + j] = cmk(au, 0.0);
        }
    }
    wave_sync();
    for (int e = lane; e < nn; e += 64) Ust[k * nn + e] = U[e];
    for (int ax = 0; ax < 3; ++ax) {
        wcopy(X, mats + (1 + ax) * ks + k * nn, nn);
        wmm(T, X, U, N, false, false);
        wmm(W, U, T, N, true, false);  // U^dag J_a U
        for (int e = lane; e < nn; e += 64) P[ax * ks + k * nn + e] = W[e];
    }
    __shared__ double occ[64];
    if (lane == 0) {
        double sum = 0.0;
        for (int m = 0; m < N; ++m) {
            double f;
            if (a.fermi) f = 1.0 / (exp((ev[m] - a.fermi_energy) / (300.0 / 3.15775024804e5)) + 1.0);
            else f = m < a.nval ? 1.0 : 0.0;
            occ[m] = f;
            sum += f;
        }
        if (a.fermi)
            for (int m = 0; m < N; ++m) occ[m] /= sum;
    }
    wave_sync();
    if (lane < N) {
        Eocc[k * 2 * N + lane] = ev[lane];
        Eocc[k * 2 * N + N + lane] = occ[lane];
    }
    for (int e = lane; e < nn; e += 64) {
        const int i = e / N, j = e % N;
        rho[k * nn + e] = cmk(i == j ? occ[i] : 0.0, 0.0);
    }
}

// Band-basis export of the LengthWannier state (SBEs.h:927-946,
// WannierToHamiltonian): U^dag rho U with U = eig(H(k + A)) from mats slot 0
// and the column phase rule of nb_vel_init_kernel.
__global__ __launch_bounds__(64) void nb_band_export_kernel(const NbArgs a, const cd *__restrict__ H,
                                                            const cd *__restrict__ rho, cd *__restrict__ out)
{
    extern __shared__ double4 smem4[];
    const int N = a.N, nn = N * N;
    cd *base = reinterpret_cast<cd *>(smem4);
    cd *A = mat_slot(base, 0, nn), *U = mat_slot(base, 1, nn), *W = mat_slot(base, 2, nn);
    cd *T = mat_slot(base, 3, nn), *R = mat_slot(base, 4, nn);
    double *ev = reinterpret_cast<double *>(mat_slot(base, 6, nn));
    double4 *rot = reinterpret_cast<double4 *>(ev + 64);
    const int64_t k = blockIdx.x;
    const int lane = threadIdx.x;
    if (k >= a.nk) return;
    wcopy(A, H + k * nn, nn);
    weigh(A, U, ev, W, N, rot, false);
    if (lane < N) {
        const int j = lane;
        int p = 0;
        while (p < N - 1 && (U[p * N + j].x * U[p * N + j].x + U[p * N + j].y * U[p * N + j].y) * (2.0 * N) < 1.0) ++p;
        const cd u = U[p * N + j];
        const double au = sqrt(u.x * u.x + u.y * u.y);
        if (au > 0.0) {
            const cd ph = cmk(u.x / au, -u.y / au);
            for (int r = 0; r < N; ++r) U[r * N + j] = cmul(U[r * N + j], ph);
            U[p * N + j] = cmk(au, 0.0);
        }
    }
    wave_sync();
    wcopy(R, rho + k * nn, nn);
    wmm(T, R, U, N, false, false);
    wmm(W, U, T, N, true, false);
    for (int e = lane; e < nn; e += 64) out[k * nn + e] = W[e];
}

// One time step of the velocity gauge for every k: the sample at t_n
// (SBEs.h:694-760 with _tkp = k: rho is band-basis and U^dag J U is pMatrix;
// the Nval A(t) term is added by the host), then the RK step with
//   f(y) = -i((H y) - (H y)^dag) - (y - rho0) o Gamma,  H = sum_a A_a P_a + diag(E)
// (SBEs.h:596-624, 787-805, 814-829, 862-871), in the reference's order of
// operations. Each wave walks k = wave, wave + G, ...: rho of its k point in
// registers (EPL elements per lane), the stage input, H, H y and the stage
// derivatives in LDS, products on FP64 MFMA (wmm). One row of partial sums
// per wave: [inter(3) intra(3) pop(N) rho(2 N^2)].
template <int EPL>
__global__ __launch_bounds__(64) void nb_vel_step_kernel(const NbArgs a, const NbVelStep v, const cd *__restrict__ P,
                                                         const double *__restrict__ Eocc, cd *__restrict__ rho,
                                                         double *__restrict__ partials, int64_t step_local, int p_lds)
{
    extern __shared__ double4 smem4[];
    const int N = a.N, nn = N * N, S = v.nstage;
    const int lane = threadIdx.x;
    const int64_t ks = a.nk_stride * nn;
    cd *base = reinterpret_cast<cd *>(smem4);
    cd *Y = mat_slot(base, 0, nn), *M = mat_slot(base, 1, nn), *T = mat_slot(base, 2, nn);
    cd *K = mat_slot(base, 3, nn);       // S stage derivatives
    cd *PL = mat_slot(base, 3 + S, nn);  // pMatrix_{x,y,z} when p_lds
    double *EL = reinterpret_cast<double *>(mat_slot(base, 3 + S + (p_lds ? 3 : 0), nn));  // E[N], occ[N]
    double racc[2 * EPL];
#pragma unroll
    for (int t = 0; t < 2 * EPL; ++t) racc[t] = 0.0;
    double cur[6] = {0, 0, 0, 0, 0, 0};
    double pop = 0.0;
    for (int64_t k = blockIdx.x; k < a.nk; k += gridDim.x) {
        const double w = a.weight[k];
        cd r[EPL];
#pragma unroll
        for (int t = 0; t < EPL; ++t) {
            const int e = lane + 64 * t;
            r[t] = e < nn ? rho[k * nn + e] : cmk(0.0, 0.0);
            if (e < nn) Y[e] = r[t];
        }
        if (lane < 2 * N) EL[lane] = Eocc[k * 2 * N + lane];
        const cd *Pk[3];
        for (int ax = 0; ax < 3; ++ax) {
            const cd *g = P + ax * ks + k * nn;
            if (p_lds) {
                for (int e = lane; e < nn; e += 64) PL[ax * nn + e] = g[e];
                Pk[ax] = PL + ax * nn;
            } else {
                Pk[ax] = g;
            }
        }
        wave_sync();
        // ---- sample at t_n
        if (lane < N) pop += Y[lane * N + lane].x * w;
#pragma unroll
        for (int t = 0; t < EPL; ++t) {
            racc[2 * t] += r[t].x * w;
            racc[2 * t + 1] += r[t].y * w;
        }
        for (int ax = 0; ax < 3; ++ax) {
            double inter = 0.0, intra = 0.0;
            for (int e = lane; e < nn; e += 64) {
                const int i = e / N, j = e % N;
                const cd y = Y[e], p = Pk[ax][j * N + i];
                const double val = y.x * p.x - y.y * p.y;
                if (i == j && a.inter_intra) intra += val;
                else inter += val;
            }
            cur[ax] += wave_sum(inter) * w;
            cur[3 + ax] += wave_sum(intra) * w;
        }
        // ---- RK stages
        for (int s = 0; s < S; ++s) {
            if (s > 0) {
#pragma unroll
                for (int t = 0; t < EPL; ++t) {
                    const int e = lane + 64 * t;
                    if (e >= nn) continue;
                    cd y = r[t];
                    for (int j = 0; j < s; ++j)
                        if (v.a[s][j] != 0.0) y = cadd(y, cscale(K[j * nn + e], v.a[s][j]));
                    Y[e] = y;
                }
            }
            const double A0 = v.A[s][0], A1 = v.A[s][1], A2 = v.A[s][2];
            for (int e = lane; e < nn; e += 64) {
                cd m = cscale(Pk[0][e], A0);
                m = cadd(m, cscale(Pk[1][e], A1));
                m = cadd(m, cscale(Pk[2][e], A2));
                const int i = e / N, j = e % N;
                if (i == j) m.x += EL[i];
                M[e] = m;
            }
            wave_sync();
            wmm(T, M, Y, N, false, false);  // H y
            for (int e = lane; e < nn; e += 64) {
                const int i = e / N, j = e % N;
                const cd d = csub(T[i * N + j], cconj(T[j * N + i]));
                cd x = cmk(d.y, -d.x);  // -i (H y - (H y)^dag)
                if (a.deph) {
                    const double g = i == j ? a.gam1 : a.gam2;
                    const cd dy = i == j ? cmk(Y[e].x - EL[N + i], Y[e].y) : Y[e];
                    x = csub(x, cscale(dy, g));
                }
                K[s * nn + e] = x;
            }
            wave_sync();
        }
#pragma unroll
        for (int t = 0; t < EPL; ++t) {
            const int e = lane + 64 * t;
            if (e >= nn) continue;
            cd y = r[t];
            for (int j = 0; j < S; ++j)
                if (v.b[j] != 0.0) y = cadd(y, cscale(K[j * nn + e], v.b[j]));
            rho[k * nn + e] = y;
        }
        wave_sync();
    }
    double *row = partials + (step_local * gridDim.x + blockIdx.x) * static_cast<int64_t>(a.nv);
    if (lane < 6) row[lane] = cur[lane];
    if (lane < N) row[6 + lane] = pop;
#pragma unroll
    for (int t = 0; t < EPL; ++t) {
        const int e = lane + 64 * t;
        if (e < nn) {
            row[6 + N + 2 * e] = racc[2 * t];
            row[6 + N + 2 * e + 1] = racc[2 * t + 1];
        }
    }
}

}  // namespace

size_t nb_lds_bytes(int N)
{
    const size_t nn = static_cast<size_t>(N) * N;
    return 6 * nn * sizeof(double2) + 64 * sizeof(double) + 64 * sizeof(double4);
}

hipError_t nb_launch_phase(const NbArgs &a, double2 *P, hipStream_t s)
{
    const int64_t total = a.nk_stride * a.nr;
    nb_phase_kernel<<<dim3(static_cast<unsigned>((total + 255) / 256)), 256, 0, s>>>(a, P);
    return hipGetLastError();
}

hipError_t nb_launch_psplit(const NbArgs &a, const double2 *P, uint16_t *p16, hipStream_t s)
{
    const int64_t total = a.nk_stride * a.nr_pad;
    nb_psplit_kernel<<<dim3(static_cast<unsigned>((total + 255) / 256)), 256, 0, s>>>(a, P, p16);
    return hipGetLastError();
}

// isDipoleZero (Wannier90.h:180-194,297-387): element (m, n) of each sum carries
// e^{ik'.(R + tau_n - tau_m)} with k' = k + A, i.e. the sum over R alone times
// e^{ik'.(tau_n - tau_m)}; J_a = -i sum (R + tau_n - tau_m)_a H e^{...} gains
// -i (tau_n - tau_m)_a H. One thread per (k, element) for every matrix of the
// evaluation: first the H and H~ (no E.D term in this form), then the J with
// the phased H of this evaluation, or hsrc (the slot before the J) if the
// evaluation has none.
__global__ void nb_tau_kernel(const NbArgs a, const NbEval ev, double2 *out, const double2 *hsrc)
{
    const int nn = a.N * a.N;
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= a.nk_stride * nn) return;
    const int64_t k = idx / nn, slot = a.nk_stride * nn;
    const int ij = static_cast<int>(idx % nn), m = ij / a.N, n = ij % a.N;
    double dt[3], ph = 0.0;
    for (int c = 0; c < 3; ++c) {
        dt[c] = a.tau[n][c] - a.tau[m][c];
        ph += (a.kpts[3 * k + c] + ev.A[c]) * dt[c];
    }
    double sn, cs;
    sincos(ph, &sn, &cs);
    const cd e = cmk(cs, sn);
    int jh = -1;
    for (int j = 0; j < ev.nmat; ++j)
        if (ev.mats[j] == kMatHtilde || ev.mats[j] == kMatH0) {
            out[j * slot + idx] = cmul(e, out[j * slot + idx]);
            if (ev.mats[j] == kMatH0 || jh < 0) jh = j;
        }
    const cd h = jh >= 0 ? out[jh * slot + idx] : hsrc[idx];
    for (int j = 0; j < ev.nmat; ++j)
        if (ev.mats[j] >= kMatJx) {
            const double d = dt[ev.mats[j] - kMatJx];
            const cd x = cmul(e, out[j * slot + idx]);
            out[j * slot + idx] = cmk(x.x + d * h.y, x.y - d * h.x);  // x - i d h
        }
}

static hipError_t nb_launch_sums(const NbArgs &a, const NbEval &ev, const double2 *P, double2 *out, hipStream_t s);

hipError_t nb_launch_fourier(const NbArgs &a, const NbEval &ev, const double2 *P, double2 *out, hipStream_t s)
{
    hipError_t e = nb_launch_sums(a, ev, P, out, s);
    if (e != hipSuccess || !a.dzero) return e;
    const int64_t total = a.nk_stride * a.N * a.N;
    bool has_h = false, has_j = false;
    for (int j = 0; j < ev.nmat; ++j) {
        has_h = has_h || ev.mats[j] < kMatJx;
        has_j = has_j || ev.mats[j] >= kMatJx;
    }
    // evaluations with J but no H follow the H slot (the sample's J after a reused H)
    const double2 *hsrc = has_j && !has_h ? out - a.nk_stride * a.N * a.N : out;
    nb_tau_kernel<<<dim3(static_cast<unsigned>((total + 255) / 256)), 256, 0, s>>>(a, ev, out, hsrc);
    return hipGetLastError();
}

static hipError_t nb_launch_sums(const NbArgs &a, const NbEval &ev, const double2 *P, double2 *out, hipStream_t s)
{
    const int64_t ncols = static_cast<int64_t>(ev.nmat) * a.N * a.N;
    if (a.split16) {
        NbScale sc{};
        const double emag = std::fabs(ev.E[0]) + std::fabs(ev.E[1]) + std::fabs(ev.E[2]);
        for (int m = 0; m < 5; ++m) {
            double bound = a.hmax;  // H0
            if (m < ev.nmat && ev.mats[m] == kMatHtilde) bound = a.hmax + emag * a.pmax;
            if (m < ev.nmat && ev.mats[m] >= kMatJx) bound = a.rmax * a.hmax;
            sc.s[m] = bound > 0 ? std::ldexp(1.0, -static_cast<int>(std::ceil(std::log2(bound)))) : 1.0;
        }
        const int64_t total = a.bcols * a.nr_pad;
        nb_bsplit_kernel<<<dim3(static_cast<unsigned>((total + 255) / 256)), 256, 0, s>>>(a, ev, sc, a.b16);
        const dim3 grid(static_cast<unsigned>(a.nk_stride / 64), static_cast<unsigned>((ncols + 63) / 64));
        nb_fourier16_kernel<<<grid, 256, 0, s>>>(a, ev, sc, a.p16, a.b16, out);
        return hipGetLastError();
    }
    if (a.box) return box_fourier(a, ev, out, s);
    const int64_t nb = static_cast<int64_t>(a.nr_pad) * a.bcols;
    nb_bform_kernel<<<dim3(static_cast<unsigned>((nb + 255) / 256)), 256, 0, s>>>(a, ev, a.b64);
    const dim3 grid(static_cast<unsigned>((a.nk_stride + GM - 1) / GM), static_cast<unsigned>((ncols + GN - 1) / GN));
    nb_fourier_kernel<<<grid, 256, 0, s>>>(a, ev, P, a.b64, out);
    return hipGetLastError();
}

hipError_t nb_launch_box_t2(const NbArgs &a, const NbEval &ev, double2 *t2, hipStream_t s)
{
    if (!a.box || a.split16 || a.dzero) return hipErrorInvalidValue;
    return box_t2(a, ev, t2, s);
}

template <int MM>
static void box_pass_multi(const NbArgs &a, const BoxMulti &bm, int pass, int64_t maxcols, hipStream_t s)
{
    const int M1 = a.boxM[0], M2 = a.boxM[1], K1 = a.boxK[0], K2 = a.boxK[1], K3 = a.boxK[2];
    const int K = pass == 0 ? K3 : pass == 1 ? K2 : K1;
    const int64_t outer = pass == 0 ? static_cast<int64_t>(M1) * M2 : pass == 1 ? M1 : 1;
    const int64_t inner = pass == 0 ? 1 : pass == 1 ? K3 : static_cast<int64_t>(K2) * K3;
    const int64_t xy = outer * inner * ((maxcols + 255) / 256) * bm.nev;
    const int kc = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(K, K * xy / 2048)));
    const int nz = (K + kc - 1) / kc;
    const dim3 grid(static_cast<unsigned>(outer * inner), static_cast<unsigned>((maxcols + 255) / 256),
                    static_cast<unsigned>(nz * bm.nev));
    if (pass == 0) nb_box_pass_multi_kernel<MM, 0><<<grid, 256, 0, s>>>(a, bm, kc, nz);
    else if (pass == 1) nb_box_pass_multi_kernel<MM, 1><<<grid, 256, 0, s>>>(a, bm, kc, nz);
    else nb_box_a1_multi_kernel<MM><<<grid, 256, 0, s>>>(a, bm, kc, nz);
}

static void box_pass_multi_m(const NbArgs &a, const BoxMulti &bm, int pass, int64_t maxcols, hipStream_t s)
{
    const int M = a.boxM[2 - pass];
    if (M <= 4) box_pass_multi<4>(a, bm, pass, maxcols, s);
    else if (M <= 8) box_pass_multi<8>(a, bm, pass, maxcols, s);
    else if (M <= 16) box_pass_multi<16>(a, bm, pass, maxcols, s);
    else box_pass_multi<32>(a, bm, pass, maxcols, s);
}

// B and the passes of up to four evaluations; out == nullptr: the passes along a3
// and a2 into t2, otherwise the one along a1 too, into out[e]
static hipError_t box_multi(const NbArgs &a, int nev, const NbEval *evs, double2 *const *b, double2 *const *t1,
                            double2 *const *t2, double2 *const *out, hipStream_t s)
{
    if (!a.box || a.split16 || a.dzero || nev < 1 || nev > 4) return hipErrorInvalidValue;
    BoxMulti bm{};
    bm.nev = nev;
    int64_t maxcols = 0;
    for (int e = 0; e < nev; ++e) {
        bm.ev[e] = evs[e];
        bm.b[e] = b[e];
        bm.t1[e] = t1[e];
        bm.t2[e] = t2[e];
        bm.out[e] = out ? out[e] : nullptr;
        maxcols = std::max<int64_t>(maxcols, static_cast<int64_t>(evs[e].nmat) * a.N * a.N);
    }
    const int64_t ncells = static_cast<int64_t>(a.boxM[0]) * a.boxM[1] * a.boxM[2];
    nb_box_bform_multi_kernel<<<dim3(static_cast<unsigned>((ncells * maxcols + 255) / 256), 1, static_cast<unsigned>(nev)),
                                256, 0, s>>>(a, bm, ncells);
    box_pass_multi_m(a, bm, 0, maxcols, s);
    box_pass_multi_m(a, bm, 1, maxcols, s);
    if (out) box_pass_multi_m(a, bm, 2, maxcols, s);
    return hipGetLastError();
}

hipError_t nb_launch_box_t2_multi(const NbArgs &a, int nev, const NbEval *evs, double2 *const *b, double2 *const *t1,
                                  double2 *const *t2, hipStream_t s)
{
    return box_multi(a, nev, evs, b, t1, t2, nullptr, s);
}

// Up to four evaluations of the dense sums in one launch, for R lists of at most MM
// vectors: a thread owns one column c of one evaluation, forms its B column
// w_R e^{iA.R} X_R[c] in registers (e^{iA.R} once per block in LDS) and sums
// out[k][c] = sum_R e^{ik.R} B_R[c] over a chunk of kc k points (e^{ik.R} read
// with uniform loads). block y: 256 columns, block z: (evaluation, k chunk).
struct DenseMulti {
    NbEval ev[4];
    int nev;
    cd *out[4];
};

// one evaluation's columns [blockIdx.y * 256, + 256) at k points [k0, k1) into out
// ([nmat][nk_stride][N][N]); phase: LDS for MM values
template <int MM>
__device__ __forceinline__ void dense_eval_body(const NbArgs &a, const NbEval &ev, const cd *__restrict__ P,
                                                cd *__restrict__ out_base, int64_t k0, int64_t k1, cd *phase)
{
    const int nr = a.nr, nn = a.N * a.N;
    if (static_cast<int>(threadIdx.x) < nr) {
        const double *R = a.rcart + 3 * threadIdx.x;
        double sn, cs;
        sincos(ev.A[0] * R[0] + ev.A[1] * R[1] + ev.A[2] * R[2], &sn, &cs);
        phase[threadIdx.x] = cmk(cs, sn);
    }
    __syncthreads();
    const int64_t ncols = static_cast<int64_t>(ev.nmat) * nn;
    const int64_t c = static_cast<int64_t>(blockIdx.y) * 256 + threadIdx.x;
    if (c >= ncols) return;  // no barriers below
    cd x[MM];
#pragma unroll
    for (int m = 0; m < MM; ++m) x[m] = m < nr ? b_element(a, ev, m, c, phase[m]) : cmk(0.0, 0.0);
    const int mat = static_cast<int>(c / nn), ij = static_cast<int>(c - static_cast<int64_t>(mat) * nn);
    cd *out = out_base + static_cast<int64_t>(mat) * a.nk_stride * nn + ij;
    for (int64_t k = k0; k < k1; ++k) {
        const cd *pk = P + k * nr;
        double re = 0.0, im = 0.0;
#pragma unroll
        for (int m = 0; m < MM; ++m) {
            if (m < nr) {
                const cd w = pk[m];
                re = fma(w.x, x[m].x, fma(-w.y, x[m].y, re));
                im = fma(w.x, x[m].y, fma(w.y, x[m].x, im));
            }
        }
        out[k * nn] = cmk(re, im);
    }
}

template <int MM>
__global__ __launch_bounds__(256) void nb_dense_multi_kernel(const NbArgs a, const DenseMulti dm,
                                                             const cd *__restrict__ P, int kc, int nz)
{
    __shared__ cd phase[MM];
    const int e = static_cast<int>(blockIdx.z) / nz, kz = static_cast<int>(blockIdx.z) - e * nz;
    const int64_t k0 = static_cast<int64_t>(kz) * kc;
    dense_eval_body<MM>(a, dm.ev[e], P, dm.out[e], k0, min(a.nk, k0 + kc), phase);
}

// The matrices of a run of RK4 steps in one launch: block z = (step, evaluation): J at
// t_n, H~ (and H with dephasing) at t_n + dt/2, H~ and H at t_n + dt, with A and E at
// those times from the host-evaluated table; block y = 64 columns (one per lane, the B
// column w_R e^{iA.R} X_R[c] formed in registers). The k points go in tiles of 64 whose
// e^{ik.R} rows are staged in LDS and read as broadcasts; wave w sums the tile's k points
// w, w + 4, ... two at a time (four independent accumulation chains).
constexpr int kDenseKTile = 64;
template <int MM>
__global__ __launch_bounds__(256) void nb_dense_steps_kernel(const NbArgs a, const NbDenseSteps d,
                                                             const cd *__restrict__ P)
{
    __shared__ cd phase[MM];
    __shared__ cd pt[kDenseKTile * MM];
    const int z = static_cast<int>(blockIdx.z), sl = z / 3, typ = z - 3 * sl;
    const int64_t step = d.s0 + sl;
    NbEval ev{};
    const double *lt = d.laser + (step * 3 + typ) * 6;
    for (int i = 0; i < 3; ++i) {
        ev.A[i] = lt[i];
        ev.E[i] = lt[3 + i];
    }
    cd *out;
    if (typ == 0) {
        ev.nmat = 3;
        ev.mats[0] = kMatJx;
        ev.mats[1] = kMatJy;
        ev.mats[2] = kMatJz;
        out = d.j + static_cast<int64_t>(sl) * 3 * d.slot;
    } else {
        ev.nmat = typ == 1 && !d.deph ? 1 : 2;
        ev.mats[0] = kMatHtilde;
        ev.mats[1] = kMatH0;
        out = typ == 1 ? d.half + static_cast<int64_t>(sl) * 2 * d.slot : d.ring + (step % d.nring) * 2 * d.slot;
    }
    const int nr = a.nr, nn = a.N * a.N;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (static_cast<int>(threadIdx.x) < nr) {
        const double *R = a.rcart + 3 * threadIdx.x;
        double sn, cs;
        sincos(ev.A[0] * R[0] + ev.A[1] * R[1] + ev.A[2] * R[2], &sn, &cs);
        phase[threadIdx.x] = cmk(cs, sn);
    }
    __syncthreads();
    const int64_t ncols = static_cast<int64_t>(ev.nmat) * nn;
    const int64_t c = static_cast<int64_t>(blockIdx.y) * 64 + lane;
    const bool live = c < ncols;
    const int mat = live ? static_cast<int>(c / nn) : 0;
    const int ij = live ? static_cast<int>(c - static_cast<int64_t>(mat) * nn) : 0;
    // the kind without indexing the evaluation's list (J_x, J_y, J_z; H~, H)
    const int type = typ == 0 ? kMatJx + mat : (mat == 0 ? kMatHtilde : kMatH0);
    cd x[MM];
#pragma unroll
    for (int m = 0; m < MM; ++m) x[m] = live && m < nr ? b_value(a, ev.E, type, m, ij, phase[m]) : cmk(0.0, 0.0);
    cd *o = out + static_cast<int64_t>(mat) * a.nk_stride * nn + ij;
    for (int64_t k0 = 0; k0 < a.nk; k0 += kDenseKTile) {
        const int kt = static_cast<int>(min<int64_t>(kDenseKTile, a.nk - k0));
        __syncthreads();  // the previous tile's reads are done
        for (int e = threadIdx.x; e < kt * nr; e += 256) pt[(e / nr) * MM + e % nr] = P[k0 * nr + e];
        __syncthreads();
        for (int kk = w; kk < kt; kk += 8) {
            const int kk2 = kk + 4 < kt ? kk + 4 : kk;  // a second k point (the same one at a ragged end)
            const cd *p1 = pt + kk * MM, *p2 = pt + kk2 * MM;
            double r1a = 0.0, i1a = 0.0, r2a = 0.0, i2a = 0.0;
#pragma unroll
            for (int m = 0; m < MM; ++m) {
                if (m < nr) {
                    const cd u = p1[m], v = p2[m];
                    r1a = fma(u.x, x[m].x, fma(-u.y, x[m].y, r1a));
                    i1a = fma(u.x, x[m].y, fma(u.y, x[m].x, i1a));
                    r2a = fma(v.x, x[m].x, fma(-v.y, x[m].y, r2a));
                    i2a = fma(v.x, x[m].y, fma(v.y, x[m].x, i2a));
                }
            }
            if (live) {
                o[(k0 + kk) * nn] = cmk(r1a, i1a);
                if (kk2 != kk) o[(k0 + kk2) * nn] = cmk(r2a, i2a);
            }
        }
    }
}

hipError_t nb_launch_dense_steps(const NbArgs &a, const NbDenseSteps &d, const double2 *P, hipStream_t s)
{
    if (a.model_kind || a.split16 || a.dzero || a.nr < 1 || a.nr > kNbDenseMultiMaxR || !P || d.ns < 1 ||
        d.nring < 1 || d.s0 < 0)
        return hipErrorInvalidValue;
    const dim3 grid(1, static_cast<unsigned>((3 * a.N * a.N + 63) / 64), static_cast<unsigned>(3 * d.ns));
    if (a.nr <= 16) nb_dense_steps_kernel<16><<<grid, 256, 0, s>>>(a, d, P);
    else nb_dense_steps_kernel<32><<<grid, 256, 0, s>>>(a, d, P);
    return hipGetLastError();
}

template <int MM>
static void dense_multi(const NbArgs &a, const DenseMulti &dm, const cd *P, int64_t maxcols, hipStream_t s)
{
    // k chunks: about two blocks per CU in all (each block forms its columns of B once)
    const int64_t ny = (maxcols + 255) / 256;
    const int64_t nz = std::max<int64_t>(1, std::min<int64_t>(a.nk, 512 / (ny * dm.nev)));
    const int kc = static_cast<int>((a.nk + nz - 1) / nz);
    const int nzc = static_cast<int>((a.nk + kc - 1) / kc);
    const dim3 grid(1, static_cast<unsigned>(ny), static_cast<unsigned>(nzc * dm.nev));
    nb_dense_multi_kernel<MM><<<grid, 256, 0, s>>>(a, dm, P, kc, nzc);
}

hipError_t nb_launch_dense_multi(const NbArgs &a, int nev, const NbEval *evs, const double2 *P, double2 *const *out,
                                 hipStream_t s)
{
    if (a.model_kind || a.split16 || a.dzero || a.nr < 1 || a.nr > kNbDenseMultiMaxR || !P || nev < 1 || nev > 4)
        return hipErrorInvalidValue;
    DenseMulti dm{};
    dm.nev = nev;
    int64_t maxcols = 0;
    for (int e = 0; e < nev; ++e) {
        if (!out[e]) return hipErrorInvalidValue;
        dm.ev[e] = evs[e];
        dm.out[e] = out[e];
        maxcols = std::max<int64_t>(maxcols, static_cast<int64_t>(evs[e].nmat) * a.N * a.N);
    }
    if (a.nr <= 16) dense_multi<16>(a, dm, P, maxcols, s);
    else dense_multi<32>(a, dm, P, maxcols, s);
    return hipGetLastError();
}

template <int MM, int PASS>
static void box_steps_pass(const NbArgs &a, const NbBoxSteps &d, int64_t maxcols, hipStream_t s)
{
    const int M1 = a.boxM[0], M2 = a.boxM[1], K2 = a.boxK[1], K3 = a.boxK[2];
    const int K = PASS == 0 ? K3 : K2;
    const int64_t outer = PASS == 0 ? static_cast<int64_t>(M1) * M2 : M1, inner = PASS == 0 ? 1 : K3;
    const int64_t nev = 3 * static_cast<int64_t>(d.ns);
    const int64_t xy = outer * inner * ((maxcols + 255) / 256) * nev;
    const int kc = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(K, K * xy / 2048)));
    const int nz = (K + kc - 1) / kc;
    const dim3 grid(static_cast<unsigned>(outer * inner), static_cast<unsigned>((maxcols + 255) / 256),
                    static_cast<unsigned>(nz * nev));
    nb_box_steps_pass_kernel<MM, PASS><<<grid, 256, 0, s>>>(a, d, kc, nz);
}

template <int PASS>
static void box_steps_pass_m(const NbArgs &a, const NbBoxSteps &d, int64_t maxcols, hipStream_t s)
{
    const int M = a.boxM[2 - PASS];
    if (M <= 4) box_steps_pass<4, PASS>(a, d, maxcols, s);
    else if (M <= 8) box_steps_pass<8, PASS>(a, d, maxcols, s);
    else if (M <= 16) box_steps_pass<16, PASS>(a, d, maxcols, s);
    else box_steps_pass<32, PASS>(a, d, maxcols, s);
}

hipError_t nb_launch_box_steps(const NbArgs &a, const NbBoxSteps &d, hipStream_t s)
{
    if (!a.box || a.split16 || a.dzero || d.ns < 1 || d.nring < 1 || d.s0 < 0 || 3 * static_cast<int64_t>(d.ns) > 65535)
        return hipErrorInvalidValue;
    const int64_t maxcols = 3 * static_cast<int64_t>(a.N) * a.N;
    const int64_t ncells = static_cast<int64_t>(a.boxM[0]) * a.boxM[1] * a.boxM[2];
    nb_box_steps_bform_kernel<<<dim3(static_cast<unsigned>((ncells * maxcols + 255) / 256), 1,
                                     static_cast<unsigned>(3 * d.ns)),
                                256, 0, s>>>(a, d, ncells);
    box_steps_pass_m<0>(a, d, maxcols, s);
    box_steps_pass_m<1>(a, d, maxcols, s);
    return hipGetLastError();
}

hipError_t nb_launch_box_mats_multi(const NbArgs &a, int nev, const NbEval *evs, double2 *const *b, double2 *const *t1,
                                    double2 *const *t2, double2 *const *out, hipStream_t s)
{
    if (!out) return hipErrorInvalidValue;
    for (int e = 0; e < nev && e < 4; ++e)
        if (!out[e]) return hipErrorInvalidValue;
    return box_multi(a, nev, evs, b, t1, t2, out, s);
}

hipError_t nb_launch_init(const NbArgs &a, const double2 *H0, double2 *rho, double2 *U, hipStream_t s)
{
    nb_init_kernel<<<dim3(static_cast<unsigned>(a.nk_stride)), 64, nb_lds_bytes(a.N), s>>>(a, H0, rho, U);
    return hipGetLastError();
}

hipError_t nb_launch_sample(const NbArgs &a, const double2 *mats, const double2 *rho, double *partials,
                            int64_t step_local, double2 *U, int ureuse, hipStream_t s)
{
    if (a.tile16) return nb16_launch_sample(a, mats, rho, partials, step_local, U, ureuse, s);
    // N > 16: four waves per k point (the products' output tiles and the element loops split
    // among them)
    const unsigned threads = a.N > 16 ? 256u : 64u;
    nb_sample_kernel<<<dim3(static_cast<unsigned>(a.sample_waves)), threads, nb_lds_bytes(a.N), s>>>(a, mats, rho, partials,
                                                                                               step_local, U, ureuse);
    return hipGetLastError();
}

hipError_t nb_launch_rhs(const NbArgs &a, const NbStage &sg, const double2 *mats, double2 *rho, double2 *kbuf,
                         double2 *U, hipStream_t s)
{
    if (a.tile16) return nb16_launch_rhs(a, sg, mats, rho, kbuf, U, s);
    const unsigned threads = a.N > 16 ? 256u : 64u;
    nb_rhs_kernel<<<dim3(static_cast<unsigned>(a.nk)), threads, nb_lds_bytes(a.N), s>>>(a, sg, mats, rho, kbuf, U);
    return hipGetLastError();
}

size_t nb_vel_lds_bytes(int N, int nstage, bool p_lds)
{
    const size_t nn = static_cast<size_t>(N) * N;
    return (3 + nstage + (p_lds ? 3 : 0)) * nn * sizeof(double2) + 2 * N * sizeof(double);
}

hipError_t nb_launch_vel_init(const NbArgs &a, const double2 *mats, double2 *P, double *Eocc, double2 *U,
                              double2 *rho, hipStream_t s)
{
    nb_vel_init_kernel<<<dim3(static_cast<unsigned>(a.nk_stride)), 64, nb_lds_bytes(a.N), s>>>(a, mats, P, Eocc, U, rho);
    return hipGetLastError();
}

hipError_t nb_launch_vel_step(const NbArgs &a, const NbVelStep &v, const double2 *P, const double *Eocc, double2 *rho,
                              double *partials, int64_t step_local, bool p_lds, hipStream_t s)
{
    if (a.tile16 && (v.nstage == 4 || v.nstage == 6))
        return nb16_launch_vel_step(a, v, P, Eocc, rho, partials, step_local, s);
    const dim3 grid(static_cast<unsigned>(a.sample_waves));
    const size_t lds = nb_vel_lds_bytes(a.N, v.nstage, p_lds);
    const int nn = a.N * a.N;
    const int pl = p_lds ? 1 : 0;
    if (nn <= 64) nb_vel_step_kernel<1><<<grid, 64, lds, s>>>(a, v, P, Eocc, rho, partials, step_local, pl);
    else if (nn <= 256) nb_vel_step_kernel<4><<<grid, 64, lds, s>>>(a, v, P, Eocc, rho, partials, step_local, pl);
    else if (nn <= 1024) nb_vel_step_kernel<16><<<grid, 64, lds, s>>>(a, v, P, Eocc, rho, partials, step_local, pl);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t nb_launch_band_export(const NbArgs &a, const double2 *H, const double2 *rho, double2 *out, hipStream_t s)
{
    nb_band_export_kernel<<<dim3(static_cast<unsigned>(a.nk_stride)), 64, nb_lds_bytes(a.N), s>>>(a, H, rho, out);
    return hipGetLastError();
}

}  // namespace sbe
