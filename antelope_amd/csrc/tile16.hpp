// tile16.hpp - complex 16x16 FP64 matrices as v_mfma_f64_16x16x4f64 register
// tiles, shared by the N-band kernels (nband16.hip, nband32.hip).
//
// A Tile: lane l = c + 16 g holds the elements (g + 4r, c), r = 0..3 - the D
// layout of v_mfma_f64_16x16x4f64. C = A B runs as four real MFMAs per K block
// of 4; the B operand of block kb is register kb of B's tile and the A operand
// register kb of A^T's tile (for Hermitian A and for A = X^dag, the conjugate of
// A's own tile).
#pragma once

#include <hip/hip_runtime.h>

namespace sbe {
namespace tile {

using cd = double2;
using d4 = double __attribute__((ext_vector_type(4)));

struct Tile {
    double re[4], im[4];
};

constexpr int kTpad = 17;  // LDS row pitch (complex) of the transpose buffer

__device__ __forceinline__ int lane_c() { return threadIdx.x & 15; }
__device__ __forceinline__ int lane_g() { return (threadIdx.x >> 4) & 3; }

// C = A B. Aop holds A^T's tile (register kb = A[c][4 kb + g]); CA and CB
// conjugate the operands. The signs go into the MFMA's operand negation (its
// last argument: 1 negates A, 2 negates B), not into separate instructions.
template <bool CA, bool CB>
__device__ __forceinline__ Tile cmm_t(const Tile &Aop, const Tile &B)
{
    // re += ar br - ai' bi',  im += ar bi' + ai' br  with ai' = (CA ? -1 : 1) ai, bi' likewise
    constexpr int kNegRe2 = (CA != CB) ? 0 : 1;  // -ai' bi' = -(+-1)(+-1) ai bi
    constexpr int kNegIm1 = CB ? 2 : 0;
    constexpr int kNegIm2 = CA ? 1 : 0;
    d4 re = d4{0.0, 0.0, 0.0, 0.0}, im = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
        const double ar = Aop.re[kb], ai = Aop.im[kb], br = B.re[kb], bi = B.im[kb];
        re = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, re, 0, 0, 0);
        re = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, bi, re, 0, 0, kNegRe2);
        im = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bi, im, 0, 0, kNegIm1);
        im = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, br, im, 0, 0, kNegIm2);
    }
    Tile C;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        C.re[r] = re[r];
        C.im[r] = im[r];
    }
    return C;
}

__device__ __forceinline__ Tile cmm(const Tile &Aop, bool conjA, const Tile &B, bool conjB = false)
{
    if (conjA) return conjB ? cmm_t<true, true>(Aop, B) : cmm_t<true, false>(Aop, B);
    return conjB ? cmm_t<false, true>(Aop, B) : cmm_t<false, false>(Aop, B);
}

__device__ __forceinline__ void wave_lds_order()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// X^T through the wave's LDS buffer (16 x kTpad complex)
__device__ __forceinline__ Tile transpose(const Tile &X, cd *buf)
{
    const int c = lane_c(), g = lane_g();
#pragma unroll
    for (int r = 0; r < 4; ++r) buf[(g + 4 * r) * kTpad + c] = make_double2(X.re[r], X.im[r]);
    wave_lds_order();
    Tile T;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const cd v = buf[c * kTpad + g + 4 * r];
        T.re[r] = v.x;
        T.im[r] = v.y;
    }
    wave_lds_order();  // the next transpose overwrites the buffer
    return T;
}

// N x N row-major matrix -> tile (zero padding; pad_one: identity padding)
__device__ __forceinline__ Tile load_tile(const cd *M, int N, bool pad_one = false)
{
    const int c = lane_c(), g = lane_g();
    Tile T;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = g + 4 * r;
        cd v = make_double2(pad_one && i == c ? 1.0 : 0.0, 0.0);
        if (i < N && c < N) v = M[i * N + c];
        T.re[r] = v.x;
        T.im[r] = v.y;
    }
    return T;
}

// the transpose of an N x N row-major matrix -> tile (zero padding)
__device__ __forceinline__ Tile load_tile_t(const cd *M, int N)
{
    const int c = lane_c(), g = lane_g();
    Tile T;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = g + 4 * r;
        cd v = make_double2(0.0, 0.0);
        if (i < N && c < N) v = M[c * N + i];
        T.re[r] = v.x;
        T.im[r] = v.y;
    }
    return T;
}

__device__ __forceinline__ void store_tile(cd *M, const Tile &T, int N)
{
    const int c = lane_c(), g = lane_g();
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = g + 4 * r;
        if (i < N && c < N) M[i * N + c] = make_double2(T.re[r], T.im[r]);
    }
}

// x and the value of lane l ^ 16 / l ^ 32 (gfx950 permlane swaps: no LDS)
__device__ __forceinline__ void pair16(double x, double &a, double &b)
{
    const auto lo = __builtin_amdgcn_permlane16_swap(__double2loint(x), __double2loint(x), false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap(__double2hiint(x), __double2hiint(x), false, false);
    a = __hiloint2double(hi[0], lo[0]);  // even rows: this row's own value in rows 0, 2
    b = __hiloint2double(hi[1], lo[1]);
}
__device__ __forceinline__ void pair32(double x, double &a, double &b)
{
    const auto lo = __builtin_amdgcn_permlane32_swap(__double2loint(x), __double2loint(x), false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap(__double2hiint(x), __double2hiint(x), false, false);
    a = __hiloint2double(hi[0], lo[0]);
    b = __hiloint2double(hi[1], lo[1]);
}
// x + x(l ^ 16) and x + x(l ^ 32), the same value in both lanes of a pair
__device__ __forceinline__ double add16(double x)
{
    double a, b;
    pair16(x, a, b);
    return a + b;
}
__device__ __forceinline__ double add32(double x)
{
    double a, b;
    pair32(x, a, b);
    return a + b;
}

__device__ __forceinline__ double wave_sum(double v)
{
    v = add32(v);
    v = add16(v);
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ double wave_max(double v)
{
    double a, b;
    pair32(v, a, b);
    v = fmax(a, b);
    pair16(v, a, b);
    v = fmax(a, b);
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Re sum_i conj(X_ic) Y_ic for this lane's column c (the diagonal of X^dag Y)
__device__ __forceinline__ double col_dot(const Tile &X, const Tile &Y)
{
    double s = 0.0;
#pragma unroll
    for (int r = 0; r < 4; ++r) s += X.re[r] * Y.re[r] + X.im[r] * Y.im[r];
    return add32(add16(s));
}

// The diagonal of a tile, for this lane's rows (lam_r) and column (lam_c).
__device__ __forceinline__ void diag_bcast(const Tile &W, double lam_r[4], double &lam_c)
{
    const int c = lane_c(), g = lane_g();
    double dv = 0.0;  // the diagonal element this lane holds, if any: (c, c) with c = g + 4r
#pragma unroll
    for (int r = 0; r < 4; ++r)
        if (g + 4 * r == c) dv = W.re[r];
    lam_c = __shfl(dv, c + 16 * (c & 3), 64);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = g + 4 * r;
        lam_r[r] = __shfl(dv, i + 16 * (i & 3), 64);
    }
}

}  // namespace tile
}  // namespace sbe
