// analytic.hpp - the analytic tight-binding classes of Materials/ as device
// (and host) functions: H(k) and the classes' GenJMatrix at one k point.
//
// Each function writes the non-zero elements of row-major N x N complex
// matrices; the caller zeroes them first. The current matrices are written
// as each class's GenJMatrix writes them, including where that differs from
// -dH/dk (the sign conventions and the elements noted per class): the
// propagation must see the reference's J, not a corrected one.
#pragma once

#include <hip/hip_runtime.h>

#include "antelope/sbe_c.h"

namespace sbe {
namespace am {

using cd = double2;

__host__ __device__ inline cd mk(double r, double i) { return make_double2(r, i); }
__host__ __device__ inline cd add(cd a, cd b) { return mk(a.x + b.x, a.y + b.y); }
__host__ __device__ inline cd mul(cd a, cd b) { return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__host__ __device__ inline cd scl(cd a, double s) { return mk(a.x * s, a.y * s); }
__host__ __device__ inline cd cj(cd a) { return mk(a.x, -a.y); }
__host__ __device__ inline cd eix(double t) { return mk(cos(t), sin(t)); }

// Materials/WilsonMass.h:40-116. p = {t, delta, mu, a0, a1}.
// H = B1 sx + B2 sy + B3 sz; B1 = delta sin(a0 kx), B2 = delta sin(a1 ky),
// B3 = -2t (cos a0 kx + cos a1 ky) - mu. J = -dH/dk.
__host__ __device__ inline void wilson_mass(const double *p, const double *k, cd *H, cd *Jx, cd *Jy)
{
    const double t = p[0], dl = p[1], mu = p[2], a0 = p[3], a1 = p[4];
    const double sx = sin(a0 * k[0]), cx = cos(a0 * k[0]), sy = sin(a1 * k[1]), cy = cos(a1 * k[1]);
    if (H) {
        const double b1 = dl * sx, b2 = dl * sy, b3 = -2.0 * t * (cx + cy) - mu;
        H[0] = mk(b3, 0.0);
        H[1] = mk(b1, -b2);
        H[2] = mk(b1, b2);
        H[3] = mk(-b3, 0.0);
    }
    if (Jx) {  // d/dkx: (B1, B2, B3)' = (a0 delta cx, 0, 2 t a0 sx)
        const double d1 = a0 * dl * cx, d3 = 2.0 * t * a0 * sx;
        Jx[0] = mk(-d3, 0.0);
        Jx[1] = mk(-d1, 0.0);
        Jx[2] = mk(-d1, 0.0);
        Jx[3] = mk(d3, 0.0);
        const double e2 = a1 * dl * cy, e3 = 2.0 * t * a1 * sy;  // d/dky: (0, a1 delta cy, 2 t a1 sy)
        Jy[0] = mk(-e3, 0.0);
        Jy[1] = mk(0.0, e2);
        Jy[2] = mk(0.0, -e2);
        Jy[3] = mk(e3, 0.0);
    }
}

// Materials/SquareSBand.h:39-101. p = {t1, M0, a0}.
// H = 2 t1 (cos a0 kx + cos a0 ky) sx + M0 sz. The class's J has only the
// (0,1) element, 2 t1 a0 sin(a0 k_a) (SquareSBand.h:69-78): not Hermitian,
// half of -dH/dk.
__host__ __device__ inline void square_sband(const double *p, const double *k, cd *H, cd *Jx, cd *Jy)
{
    const double t1 = p[0], M0 = p[1], a0 = p[2];
    if (H) {
        const double b1 = 2.0 * t1 * (cos(a0 * k[0]) + cos(a0 * k[1]));
        H[0] = mk(M0, 0.0);
        H[1] = mk(b1, 0.0);
        H[2] = mk(b1, 0.0);
        H[3] = mk(-M0, 0.0);
    }
    if (Jx) {
        Jx[1] = mk(2.0 * t1 * a0 * sin(a0 * k[0]), 0.0);
        Jy[1] = mk(2.0 * t1 * a0 * sin(a0 * k[1]), 0.0);
    }
}

// Honeycomb vectors of Haldane.h:134-143 / KaneMele.h:176-187: NN a_i, NNN b_i.
__host__ __device__ inline void honeycomb(double a0, double va[3][2], double vb[3][2])
{
    const double r3 = sqrt(3.0);
    va[0][0] = 0.0;           va[0][1] = a0;
    va[1][0] = -r3 / 2 * a0;  va[1][1] = -0.5 * a0;
    va[2][0] = r3 / 2 * a0;   va[2][1] = -0.5 * a0;
    vb[0][0] = r3 * a0;       vb[0][1] = 0.0;
    vb[1][0] = -r3 / 2 * a0;  vb[1][1] = 1.5 * a0;
    vb[2][0] = -r3 / 2 * a0;  vb[2][1] = -1.5 * a0;
}

// Materials/KaneMele.h:47-248. p = {t1, t2, tR, M0, phi0, a0}; basis
// (A up, B up, A down, B down) in blocks: Haldane blocks with +phi0 (rows
// 0-1) and -phi0 (rows 2-3); Rashba hopping R0 = i tR sum_i (c_i) e^{i k.a_i},
// R1 = i tR sum_i (c_i) e^{-i k.a_i} with c_i = cx_i - i cy_i
// (cx, cy) = (1, 0), (-1/2, sqrt3/2), (-1/2, -sqrt3/2) at (0,3), (1,2) and
// their conjugates at (2,1), (3,0). J: the Haldane blocks are -dH/dk; the
// Rashba elements are written as KaneMele.h:138-176 does: (0,3), (1,2), (2,1)
// equal +dH/dk and (3,0) is -sum_i conj(c_i) e^{-i k.a_i} (-i a_i) without
// the i tR factor.
__host__ __device__ inline void kane_mele(const double *p, const double *k, cd *H, cd *Jx, cd *Jy)
{
    const double t1 = p[0], t2 = p[1], tR = p[2], M0 = p[3], phi = p[4], a0 = p[5];
    double va[3][2], vb[3][2];
    honeycomb(a0, va, vb);
    const double r3 = sqrt(3.0);
    const double cr[3][2] = {{1.0, 0.0}, {-0.5, r3 / 2}, {-0.5, -r3 / 2}};
    double tha[3], thb[3];
    for (int i = 0; i < 3; ++i) {
        tha[i] = k[0] * va[i][0] + k[1] * va[i][1];
        thb[i] = k[0] * vb[i][0] + k[1] * vb[i][1];
    }
    if (H) {
        double sca = 0, ssa = 0, scb = 0, ssb = 0;
        for (int i = 0; i < 3; ++i) {
            sca += cos(tha[i]);
            ssa += sin(tha[i]);
            scb += cos(thb[i]);
            ssb += sin(thb[i]);
        }
        for (int b = 0; b < 2; ++b) {
            const double ph = b == 0 ? phi : -phi;
            const double B0 = 2.0 * t2 * cos(ph) * scb, B1 = t1 * sca, B2 = t1 * ssa;
            const double B3 = M0 - 2.0 * t2 * sin(ph) * ssb;
            const int o = 2 * b;
            H[o * 4 + o] = mk(B0 + B3, 0.0);
            H[o * 4 + o + 1] = mk(B1, -B2);
            H[(o + 1) * 4 + o] = mk(B1, B2);
            H[(o + 1) * 4 + o + 1] = mk(B0 - B3, 0.0);
        }
        cd R0 = mk(0, 0), R1 = mk(0, 0);
        for (int i = 0; i < 3; ++i) {
            const cd c = mk(cr[i][0], -cr[i][1]);
            R0 = add(R0, mul(c, eix(tha[i])));
            R1 = add(R1, mul(c, eix(-tha[i])));
        }
        R0 = mul(mk(0.0, tR), R0);
        R1 = mul(mk(0.0, tR), R1);
        H[3] = R0;
        H[6] = R1;
        H[9] = cj(R1);
        H[12] = cj(R0);
    }
    if (Jx) {
        for (int ax = 0; ax < 2; ++ax) {
            cd *J = ax == 0 ? Jx : Jy;
            double x0 = 0, x1 = 0, x2 = 0, x3 = 0;  // derivatives of B0..B3 (phi0 block)
            for (int i = 0; i < 3; ++i) {
                x0 += vb[i][ax] * sin(thb[i]);
                x1 += va[i][ax] * sin(tha[i]);
                x2 += va[i][ax] * cos(tha[i]);
                x3 += vb[i][ax] * cos(thb[i]);
            }
            x0 *= -2.0 * t2 * cos(phi);
            x1 *= -t1;
            x2 *= t1;
            x3 *= -2.0 * t2 * sin(phi);
            J[0] = mk(-(x0 + x3), 0.0);
            J[1] = mk(-x1, x2);
            J[4] = mk(-x1, -x2);
            J[5] = mk(-(x0 - x3), 0.0);
            J[10] = mk(-(x0 - x3), 0.0);  // -phi0 block: the sin(phi0) terms change sign
            J[11] = mk(-x1, x2);
            J[14] = mk(-x1, -x2);
            J[15] = mk(-(x0 + x3), 0.0);
            cd r[4] = {mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0)};
            for (int i = 0; i < 3; ++i) {
                const cd cm = mk(cr[i][0], -cr[i][1]), cp = mk(cr[i][0], cr[i][1]);
                const cd ip = mk(0.0, va[i][ax]), im = mk(0.0, -va[i][ax]);  // +-i a_i
                r[0] = add(r[0], mul(mul(cm, eix(tha[i])), ip));
                r[1] = add(r[1], mul(mul(cm, eix(-tha[i])), im));
                r[2] = add(r[2], scl(mul(mul(cp, eix(tha[i])), ip), -1.0));
                r[3] = add(r[3], scl(mul(mul(cp, eix(-tha[i])), im), -1.0));
            }
            for (int i = 0; i < 3; ++i) r[i] = mul(r[i], mk(0.0, tR));
            J[3] = r[0];
            J[6] = r[1];
            J[9] = r[2];
            J[12] = r[3];
        }
    }
}

// Materials/XYHlattice.h:49-183. p = {tl, tr, lambda, a0}; a_i = (0, a0/sqrt3),
// (-a0/2, -a0/(2 sqrt3)), (a0/2, -a0/(2 sqrt3)); phases e_i = e^{-i k.a_i},
// alpha+- = e_0 +- e_1, beta = e_2. H: diag(l, l, -l, -l) and the (0-1) x (2-3)
// block; J is the same block with e_i -> d e_i/dk = -i a_i e_i, i.e. +dH/dk.
__host__ __device__ inline void xyh_block(const double *p, cd e0, cd e1, cd e2, cd *M)
{
    const double tl = p[0], tr = p[1];
    const cd ap = add(e0, e1), am = add(e0, scl(e1, -1.0));
    M[2] = add(scl(ap, 0.75 * tl + 0.25 * tr), scl(e2, tr));
    M[3] = scl(am, sqrt(3.0) / 4 * (tl - tr));
    M[6] = M[3];
    M[7] = add(scl(ap, 0.25 * tl + 0.75 * tr), scl(e2, tl));
    M[8] = cj(M[2]);
    M[9] = cj(M[6]);
    M[12] = cj(M[3]);
    M[13] = cj(M[7]);
}

__host__ __device__ inline void xyh_lattice(const double *p, const double *k, cd *H, cd *Jx, cd *Jy)
{
    const double a0 = p[3], r3 = sqrt(3.0);
    const double va[3][2] = {{0.0, a0 / r3}, {-0.5 * a0, -0.5 * a0 / r3}, {0.5 * a0, -0.5 * a0 / r3}};
    cd e[3];
    for (int i = 0; i < 3; ++i) e[i] = eix(-(k[0] * va[i][0] + k[1] * va[i][1]));
    if (H) {
        H[0] = mk(p[2], 0.0);
        H[5] = mk(p[2], 0.0);
        H[10] = mk(-p[2], 0.0);
        H[15] = mk(-p[2], 0.0);
        xyh_block(p, e[0], e[1], e[2], H);
    }
    if (Jx) {
        for (int ax = 0; ax < 2; ++ax) {
            cd d[3];
            for (int i = 0; i < 3; ++i) d[i] = mul(mk(0.0, -va[i][ax]), e[i]);
            xyh_block(p, d[0], d[1], d[2], ax == 0 ? Jx : Jy);
        }
    }
}

// Materials/TMDCs.h:48-320: three-band d-orbital model (dz2, dxy, dx2-y2) of
// transition-metal dichalcogenide monolayers with up to third-neighbour
// hopping; p = {a0, soc, e1, e2, t0, t1, t2, t11, t12, t22, r0, r1, r2, r11,
// r12, u0, u1, u2, u11, u12, u22}. N = 6 adds spin-orbit i soc on the
// (dxy, dx2-y2) pair, +soc for spin up and -soc for spin down. J is the
// class's derivative matrix: +dH/dk (TMDCs.h:214-283).
struct TmdcTerms {
    double V0, V11, V22;
    cd V1, V2, V12;
};

// V's of TMDCs.h:166-192, or their derivative along x (dir 0) or y (dir 1)
// when dir >= 0: every trigonometric factor f(n alpha) or f(n beta) is
// replaced by its derivative in the product rule.
__host__ __device__ inline TmdcTerms tmdc_terms(const double *p, const double *k, int dir)
{
    const double a0 = p[0], e1 = p[2], e2 = p[3], t0 = p[4], t1 = p[5], t2 = p[6], t11 = p[7], t12 = p[8],
                 t22 = p[9], r0 = p[10], r1 = p[11], r2 = p[12], r11 = p[13], r12 = p[14], u0 = p[15], u1 = p[16],
                 u2 = p[17], u11 = p[18], u12 = p[19], u22 = p[20];
    const double r3 = sqrt(3.0);
    const double al = k[0] * a0 / 2, be = k[1] * a0 * r3 / 2;
    double ca[5], sa[5], cb[5], sb[5], dca[5], dsa[5], dcb[5], dsb[5];
    for (int n = 1; n <= 4; ++n) {
        ca[n] = cos(n * al);
        sa[n] = sin(n * al);
        cb[n] = cos(n * be);
        sb[n] = sin(n * be);
        dca[n] = dir == 0 ? -n * a0 / 2 * sa[n] : 0.0;
        dsa[n] = dir == 0 ? n * a0 / 2 * ca[n] : 0.0;
        dcb[n] = dir == 1 ? -n * a0 * r3 / 2 * sb[n] : 0.0;
        dsb[n] = dir == 1 ? n * a0 * r3 / 2 * cb[n] : 0.0;
    }
    // product rule for a product of two factors: (f g)' = f' g + f g'
    auto pr = [&](double f, double df, double g, double dg) { return dir < 0 ? f * g : df * g + f * dg; };
    auto one = [&](double f, double df) { return dir < 0 ? f : df; };
    const double c0 = dir < 0 ? 1.0 : 0.0;  // constants drop out of derivatives
    TmdcTerms T;
    T.V0 = c0 * e1 + 2 * t0 * (2 * pr(ca[1], dca[1], cb[1], dcb[1]) + one(ca[2], dca[2])) +
           2 * r0 * (2 * pr(ca[3], dca[3], cb[1], dcb[1]) + one(cb[2], dcb[2])) +
           2 * u0 * (2 * pr(ca[2], dca[2], cb[2], dcb[2]) + one(ca[4], dca[4]));
    T.V1 = mk(-2 * r3 * t2 * pr(sa[1], dsa[1], sb[1], dsb[1]) + 2 * (r1 + r2) * pr(sa[3], dsa[3], sb[1], dsb[1]) -
                  2 * r3 * u2 * pr(sa[2], dsa[2], sb[2], dsb[2]),
              2 * t1 * pr(sa[1], dsa[1], 2 * ca[1] + cb[1], 2 * dca[1] + dcb[1]) +
                  2 * (r1 - r2) * pr(sa[3], dsa[3], cb[1], dcb[1]) +
                  2 * u1 * pr(sa[2], dsa[2], 2 * ca[2] + cb[2], 2 * dca[2] + dcb[2]));
    T.V2 = mk(2 * t2 * (one(ca[2], dca[2]) - pr(ca[1], dca[1], cb[1], dcb[1])) -
                  2 / r3 * (r1 + r2) * (pr(ca[3], dca[3], cb[1], dcb[1]) - one(cb[2], dcb[2])) +
                  2 * u2 * (one(ca[4], dca[4]) - pr(ca[2], dca[2], cb[2], dcb[2])),
              2 * r3 * t1 * pr(ca[1], dca[1], sb[1], dsb[1]) +
                  2 / r3 * (r1 - r2) * pr(sb[1], dsb[1], ca[3] + 2 * cb[1], dca[3] + 2 * dcb[1]) +
                  2 * r3 * u1 * pr(ca[2], dca[2], sb[2], dsb[2]));
    T.V11 = c0 * e2 + (t11 + 3 * t22) * pr(ca[1], dca[1], cb[1], dcb[1]) + 2 * t11 * one(ca[2], dca[2]) +
            4 * r11 * pr(ca[3], dca[3], cb[1], dcb[1]) + 2 * (r11 + r3 * r12) * one(cb[2], dcb[2]) +
            (u11 + 3 * u22) * pr(ca[2], dca[2], cb[2], dcb[2]) + 2 * u11 * one(ca[4], dca[4]);
    T.V12 = mk(r3 * (t22 - t11) * pr(sa[1], dsa[1], sb[1], dsb[1]) + 4 * r12 * pr(sa[3], dsa[3], sb[1], dsb[1]) +
                   r3 * (u22 - u11) * pr(sa[2], dsa[2], sb[2], dsb[2]),
               4 * t12 * pr(sa[1], dsa[1], ca[1] - cb[1], dca[1] - dcb[1]) +
                   4 * u12 * pr(sa[2], dsa[2], ca[2] - cb[2], dca[2] - dcb[2]));
    T.V22 = c0 * e2 + (3 * t11 + t22) * pr(ca[1], dca[1], cb[1], dcb[1]) + 2 * t22 * one(ca[2], dca[2]) +
            2 * r11 * (2 * pr(ca[3], dca[3], cb[1], dcb[1]) + one(cb[2], dcb[2])) +
            2 / r3 * r12 * (4 * pr(ca[3], dca[3], cb[1], dcb[1]) - one(cb[2], dcb[2])) +
            (3 * u11 + u22) * pr(ca[2], dca[2], cb[2], dcb[2]) + 2 * u22 * one(ca[4], dca[4]);
    return T;
}

// writes one 3x3 block at (o, o) of an N x N matrix; s adds i s to (1,2) and -i s to (2,1)
__host__ __device__ inline void tmdc_block(const TmdcTerms &T, double s, int o, int N, cd *M)
{
    auto at = [&](int i, int j) -> cd & { return M[(o + i) * N + o + j]; };
    at(0, 0) = mk(T.V0, 0.0);
    at(0, 1) = T.V1;
    at(0, 2) = T.V2;
    at(1, 0) = cj(T.V1);
    at(1, 1) = mk(T.V11, 0.0);
    at(1, 2) = add(T.V12, mk(0.0, s));
    at(2, 0) = cj(T.V2);
    at(2, 1) = add(cj(T.V12), mk(0.0, -s));
    at(2, 2) = mk(T.V22, 0.0);
}

__host__ __device__ inline void tmdc(const double *p, int N, const double *k, cd *H, cd *Jx, cd *Jy)
{
    const bool soc = N == 6;
    if (H) {
        const TmdcTerms T = tmdc_terms(p, k, -1);
        tmdc_block(T, soc ? p[1] : 0.0, 0, N, H);
        if (soc) tmdc_block(T, -p[1], 3, N, H);
    }
    if (Jx) {
        for (int ax = 0; ax < 2; ++ax) {
            const TmdcTerms T = tmdc_terms(p, k, ax);
            cd *J = ax == 0 ? Jx : Jy;
            tmdc_block(T, 0.0, 0, N, J);
            if (soc) tmdc_block(T, 0.0, 3, N, J);
        }
    }
}

// Materials/Haldane2L.h:52-311: two Haldane layers (A top, B top, A bottom,
// B bottom) with real interlayer hoppings h (polytype), B2 signs s_l (+-1,
// the complex-conjugate layer of the primed stackings) and per-layer
// parameters. p = {t1, t2, M0, phi0 (top), t1, t2, M0, phi0 (bottom), a0,
// same, s0, s1, h0, h1, h2, h3}; same = 1 when no bottom parameter was
// given. With different layers the class forms the bottom B3 from the
// top's finished B3 (Haldane2L.h:307): M0_b - 2 t2_b sin(phi0_b) B3_top,
// restated as written; J is the intended -dH/dk of each layer.
__host__ __device__ inline void haldane2l(const double *p, const double *k, cd *H, cd *Jx, cd *Jy)
{
    double va[3][2], vb[3][2];
    honeycomb(p[8], va, vb);
    double tha[3], thb[3];
    for (int i = 0; i < 3; ++i) {
        tha[i] = k[0] * va[i][0] + k[1] * va[i][1];
        thb[i] = k[0] * vb[i][0] + k[1] * vb[i][1];
    }
    const bool same = p[9] != 0.0;
    if (H) {
        double sca = 0, ssa = 0, scb = 0, ssb = 0;
        for (int i = 0; i < 3; ++i) {
            sca += cos(tha[i]);
            ssa += sin(tha[i]);
            scb += cos(thb[i]);
            ssb += sin(thb[i]);
        }
        double B[2][4];
        for (int l = 0; l < 2; ++l) {
            const double *q = p + 4 * l;  // t1, t2, M0, phi0 of layer l
            B[l][0] = 2.0 * q[1] * cos(q[3]) * scb;
            B[l][1] = q[0] * sca;
            B[l][2] = q[0] * ssa;
            B[l][3] = q[2] - 2.0 * q[1] * sin(q[3]) * ssb;
        }
        if (same) {
            for (int c = 0; c < 4; ++c) B[1][c] = B[0][c];
        } else {
            B[1][3] = p[6] - 2.0 * p[5] * sin(p[7]) * B[0][3];
        }
        for (int l = 0; l < 2; ++l) {
            const int o = 2 * l;
            const double sg = p[10 + l];
            H[o * 4 + o] = mk(B[l][0] + B[l][3], 0.0);
            H[o * 4 + o + 1] = mk(B[l][1], -sg * B[l][2]);
            H[(o + 1) * 4 + o] = mk(B[l][1], sg * B[l][2]);
            H[(o + 1) * 4 + o + 1] = mk(B[l][0] - B[l][3], 0.0);
        }
        H[2] = mk(p[12], 0.0);
        H[3] = mk(p[13], 0.0);
        H[6] = mk(p[14], 0.0);
        H[7] = mk(p[15], 0.0);
        H[8] = mk(p[12], 0.0);
        H[9] = mk(p[14], 0.0);
        H[12] = mk(p[13], 0.0);
        H[13] = mk(p[15], 0.0);
    }
    if (Jx) {
        for (int ax = 0; ax < 2; ++ax) {
            cd *J = ax == 0 ? Jx : Jy;
            for (int l = 0; l < 2; ++l) {
                const double *q = p + 4 * l;
                double x0 = 0, x1 = 0, x2 = 0, x3 = 0;
                for (int i = 0; i < 3; ++i) {
                    x0 += vb[i][ax] * sin(thb[i]);
                    x1 += va[i][ax] * sin(tha[i]);
                    x2 += va[i][ax] * cos(tha[i]);
                    x3 += vb[i][ax] * cos(thb[i]);
                }
                x0 *= -2.0 * q[1] * cos(q[3]);
                x1 *= -q[0];
                x2 *= q[0];
                x3 *= -2.0 * q[1] * sin(q[3]);
                const int o = 2 * l;
                const double sg = p[10 + l];
                J[o * 4 + o] = mk(-(x0 + x3), 0.0);
                J[o * 4 + o + 1] = mk(-x1, sg * x2);
                J[(o + 1) * 4 + o] = mk(-x1, -sg * x2);
                J[(o + 1) * 4 + o + 1] = mk(-(x0 - x3), 0.0);
            }
        }
    }
}

// Materials/Bi2Se3surf.h (class BieSe3surf): surface bands of Bi2Se3 on a
// triangular lattice, parameters hard-coded (eV, a = 4.14 Angstrom). The
// class's GenJMatrix writes its sixth derivative component past the end of
// five-element arrays (Bi2Se3surf.h:104,159,176): undefined behaviour in the
// reference; here that component is the derivative of h5, as the formulas
// intend, giving J = -dH/dk.
__host__ __device__ inline void bi2se3_consts(double c[10])
{
    const double ev = 27.211386245988;
    c[0] = -0.0255 / ev; c[1] = 0.1937 / ev; c[2] = 0.2240 / ev; c[3] = 0.0551 / ev;  // A0, A11, A12, A14
    c[4] = 0.0164 / ev;  c[5] = 0.1203 / ev; c[6] = 0.3264 / ev; c[7] = 0.0;          // B0, B11, B12, B14
    c[8] = -1.6978 / ev;                                                                // m11
    c[9] = 4.14 / 0.529177210903;                                                       // a (bohr)
}

__host__ __device__ inline void bi2se3_surf(const double *k, cd *H, cd *Jx, cd *Jy)
{
    double c[10];
    bi2se3_consts(c);
    const double A0 = c[0], A11 = c[1], A12 = c[2], A14 = c[3], B0 = c[4], B11 = c[5], B12 = c[6], B14 = c[7];
    const double m11 = c[8], a = c[9], r3 = sqrt(3.0);
    const double va[3][2] = {{a, 0.0}, {-a / 2, r3 / 2 * a}, {-a / 2, -r3 / 2 * a}};
    const double vb[3][2] = {{0.0, r3 * a / 3}, {-a / 2, -r3 / 6 * a}, {a / 2, -r3 / 6 * a}};
    double tha[3], thb[3];
    for (int i = 0; i < 3; ++i) {
        tha[i] = k[0] * va[i][0] + k[1] * va[i][1];
        thb[i] = k[0] * vb[i][0] + k[1] * vb[i][1];
    }
    const double w = -2.0 * 3.14159265358979323846 / 3.0, sw = sin(w), cw = cos(w);
    const double bc = sqrt(1.0 - B0 * B0 / (B11 * B11));
    const double g5 = 6.0 * A11 + 6.0 * B11 + m11;  // h5 at Gamma
    // h0..h5 (or their derivative along axis ax when ax >= 0)
    auto comps = [&](int ax, double h[6]) {
        double ca = 0, sa = 0, cb = 0, sb = 0;
        double fa[3], ga[3], fb[3], gb[3];  // sin/cos of the phases or their derivatives
        for (int i = 0; i < 3; ++i) {
            const double da = ax < 0 ? 1.0 : va[i][ax], db = ax < 0 ? 1.0 : vb[i][ax];
            fa[i] = ax < 0 ? sin(tha[i]) : da * cos(tha[i]);   // (sin tha)'
            ga[i] = ax < 0 ? cos(tha[i]) : -da * sin(tha[i]);  // (cos tha)'
            fb[i] = ax < 0 ? sin(thb[i]) : db * cos(thb[i]);
            gb[i] = ax < 0 ? cos(thb[i]) : -db * sin(thb[i]);
            ca += ga[i];
            sa += fa[i];
            cb += gb[i];
            sb += fb[i];
        }
        h[0] = 2.0 * A0 * ca + 2.0 * B0 * cb;
        h[1] = -2.0 * A14 * sw * (fa[1] - fa[2]) + 2.0 * B14 * (fb[0] + cw * (fb[1] - fb[2]));
        h[2] = -2.0 * B14 * sw * (fb[1] - fb[2]) - 2.0 * A14 * (fa[0] + cw * (fa[1] + fa[2]));
        h[3] = 2.0 * A12 * sa;
        h[4] = -2.0 * B12 * sb;
        h[5] = 2.0 * A11 * ca + 2.0 * B11 * cb + (ax < 0 ? m11 : 0.0);
    };
    if (H) {
        double h[6];
        comps(-1, h);
        const double b0 = h[0] + B0 / B11 * (-h[5] + g5), b1 = bc * h[1], b2 = bc * h[2], b3 = bc * h[3];
        H[0] = mk(b0 + b3, 0.0);
        H[1] = mk(b1, -b2);
        H[2] = mk(b1, b2);
        H[3] = mk(b0 - b3, 0.0);
    }
    if (Jx) {
        for (int ax = 0; ax < 2; ++ax) {
            double h[6];
            comps(ax, h);
            const double d0 = h[0] - B0 / B11 * h[5], d1 = bc * h[1], d2 = bc * h[2], d3 = bc * h[3];
            cd *J = ax == 0 ? Jx : Jy;
            J[0] = mk(-(d0 + d3), 0.0);
            J[1] = mk(-d1, d2);
            J[2] = mk(-d1, -d2);
            J[3] = mk(-(d0 - d3), 0.0);
        }
    }
}

// Materials/MultiWSM.h:47-200: multi-Weyl k.p model, p = {n, v, alphan, phi0, a0};
// B1 + i B2 = alphan |k_perp|^n e^{i n atan2(ky, kx)}, B3 = v kz (phi0 and a0
// enter only the Brillouin zone). J = -dH/dk in the class's closed form,
// which divides by |k_perp|^2 (non-finite on the kz axis, as in the reference).
__host__ __device__ inline void multi_wsm(const double *p, const double *k, cd *H, cd *Jx, cd *Jy, cd *Jz)
{
    const double n = p[0], v = p[1], al = p[2];
    const double ph = atan2(k[1], k[0]), ps = k[0] * k[0] + k[1] * k[1];
    if (H) {
        const double r = al * pow(ps, n / 2.0);
        const double b1 = r * cos(n * ph), b2 = r * sin(n * ph), b3 = v * k[2];
        H[0] = mk(b3, 0.0);
        H[1] = mk(b1, -b2);
        H[2] = mk(b1, b2);
        H[3] = mk(-b3, 0.0);
    }
    if (Jx) {
        const double c = cos(n * ph), s = sin(n * ph), pm = pow(ps, n / 2.0 - 1.0), pn = pow(ps, n / 2.0);
        const double x1 = n * al * k[0] * c * pm + n * al * s * pn * k[1] / ps;
        const double y1 = n * al * k[1] * c * pm - n * al * s * pn * k[0] / ps;
        const double x2 = n * al * k[0] * s * pm - n * al * k[1] * c * pn / ps;
        const double y2 = n * al * k[1] * s * pm + n * al * k[0] * c * pn / ps;
        Jx[1] = mk(-x1, x2);
        Jx[2] = mk(-x1, -x2);
        Jy[1] = mk(-y1, y2);
        Jy[2] = mk(-y1, -y2);
        if (Jz) {
            Jz[0] = mk(-v, 0.0);
            Jz[3] = mk(v, 0.0);
        }
    }
}

// Materials/WDSM_TaAs.h:41-255: four-band Weyl/Dirac semimetal model, p = {t,
// delta, my, mz, a0}. J as the class writes it: the x component is
// -d/dkx without the factor a0, the y slot is overwritten by the z
// component (-d/dkz, again without a0; WDSM_TaAs.h:170-185), the z slot
// stays zero.
__host__ __device__ inline void wdsm_taas(const double *p, const double *k, cd *H, cd *Jx, cd *Jy)
{
    const double t = p[0], dl = p[1], my = p[2], mz = p[3], a0 = p[4];
    const double sx = sin(k[0] * a0), cx = cos(k[0] * a0), sy = sin(k[1] * a0), cy = cos(k[1] * a0);
    const double sz = sin(k[2] * a0), cz = cos(k[2] * a0);
    if (H) {
        const double m = t * cx + t * my * (1.0 - cy) + t * mz * (1.0 - cz);
        H[1] = mk(t * sz, 0.0);
        H[4] = mk(t * sz, 0.0);
        H[11] = mk(-t * sz, 0.0);
        H[14] = mk(-t * sz, 0.0);
        H[2] = mk(m, -t * sy);
        H[7] = mk(m, -t * sy);
        H[8] = mk(m, t * sy);
        H[13] = mk(m, t * sy);
        H[3] = mk(0.0, -t * cy * dl);
        H[6] = mk(0.0, -t * cy * dl);
        H[9] = mk(0.0, t * cy * dl);
        H[12] = mk(0.0, t * cy * dl);
    }
    if (Jx) {
        const double dx = -t * sx;
        Jx[2] = mk(-dx, 0.0);
        Jx[7] = mk(-dx, 0.0);
        Jx[8] = mk(-dx, 0.0);
        Jx[13] = mk(-dx, 0.0);
        const double z1 = t * mz * sz, z2 = t * cz;  // z derivatives in the y slot
        Jy[2] = mk(-z1, 0.0);
        Jy[7] = mk(-z1, 0.0);
        Jy[8] = mk(-z1, 0.0);
        Jy[13] = mk(-z1, 0.0);
        Jy[1] = mk(-z2, 0.0);
        Jy[4] = mk(-z2, 0.0);
        Jy[11] = mk(z2, 0.0);
        Jy[14] = mk(z2, 0.0);
    }
}

// Dispatch: H and (Jx, Jy, Jz) of the analytic class `kind`. H may be null;
// Jx, Jy are written together (Jz may be null: only MultiWSM has one).
__host__ __device__ inline void analytic_eval(int kind, int N, const double *p, const double *k, cd *H, cd *Jx, cd *Jy,
                                              cd *Jz)
{
    switch (kind) {
    case SBE_MODEL_WILSON_MASS: wilson_mass(p, k, H, Jx, Jy); break;
    case SBE_MODEL_SQUARE_SBAND: square_sband(p, k, H, Jx, Jy); break;
    case SBE_MODEL_KANE_MELE: kane_mele(p, k, H, Jx, Jy); break;
    case SBE_MODEL_XYH_LATTICE: xyh_lattice(p, k, H, Jx, Jy); break;
    case SBE_MODEL_TMDC: tmdc(p, N, k, H, Jx, Jy); break;
    case SBE_MODEL_HALDANE_2L: haldane2l(p, k, H, Jx, Jy); break;
    case SBE_MODEL_BI2SE3_SURF: bi2se3_surf(k, H, Jx, Jy); break;
    case SBE_MODEL_MULTI_WSM: multi_wsm(p, k, H, Jx, Jy, Jz); break;
    case SBE_MODEL_WDSM_TAAS: wdsm_taas(p, k, H, Jx, Jy); break;
    default: break;
    }
}

__host__ __device__ inline bool is_analytic(int kind) { return kind >= SBE_MODEL_WILSON_MASS && kind <= SBE_MODEL_WDSM_TAAS; }

}  // namespace am
}  // namespace sbe
