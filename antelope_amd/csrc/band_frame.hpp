// band_frame.hpp - eigenvectors of a 2x2 Hermitian H = d0 I + d.sigma as a
// frame of Bloch axes (host and device).
//
// The velocity gauge propagates rho in the band basis of H(k) (SBEs.h:284-373),
// whose off-diagonal elements depend on the phases of the eigenvectors. The
// reference keeps LAPACK's (unspecified) phases; here they are fixed by one
// rule for every band count: column j of U is multiplied by
// conj(U[p][j]) / |U[p][j]|, p being the first row with |U[p][j]|^2 >= 1/(2N).
// For N = 2 with h = d/|d|: the lower vector u1 pivots on row 0 when
// h_z <= 1/2, the upper u2 when h_z >= -1/2.
//
// With U = [u1 u2] (lower band first), U^dag (v.sigma) U = sum_i (v.f_i) sigma_i
// where f_z = -h and f_x - i f_y = (u1^dag sigma_j u2)_j. So a band-basis Bloch
// vector is (s.f_x, s.f_y, s.f_z) of the orbital one. At d = 0 U = I.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace sbe {

struct BandFrame {
    double fx[3], fy[3], fz[3];
};

__host__ __device__ inline BandFrame band_frame(double dx, double dy, double dz)
{
    BandFrame F;
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    if (!(r > 0.0)) {
        for (int i = 0; i < 3; ++i) F.fx[i] = F.fy[i] = F.fz[i] = 0.0;
        F.fx[0] = F.fy[1] = F.fz[2] = 1.0;
        return F;
    }
    const double hx = dx / r, hy = dy / r, hz = dz / r;
    // u1 = (a1, b1), u2 = (a2, b2) as (re, im) pairs
    double a1r, a1i, b1r, b1i, a2r, a2i, b2r, b2i;
    if (hz <= 0.5) {  // ((1 - hz), -(hx + i hy)) / sqrt(2 (1 - hz))
        const double n = 1.0 / sqrt(2.0 * (1.0 - hz));
        a1r = (1.0 - hz) * n; a1i = 0.0;
        b1r = -hx * n;        b1i = -hy * n;
    } else {          // (-(hx - i hy), 1 + hz) / sqrt(2 (1 + hz))
        const double n = 1.0 / sqrt(2.0 * (1.0 + hz));
        a1r = -hx * n;        a1i = hy * n;
        b1r = (1.0 + hz) * n; b1i = 0.0;
    }
    if (hz >= -0.5) { // (1 + hz, hx + i hy) / sqrt(2 (1 + hz))
        const double n = 1.0 / sqrt(2.0 * (1.0 + hz));
        a2r = (1.0 + hz) * n; a2i = 0.0;
        b2r = hx * n;         b2i = hy * n;
    } else {          // (hx - i hy, 1 - hz) / sqrt(2 (1 - hz))
        const double n = 1.0 / sqrt(2.0 * (1.0 - hz));
        a2r = hx * n;         a2i = -hy * n;
        b2r = (1.0 - hz) * n; b2i = 0.0;
    }
    // p = conj(a1) b2, q = conj(b1) a2, u = conj(a1) a2, v = conj(b1) b2
    const double pr = a1r * b2r + a1i * b2i, pi = a1r * b2i - a1i * b2r;
    const double qr = b1r * a2r + b1i * a2i, qi = b1r * a2i - b1i * a2r;
    const double ur = a1r * a2r + a1i * a2i, ui = a1r * a2i - a1i * a2r;
    const double vr = b1r * b2r + b1i * b2i, vi = b1r * b2i - b1i * b2r;
    // w_x = p + q, w_y = -i p + i q, w_z = u - v; f_x = Re w, f_y = -Im w
    const double wxr = pr + qr, wxi = pi + qi;
    const double wyr = pi - qi, wyi = -pr + qr;
    const double wzr = ur - vr, wzi = ui - vi;
    F.fx[0] = wxr; F.fx[1] = wyr; F.fx[2] = wzr;
    F.fy[0] = -wxi; F.fy[1] = -wyi; F.fy[2] = -wzi;
    F.fz[0] = -hx; F.fz[1] = -hy; F.fz[2] = -hz;
    return F;
}

}  // namespace sbe
