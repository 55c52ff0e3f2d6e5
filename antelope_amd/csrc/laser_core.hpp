// laser_core.hpp - one laser pulse (reference laser.h:68-238), shared by the
// host (sample-time tables, output files) and the device (the adaptive
// stepper evaluates A at its own stage times). FP64 throughout.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "antelope/sbe_c.h"

namespace sbe {

// One pulse with its derived constants (laser.h:68-141).
struct PulseConst {
    double E0x, E0y, A0x, A0y, w0, cep, t0;
    double twidth, twidth_clipped, envfactor, period0;
    double xdir[3], ydir[3];
    int env;
};

// Envelope f(t) and f'(t) (laser.h:150-206). SBE_ENV_SIN2 is
// cos^2(t w0/(2 ncycles)) on |t| <= ncycles T/2, a sin^2 pulse of total
// length ncycles periods.
__host__ __device__ inline void envelope_cos4(double x, double ef, double &f, double &fd)
{
    const double c = cos(x * ef), s = sin(x * ef);
    f = pow(c, 4.0);
    fd = -4.0 * ef * s * pow(c, 3.0);
}

__host__ __device__ inline void envelope_gauss(double x, double ef, double &f, double &fd)
{
    const double g = exp(ef * x * x);
    f = g;
    fd = 2.0 * x * ef * g;
}

__host__ __device__ inline void envelope(const PulseConst &p, double t, double &f, double &fd)
{
    const double ef = p.envfactor;
    f = 0.0;
    fd = 0.0;
    switch (p.env) {
    case SBE_ENV_COS4:
        if (fabs(t) <= p.twidth) envelope_cos4(t, ef, f, fd);
        return;
    case SBE_ENV_SIN2:
        if (fabs(t) <= p.twidth) {
            const double c = cos(t * ef), s = sin(t * ef);
            f = c * c;
            fd = -2.0 * ef * s * c;
        }
        return;
    case SBE_ENV_COS4_CLIPPED:
        if (fabs(t) > p.twidth) return;
        if (t > p.twidth_clipped) envelope_cos4(t - p.twidth_clipped, ef, f, fd);
        else if (t > -p.twidth_clipped) f = 1.0;
        else envelope_cos4(t + p.twidth_clipped, ef, f, fd);
        return;
    case SBE_ENV_GAUSS:
        envelope_gauss(t, ef, f, fd);
        return;
    case SBE_ENV_GAUSS_CLIPPED:
        if (t > p.twidth_clipped) envelope_gauss(t - p.twidth_clipped, ef, f, fd);
        else if (t > -p.twidth_clipped) f = 1.0;
        else envelope_gauss(t + p.twidth_clipped, ef, f, fd);
        return;
    default:
        return;
    }
}

// A(t) and E(t) of one pulse in the laser plane (laser.h:209-238, 133-140).
// consistent = false keeps the reference's E, whose f' terms have the
// opposite sign to -dA/dt (laser.h:217-224).
__host__ __device__ inline void pulse_field(const PulseConst &p, double t, bool consistent, double A[3], double E[3])
{
    const double kt = t - p.t0;
    const double ph = kt * p.w0 - p.cep;
    const double sn = sin(ph), cs = cos(ph);
    double f, fd;
    envelope(p, kt, f, fd);
    const double ax = -p.A0x * sn * f, ay = p.A0y * cs * f;
    const double sgn = consistent ? -1.0 : 1.0;
    const double ex = p.E0x * cs * f - sgn * p.A0x * sn * fd;
    const double ey = p.E0y * sn * f + sgn * p.A0y * cs * fd;
    for (int j = 0; j < 3; ++j) {
        A[j] = ax * p.xdir[j] + ay * p.ydir[j];
        E[j] = ex * p.xdir[j] + ey * p.ydir[j];
    }
}

}  // namespace sbe
