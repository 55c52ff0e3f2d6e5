// host_setup.cpp - k-independent set-up done once per context on the host:
// laser pulses (FP64), the k-grid block, and the Runge-Kutta tableau.
#include <cmath>
#include <cstring>

#include "sbe_internal.hpp"

namespace sbe {

namespace {
constexpr double kPi = 3.14159265358979323846;

}  // namespace

int Laser::init(const sbe_config &cfg, std::string &err)
{
    if (cfg.npulses < 1 || cfg.pulses == nullptr) {
        err = "laser: no pulses";
        return SBE_ERR_CONFIG;
    }
    if (!(cfg.dt > 0.0) || !std::isfinite(cfg.dt)) {
        err = "calc.dt must be a positive finite number";
        return SBE_ERR_CONFIG;
    }
    pulses.clear();
    consistent = cfg.efield_mode == SBE_EFIELD_CONSISTENT;
    double tstart = 0, tend = 0;
    for (int i = 0; i < cfg.npulses; ++i) {
        const sbe_pulse &s = cfg.pulses[i];
        if (!(s.w0 > 0.0) || !(s.ncycles > 0.0)) {
            err = "laser: pulse " + std::to_string(i) + " needs w0 > 0 and ncycles > 0";
            return SBE_ERR_CONFIG;
        }
        PulseConst p{};
        p.w0 = s.w0;
        p.cep = s.cep;
        p.t0 = s.t0;
        p.env = s.envelope;
        p.period0 = 2.0 * kPi / s.w0;
        const double span = s.ncycles * p.period0;
        switch (s.envelope) {
        case SBE_ENV_GAUSS:
            p.twidth = 2.0 * span;
            p.envfactor = -4.0 * std::log(2.0) / (span * span);
            break;
        case SBE_ENV_GAUSS_CLIPPED:
            p.twidth_clipped = s.ncycles_clipped * p.period0 / 2;
            p.twidth = 2.0 * span + p.twidth_clipped;
            p.envfactor = -4.0 * std::log(2.0) / (span * span);
            break;
        case SBE_ENV_COS4:
        case SBE_ENV_SIN2:
            p.twidth = span / 2;
            p.envfactor = s.w0 / (2 * s.ncycles);
            break;
        case SBE_ENV_COS4_CLIPPED:
            p.twidth_clipped = s.ncycles_clipped * p.period0 / 2;
            p.twidth = (s.ncycles + s.ncycles_clipped) * p.period0 / 2;
            p.envfactor = s.w0 / (2 * s.ncycles);
            break;
        default:
            err = "invalid envelope type";
            return SBE_ERR_CONFIG;
        }
        const double norm = std::sqrt(1.0 / (1.0 + s.ellip * s.ellip));
        p.E0x = norm * s.E0;
        p.E0y = s.ellip * norm * s.E0;
        p.A0x = p.E0x / s.w0;
        p.A0y = p.E0y / s.w0;
        // laser plane axes from (thetaz, phiz, phix), laser.h:129-140
        const double x0[3] = {std::cos(s.thetaz) * std::cos(s.phiz),
                              std::cos(s.thetaz) * std::sin(s.phiz), -std::sin(s.thetaz)};
        const double y0[3] = {-std::sin(s.phiz), std::cos(s.phiz), 0.0};
        for (int j = 0; j < 3; ++j) {
            p.xdir[j] = std::cos(s.phix) * x0[j] + std::sin(s.phix) * y0[j];
            p.ydir[j] = -std::sin(s.phix) * x0[j] + std::cos(s.phix) * y0[j];
        }
        const double a = s.t0 - p.twidth, b = s.t0 + p.twidth;
        if (i == 0 || a < tstart) tstart = a;
        if (i == 0 || b > tend) tend = b;
        pulses.push_back(p);
    }
    dt = cfg.dt;
    atmin = tstart - cfg.offset_before;
    atmax = tend + cfg.offset_after;
    for (int j = 0; j < 3; ++j) A_at_min[j] = E_at_min[j] = 0.0;
    double A[3], E[3];
    eval(atmin, A, E);
    for (int j = 0; j < 3; ++j) {
        A_at_min[j] = A[j];
        E_at_min[j] = E[j];
    }
    nt = static_cast<int64_t>(std::floor((atmax - atmin) / dt));
    if (nt < 1) {
        err = "laser: time window shorter than one step";
        return SBE_ERR_CONFIG;
    }
    return SBE_OK;
}

void Laser::eval(double t, double A[3], double E[3]) const
{
    double a[3] = {0, 0, 0}, e[3] = {0, 0, 0};
    for (const PulseConst &p : pulses) {
        double pa[3], pe[3];
        pulse_field(p, t, consistent, pa, pe);
        for (int j = 0; j < 3; ++j) {
            a[j] += pa[j];
            e[j] += pe[j];
        }
    }
    for (int j = 0; j < 3; ++j) {
        A[j] = a[j] - A_at_min[j];
        E[j] = e[j] - E_at_min[j];
    }
}

// ----------------------------------------------------------------- grid --
int Grid::build(const sbe_config &cfg, const sbe_model &m, int64_t kb, int64_t ke,
                std::string &err)
{
    for (int d = 0; d < 3; ++d) {
        n[d] = cfg.npoints[d];
        if (n[d] <= 0) {
            err = "calc.Npoints entries must be > 0";
            return SBE_ERR_CONFIG;
        }
    }
    ntotal = static_cast<int64_t>(n[0]) * n[1] * n[2];
    if (kb == -1 && ke == -1) {
        kb = 0;
        ke = ntotal;
    }
    if (kb < 0 || ke > ntotal || kb >= ke) {
        err = "k range [" + std::to_string(kb) + "," + std::to_string(ke) +
              ") outside the grid of " + std::to_string(ntotal) + " points";
        return SBE_ERR_RANGE;
    }
    k_begin = kb;
    k_end = ke;

    // Material Brillouin zone: Haldane.h:77-81 (rectangle centred at 0) or the
    // reciprocal vectors of the Wannier lattice (Wannier90.h:508-521).
    std::memset(axes, 0, sizeof axes);
    std::memset(origin, 0, sizeof origin);
    if (m.kind == SBE_MODEL_HALDANE) {
        axes[0] = 2.0 * kPi / std::sqrt(3.0) / m.a0;
        axes[4] = 2.0 * (2.0 * kPi / 3.0 / m.a0);
    } else if (m.kind != SBE_MODEL_WANNIER) {
        // boxes 2 kxMax x 2 kyMax (x 2 kzMax) centred at 0 of the analytic
        // classes: WilsonMass.h:56-60, SquareSBand.h:54-58, KaneMele.h:68-72,
        // XYHlattice.h:65-69, TMDCs.h:143-148 and below
        const double *p = m.params, r3 = std::sqrt(3.0);
        double kx = 0, ky = 0;
        switch (m.kind) {
        case SBE_MODEL_WILSON_MASS: kx = kPi / p[3]; ky = kPi / p[4]; break;
        case SBE_MODEL_SQUARE_SBAND: kx = ky = kPi / p[2]; break;
        case SBE_MODEL_KANE_MELE: kx = kPi / r3 / p[5]; ky = 2.0 * kPi / 3.0 / p[5]; break;
        case SBE_MODEL_XYH_LATTICE: kx = kPi / p[3]; ky = 2.0 * kPi / r3 / p[3]; break;
        case SBE_MODEL_TMDC: kx = kPi / p[0]; ky = 2.0 * kPi / r3 / p[0]; break;
        case SBE_MODEL_HALDANE_2L: kx = kPi / r3 / p[8]; ky = 2.0 * kPi / 3.0 / p[8]; break;  // Haldane2L.h:151-154
        case SBE_MODEL_BI2SE3_SURF: {  // Bi2Se3surf.h:71-77, a = 4.14 Angstrom
            const double a = 4.14 / 0.529177210903;
            kx = kPi / a;
            ky = 2.0 * kPi / r3 / a;
            break;
        }
        case SBE_MODEL_MULTI_WSM:  // MultiWSM.h:65-71, WDSM_TaAs.h:61-65: cubes
        case SBE_MODEL_WDSM_TAAS: kx = ky = kPi / p[4]; axes[8] = 2.0 * kPi / p[4]; break;
        default: err = "unknown model kind"; return SBE_ERR_MODEL;
        }
        axes[0] = 2.0 * kx;
        axes[4] = 2.0 * ky;
    } else {
        const double *L = m.lattice;
        const double vol = L[0] * (L[4] * L[8] - L[5] * L[7]) + L[1] * (L[5] * L[6] - L[3] * L[8]) +
                           L[2] * (L[3] * L[7] - L[4] * L[6]);
        if (vol == 0.0) {
            err = "Wannier lattice vectors are linearly dependent";
            return SBE_ERR_MODEL;
        }
        for (int i = 0; i < 3; ++i) {
            const double *u = L + 3 * ((i + 1) % 3), *v = L + 3 * ((i + 2) % 3);
            const double cr[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2],
                                  u[0] * v[1] - u[1] * v[0]};
            for (int j = 0; j < 3; ++j) axes[3 * i + j] = cr[j] * 2.0 * kPi / vol;
        }
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) axes[3 * i + j] *= cfg.ksfactor[i];  // SBEs.h:222-224
    if (cfg.has_bz_axes) std::memcpy(axes, cfg.bz_axes, sizeof axes);      // SBEs.h:226-233
    if (cfg.has_bz_origin) std::memcpy(origin, cfg.bz_origin, sizeof origin);

    // Volume over the non-singleton axes (momaxis.h:208-236, including the
    // reference's sqrt of the triple product in 3D), dV = |V|/Ntotal.
    std::vector<const double *> live;
    for (int i = 0; i < 3; ++i)
        if (n[i] != 1) live.push_back(axes + 3 * i);
    double vol = 1.0;
    if (live.size() == 1) {
        const double *a = live[0];
        vol = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    } else if (live.size() >= 2) {
        const double *a = live[0], *b = live[1];
        const double cr[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                              a[0] * b[1] - b[0] * a[1]};
        if (live.size() == 2) vol = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
        else vol = std::sqrt(cr[0] * live[2][0] + cr[1] * live[2][1] + cr[2] * live[2][2]);
    }
    dV = std::fabs(vol) / static_cast<double>(ntotal);

    const int64_t nk = ke - kb;
    k.assign(static_cast<size_t>(nk) * 3, 0.0);
    w.assign(static_cast<size_t>(nk), 1.0);
    total_weight = 0.0;
    const bool mp = cfg.grid == SBE_GRID_MP;
    for (int64_t g = kb; g < ke; ++g) {
        const int64_t plane = static_cast<int64_t>(n[1]) * n[2];
        const int idx[3] = {static_cast<int>(g / plane), static_cast<int>((g / n[2]) % n[1]),
                            static_cast<int>(g % n[2])};
        double frac[3], wt = 1.0;
        for (int d = 0; d < 3; ++d) {
            if (mp) {
                frac[d] = (idx[d] + 0.5) / n[d] - 0.5;
            } else if (n[d] == 1) {
                frac[d] = 0.0;
            } else {
                frac[d] = static_cast<double>(idx[d]) / (n[d] - 1) - 0.5;
                if (idx[d] == 0 || idx[d] == n[d] - 1) wt *= 0.5;  // trapezoid ends
            }
        }
        const int64_t o = g - kb;
        for (int c = 0; c < 3; ++c)
            k[3 * o + c] = origin[c] + frac[0] * axes[c] + frac[1] * axes[3 + c] + frac[2] * axes[6 + c];
        w[o] = wt;
        total_weight += wt;
    }
    return SBE_OK;
}

// -------------------------------------------------------------- tableau --
int Tableau::init(int rk_order, std::string &err)
{
    *this = Tableau{};
    switch (rk_order) {
    case SBE_RK4:  // classical RK4
        nstage = 4;
        order = 4;
        c[1] = 0.5; c[2] = 0.5; c[3] = 1.0;
        a[1][0] = 0.5;
        a[2][1] = 0.5;
        a[3][2] = 1.0;
        b[0] = 1.0 / 6.0; b[1] = 2.0 / 6.0; b[2] = 2.0 / 6.0; b[3] = 1.0 / 6.0;
        return SBE_OK;
    case SBE_RK5_FEHLBERG:  // fixed-step 6-stage Fehlberg 5th-order weights
        nstage = 6;
        order = 5;
        c[1] = 0.25; c[2] = 3. / 8.; c[3] = 12. / 13.; c[4] = 1.0; c[5] = 0.5;
        a[1][0] = 0.25;
        a[2][0] = 3. / 32.; a[2][1] = 9. / 32.;
        a[3][0] = 1932. / 2197.; a[3][1] = -7200. / 2197.; a[3][2] = 7296. / 2197.;
        a[4][0] = 439. / 216.; a[4][1] = -8.; a[4][2] = 3680. / 513.; a[4][3] = -845. / 4104.;
        a[5][0] = -8. / 27.; a[5][1] = 2.; a[5][2] = -3544. / 2565.; a[5][3] = 1859. / 4104.;
        a[5][4] = -11. / 40.;
        b[0] = 16. / 135.; b[2] = 6656. / 12825.; b[3] = 28561. / 56430.; b[4] = -9. / 50.;
        b[5] = 2. / 55.;
        return SBE_OK;
    case SBE_RK_CASH_KARP:  // J.R. Cash, A.H. Karp, ACM TOMS 16 (1990) 201
        nstage = 6;
        order = 5;
        embedded = true;
        c[1] = 0.2; c[2] = 0.3; c[3] = 0.6; c[4] = 1.0; c[5] = 0.875;
        a[1][0] = 0.2;
        a[2][0] = 3. / 40.; a[2][1] = 9. / 40.;
        a[3][0] = 0.3; a[3][1] = -0.9; a[3][2] = 1.2;
        a[4][0] = -11. / 54.; a[4][1] = 2.5; a[4][2] = -70. / 27.; a[4][3] = 35. / 27.;
        a[5][0] = 1631. / 55296.; a[5][1] = 175. / 512.; a[5][2] = 575. / 13824.;
        a[5][3] = 44275. / 110592.; a[5][4] = 253. / 4096.;
        b[0] = 37. / 378.; b[2] = 250. / 621.; b[3] = 125. / 594.; b[5] = 512. / 1771.;
        bhat[0] = 2825. / 27648.; bhat[2] = 18575. / 48384.; bhat[3] = 13525. / 55296.;
        bhat[4] = 277. / 14336.; bhat[5] = 0.25;
        return SBE_OK;
    default:
        err = "calc.RKorder must be 4 or 5 (or the adaptive Cash-Karp pair 45), got " +
              std::to_string(rk_order);
        return SBE_ERR_CONFIG;
    }
}

}  // namespace sbe
