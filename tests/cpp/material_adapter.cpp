// Host program for tests/test_material_adapter.py: two material classes
// written against the reference's material interface (BaseMaterial.h:22-78),
// sampled with antelope::SampledMaterial. Prints, one item per line:
//   H <case> <kx> <ky> <kz> <re,im of H(k) row-major ...>      the class's own H(k)
//   M <case> <nrpts> <max_dev>
//   R <case> <n1> <n2> <n3> <ham re,im ...> | <pos re,im ...>
//   Z <case> <bz axes 9> <origin 3>                            after apply_brillouin_zone
//   E <case> <message>                                          an expected error
#include <cmath>
#include <complex>
#include <cstdio>
#include <stdexcept>

#include "antelope/sbe_material.hpp"

using cd = std::complex<double>;

// Haldane's model in the bond-phase form of Haldane.h:86-92,157-170: A-B bonds
// va_i, A-A / B-B second neighbours vb_i (a0 the bond length).
class Haldane : public antelope::BaseMaterial {
public:
    double t1, t2, M0, phi0, a0;
    double va[3][2], vb[3][2];
    Haldane(double t1_, double t2_, double M0_, double phi0_, double a0_)
        : t1(t1_), t2(t2_), M0(M0_), phi0(phi0_), a0(a0_)
    {
        Nband = 2;
        Nval = 1;
        isDipoleZero = true;
        const double s = std::sqrt(3.0);
        const double A[3][2] = {{0.0, a0}, {-s / 2 * a0, -0.5 * a0}, {s / 2 * a0, -0.5 * a0}};
        const double B[3][2] = {{s * a0, 0.0}, {-s / 2 * a0, 1.5 * a0}, {-s / 2 * a0, -1.5 * a0}};
        for (int i = 0; i < 3; ++i)
            for (int c = 0; c < 2; ++c) va[i][c] = A[i][c], vb[i][c] = B[i][c];
    }
    void GenHamiltonian(cd *h, std::array<double, 3> k) override
    {
        double c2 = 0, s2 = 0, c1 = 0, s1 = 0;
        for (int i = 0; i < 3; ++i) {
            const double xa = k[0] * va[i][0] + k[1] * va[i][1], xb = k[0] * vb[i][0] + k[1] * vb[i][1];
            c1 += std::cos(xa), s1 += std::sin(xa), c2 += std::cos(xb), s2 += std::sin(xb);
        }
        const double e0 = 2 * t2 * std::cos(phi0) * c2, ez = M0 - 2 * t2 * std::sin(phi0) * s2;
        h[0] = e0 + ez;
        h[1] = cd(t1 * c1, -t1 * s1);
        h[2] = cd(t1 * c1, t1 * s1);
        h[3] = e0 - ez;
    }
    void GenDipole(cd **, std::array<double, 3>) override {}  // no dipoles: writes nothing
    void GenJMatrix(cd **, std::array<double, 3>) override {}
    std::tuple<std::array<double, 9>, std::array<double, 3>> GenBrillouinzone() override
    {
        const double p = 2 * M_PI;
        return {{p / std::sqrt(3.0) / a0, 0, 0, 0, 2 * p / 3 / a0, 0, 0, 0, 0}, {0, 0, 0}};
    }
    void PrintMaterialInformation() override {}
};

// A two-band chain along a1 with dipoles: H(k) = [[eps, t (1 + e^{-ik.a1})], [h.c., -eps]],
// D_x(k) = [[0, d e^{-ik.a1}], [h.c., x_B]]. Blocks: h(0) = [[eps, t], [t, -eps]],
// h_01(-a1) = t, h_10(+a1) = t; r_x,01(-a1) = d, r_x,10(+a1) = d, r_x,11(0) = x_B.
// Not derived from BaseMaterial: the adapter needs only the members.
struct Chain {
    int Nband = 2, Nval = 1;
    bool isDipoleZero = false;
    double eps = 0.07, t = -0.03, d = 0.4, xB = 1.3, a = 4.0;
    bool fail = false;
    void GenHamiltonian(cd *h, std::array<double, 3> k)
    {
        if (fail) throw std::runtime_error("chain: requested failure");
        const cd e = std::exp(cd(0, -k[0] * a));
        h[0] = eps, h[1] = t * (1.0 + e), h[2] = std::conj(h[1]), h[3] = -eps;
    }
    void GenDipole(cd **D, std::array<double, 3> k)
    {
        const cd e = std::exp(cd(0, -k[0] * a));
        D[0][1] = d * e, D[0][2] = std::conj(D[0][1]), D[0][3] = xB;  // D_y = D_z = 0: left unwritten
    }
};

static void print_blocks(const char *tag, const antelope::SampledMaterial &s)
{
    const sbe_model &m = s.model();
    const int nn = m.nband * m.nband;
    std::printf("M %s %d %.17g\n", tag, m.nrpts, s.max_deviation());
    for (int q = 0; q < m.nrpts; ++q) {
        std::printf("R %s %d %d %d", tag, m.rvec[3 * q], m.rvec[3 * q + 1], m.rvec[3 * q + 2]);
        for (int e = 0; e < 2 * nn; ++e) std::printf(" %.17g", m.ham[2 * q * nn + e]);
        std::printf(" |");
        for (int e = 0; e < 6 * nn; ++e) std::printf(" %.17g", m.pos[6 * q * nn + e]);
        std::printf("\n");
    }
}

int main()
{
    const double a0 = 2.683;
    Haldane hal(0.075, 0.025, 2.54 * 0.025, 0.3 * M_PI, a0);
    const double ks[3][3] = {{0.3, -0.7, 0.0}, {1.1, 0.4, 0.0}, {-0.9, 1.3, 0.0}};
    for (const auto &k : ks) {
        cd h[4];
        hal.GenHamiltonian(h, {k[0], k[1], k[2]});
        std::printf("H haldane %.17g %.17g %.17g", k[0], k[1], k[2]);
        for (const cd &v : h) std::printf(" %.17g %.17g", v.real(), v.imag());
        std::printf("\n");
    }
    const double s3 = std::sqrt(3.0);
    const std::array<double, 9> L = {s3 * a0, 0, 0, s3 / 2 * a0, 1.5 * a0, 0, 0, 0, 10.0};
    {
        antelope::SampledMaterial s(hal, L, {1, 1, 0}, {{0.0, 0.0, 0.0}, {0.0, -a0, 0.0}});
        print_blocks("haldane", s);
        sbe_config cfg{};
        cfg.ksfactor[0] = cfg.ksfactor[1] = cfg.ksfactor[2] = 1.0;
        cfg.ksfactor[1] = 0.5;
        antelope::apply_brillouin_zone(hal, cfg);
        std::printf("Z haldane");
        for (double v : cfg.bz_axes) std::printf(" %.17g", v);
        for (double v : cfg.bz_origin) std::printf(" %.17g", v);
        std::printf("\n");
    }
    {
        antelope::SampledMaterial s(hal, L, {0, 0, 0}, {{0.0, 0.0, 0.0}, {0.0, -a0, 0.0}});
        std::printf("M haldane_short %d %.17g\n", s.model().nrpts, s.max_deviation());
    }
    Chain ch;
    const std::array<double, 9> Lc = {ch.a, 0, 0, 0, 10.0, 0, 0, 0, 10.0};
    {
        antelope::SampledMaterial s(ch, Lc, {1, 0, 0});
        antelope::SampledMaterial moved(std::move(s));  // ownership moves; s frees nothing
        print_blocks("chain", moved);
    }
    try {
        ch.fail = true;
        antelope::SampledMaterial s(ch, Lc, {1, 0, 0});
        std::printf("E chain no error\n");
    } catch (const std::runtime_error &e) {
        std::printf("E chain %s\n", e.what());
    }
    try {
        ch.fail = false;
        antelope::SampledMaterial s(ch, Lc, {1, 0, 0}, {{0, 0, 0}});
        std::printf("E tau no error\n");
    } catch (const std::runtime_error &e) {
        std::printf("E tau %s\n", e.what());
    }
    return 0;
}
