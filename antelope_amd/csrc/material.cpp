// material.cpp - a material given as code (its H(k), and D(k) when it has
// dipoles) turned into the tight-binding blocks the GPU path takes.
//
// The reference's materials are C++ classes with virtual GenHamiltonian /
// GenDipole / GenJMatrix (BaseMaterial.h:22-78), evaluated per k on the CPU.
// A GPU kernel cannot call them per k point, so sbe_sample_material turns
// one into data: it evaluates the callbacks on the (2 n_i + 1)-point grid of
// the reciprocal lattice and inverts the discrete Fourier transform,
//   h_mn(R) = (1/M) sum_k H_mn(k) e^{-ik.(R + tau_n - tau_m)}   (dipole-free form)
//   h_mn(R) = (1/M) sum_k H_mn(k) e^{-ik.R},  r_mn(R) likewise from D(k)
// for R = n1 a1 + n2 a2 + n3 a3 with |n_i| <= nmax_i. For a model whose
// hoppings reach no further than nmax_i cells along a_i (every tight-binding
// class of Materials/), the blocks reproduce H(k) exactly (a trigonometric
// polynomial of degree <= nmax_i per axis is fixed by 2 nmax_i + 1 samples),
// and J = -dH/dk follows as the reference's Wannier current (Wannier90.h:347-387).
// The blocks are then summed back at 16 k points off the grid and compared with
// the callbacks (*max_dev), which shows a hopping range set too short.
// The result is an SBE_MODEL_WANNIER model; the dipole-free form carries the
// orbital offsets tau in the diagonal of the R = 0 position block, as the
// reference's isDipoleZero branch reads them (Wannier90.h:178-194).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "sbe_internal.hpp"

namespace {

using cplx = std::complex<double>;

bool invert3(const double a[9], double inv[9])
{
    const double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                       a[2] * (a[3] * a[7] - a[4] * a[6]);
    if (!(std::fabs(det) > 0.0) || !std::isfinite(det)) return false;
    inv[0] = (a[4] * a[8] - a[5] * a[7]) / det;
    inv[1] = (a[2] * a[7] - a[1] * a[8]) / det;
    inv[2] = (a[1] * a[5] - a[2] * a[4]) / det;
    inv[3] = (a[5] * a[6] - a[3] * a[8]) / det;
    inv[4] = (a[0] * a[8] - a[2] * a[6]) / det;
    inv[5] = (a[2] * a[3] - a[0] * a[5]) / det;
    inv[6] = (a[3] * a[7] - a[4] * a[6]) / det;
    inv[7] = (a[1] * a[6] - a[0] * a[7]) / det;
    inv[8] = (a[0] * a[4] - a[1] * a[3]) / det;
    return true;
}

}  // namespace

extern "C" {

int sbe_sample_material(int32_t nband, int32_t nval, int32_t is_dipole_zero, const double lattice[9],
                        const double *tau, const int32_t nmax[3], sbe_gen_matrix_fn gen_h, sbe_gen_matrix_fn gen_d,
                        void *user, sbe_model *out, double *max_dev)
{
    if (!lattice || !nmax || !gen_h || !out) return sbe::set_thread_error(SBE_ERR_CONFIG, "null argument");
    std::memset(out, 0, sizeof *out);
    if (max_dev) *max_dev = 0.0;
    if (nband < 1 || nband > 32 || nval < 0 || nval > nband)
        return sbe::set_thread_error(SBE_ERR_MODEL, "1 <= nband <= 32 and 0 <= nval <= nband");
    if (!is_dipole_zero && !gen_d)
        return sbe::set_thread_error(SBE_ERR_MODEL, "a material with dipoles needs its D(k) (gen_d)");
    for (int i = 0; i < 3; ++i)
        if (nmax[i] < 0 || nmax[i] > 16) return sbe::set_thread_error(SBE_ERR_RANGE, "nmax entries must be 0..16");
    // reciprocal vectors b_i (rows): b_i . a_j = 2 pi delta_ij, B = 2 pi (A^-1)^T
    double inv[9], b[9];
    if (!invert3(lattice, inv)) return sbe::set_thread_error(SBE_ERR_MODEL, "lattice vectors are not independent");
    for (int i = 0; i < 3; ++i)
        for (int c = 0; c < 3; ++c) b[3 * i + c] = 2.0 * M_PI * inv[3 * c + i];
    const int N = nband, nn = N * N;
    const int M[3] = {2 * nmax[0] + 1, 2 * nmax[1] + 1, 2 * nmax[2] + 1};
    const int64_t nk = static_cast<int64_t>(M[0]) * M[1] * M[2];
    const int nr = static_cast<int>(nk);  // one R per grid point: |n_i| <= nmax_i
    std::vector<int32_t> rv(3 * static_cast<size_t>(nr));
    std::vector<double> rc(3 * static_cast<size_t>(nr));
    int r = 0;
    for (int n1 = -nmax[0]; n1 <= nmax[0]; ++n1)
        for (int n2 = -nmax[1]; n2 <= nmax[1]; ++n2)
            for (int n3 = -nmax[2]; n3 <= nmax[2]; ++n3, ++r) {
                rv[3 * r] = n1, rv[3 * r + 1] = n2, rv[3 * r + 2] = n3;
                for (int c = 0; c < 3; ++c) rc[3 * r + c] = n1 * lattice[c] + n2 * lattice[3 + c] + n3 * lattice[6 + c];
            }
    std::vector<cplx> hsum(static_cast<size_t>(nr) * nn, 0.0), psum;
    if (!is_dipole_zero) psum.assign(static_cast<size_t>(nr) * 3 * nn, 0.0);
    std::vector<double> hk(2 * static_cast<size_t>(nn)), dk(6 * static_cast<size_t>(nn));
    for (int m1 = 0; m1 < M[0]; ++m1)
        for (int m2 = 0; m2 < M[1]; ++m2)
            for (int m3 = 0; m3 < M[2]; ++m3) {
                double k[3];
                for (int c = 0; c < 3; ++c)
                    k[c] = (m1 * b[c]) / M[0] + (m2 * b[3 + c]) / M[1] + (m3 * b[6 + c]) / M[2];
                std::fill(hk.begin(), hk.end(), 0.0);  // a callback may write only its non-zero entries
                std::fill(dk.begin(), dk.end(), 0.0);
                gen_h(user, k, hk.data());
                if (!is_dipole_zero) gen_d(user, k, dk.data());
                for (double v : hk)
                    if (!std::isfinite(v)) return sbe::set_thread_error(SBE_ERR_NAN, "gen_h wrote a non-finite value");
                for (double v : dk)
                    if (!std::isfinite(v)) return sbe::set_thread_error(SBE_ERR_NAN, "gen_d wrote a non-finite value");
                for (int q = 0; q < nr; ++q) {
                    const double *R = rc.data() + 3 * q;
                    const double kr = k[0] * R[0] + k[1] * R[1] + k[2] * R[2];
                    for (int e = 0; e < nn; ++e) {
                        const int mi = e / N, ni = e % N;
                        double ph = -kr;
                        if (is_dipole_zero && tau)
                            for (int c = 0; c < 3; ++c) ph -= k[c] * (tau[3 * ni + c] - tau[3 * mi + c]);
                        const cplx w(std::cos(ph), std::sin(ph));
                        hsum[static_cast<size_t>(q) * nn + e] += w * cplx(hk[2 * e], hk[2 * e + 1]);
                        if (!is_dipole_zero)
                            for (int c = 0; c < 3; ++c)
                                psum[(static_cast<size_t>(q) * 3 + c) * nn + e] +=
                                    w * cplx(dk[2 * (c * nn + e)], dk[2 * (c * nn + e) + 1]);
                    }
                }
            }
    // owned by the model until sbe_free_sampled_material
    auto *mrv = static_cast<int32_t *>(std::malloc(sizeof(int32_t) * 3 * nr));
    auto *mw = static_cast<double *>(std::malloc(sizeof(double) * nr));
    auto *mh = static_cast<double *>(std::malloc(sizeof(double) * 2 * static_cast<size_t>(nr) * nn));
    auto *mp = static_cast<double *>(std::calloc(6 * static_cast<size_t>(nr) * nn, sizeof(double)));
    if (!mrv || !mw || !mh || !mp) {
        std::free(mrv), std::free(mw), std::free(mh), std::free(mp);
        return sbe::set_thread_error(SBE_ERR_OOM, "host allocation failed");
    }
    std::memcpy(mrv, rv.data(), sizeof(int32_t) * 3 * nr);
    for (int q = 0; q < nr; ++q) {
        mw[q] = 1.0;
        for (int e = 0; e < nn; ++e) {
            const cplx v = hsum[static_cast<size_t>(q) * nn + e] / static_cast<double>(nk);
            mh[2 * (static_cast<size_t>(q) * nn + e)] = v.real();
            mh[2 * (static_cast<size_t>(q) * nn + e) + 1] = v.imag();
            if (!is_dipole_zero)
                for (int c = 0; c < 3; ++c) {
                    const size_t i = (static_cast<size_t>(q) * 3 + c) * nn + e;
                    mp[2 * i] = (psum[i] / static_cast<double>(nk)).real();
                    mp[2 * i + 1] = (psum[i] / static_cast<double>(nk)).imag();
                }
        }
        const bool r0 = rv[3 * q] == 0 && rv[3 * q + 1] == 0 && rv[3 * q + 2] == 0;
        if (is_dipole_zero && tau && r0)  // the offsets, where the dipole-free form reads them
            for (int n = 0; n < N; ++n)
                for (int c = 0; c < 3; ++c) mp[2 * ((static_cast<size_t>(q) * 3 + c) * nn + n * N + n)] = tau[3 * n + c];
    }
    if (max_dev) {
        // the sampled series against the callbacks at k points off the grid
        // (fractions of the reciprocal axes from the additive recurrence of the
        // plastic number, which fills the unit cube evenly)
        const double g = 1.32471795724474602596, f[3] = {1.0 / g, 1.0 / (g * g), 1.0 / (g * g * g)};
        double dev = 0.0;
        std::vector<cplx> hs(nn), ds(3 * static_cast<size_t>(nn));
        for (int s = 1; s <= 16; ++s) {
            double k[3] = {0.0, 0.0, 0.0};
            for (int i = 0; i < 3; ++i) {
                const double x = std::fmod(0.5 + s * f[i], 1.0) - 0.5;
                for (int c = 0; c < 3; ++c) k[c] += x * b[3 * i + c];
            }
            std::fill(hs.begin(), hs.end(), 0.0);
            std::fill(ds.begin(), ds.end(), 0.0);
            for (int q = 0; q < nr; ++q) {
                const double *R = rc.data() + 3 * q;
                const double kr = k[0] * R[0] + k[1] * R[1] + k[2] * R[2];
                for (int e = 0; e < nn; ++e) {
                    const int mi = e / N, ni = e % N;
                    double ph = kr;
                    if (is_dipole_zero && tau)
                        for (int c = 0; c < 3; ++c) ph += k[c] * (tau[3 * ni + c] - tau[3 * mi + c]);
                    const cplx w(std::cos(ph), std::sin(ph));
                    const size_t i = static_cast<size_t>(q) * nn + e;
                    hs[e] += w * cplx(mh[2 * i], mh[2 * i + 1]);
                    if (!is_dipole_zero)
                        for (int c = 0; c < 3; ++c) {
                            const size_t j = (static_cast<size_t>(q) * 3 + c) * nn + e;
                            ds[static_cast<size_t>(c) * nn + e] += w * cplx(mp[2 * j], mp[2 * j + 1]);
                        }
                }
            }
            std::fill(hk.begin(), hk.end(), 0.0);
            gen_h(user, k, hk.data());
            for (int e = 0; e < nn; ++e) {
                const double x = std::abs(hs[e] - cplx(hk[2 * e], hk[2 * e + 1]));
                if (!(x <= dev)) dev = x;  // a NaN stays
            }
            if (!is_dipole_zero) {
                std::fill(dk.begin(), dk.end(), 0.0);
                gen_d(user, k, dk.data());
                for (int e = 0; e < 3 * nn; ++e) {
                    const double x = std::abs(ds[e] - cplx(dk[2 * e], dk[2 * e + 1]));
                    if (!(x <= dev)) dev = x;
                }
            }
        }
        *max_dev = dev;
    }
    out->kind = SBE_MODEL_WANNIER;
    out->nband = N;
    out->nval = nval;
    out->is_dipole_zero = is_dipole_zero ? 1 : 0;
    out->nrpts = nr;
    std::memcpy(out->lattice, lattice, sizeof out->lattice);
    out->rvec = mrv;
    out->rweight = mw;
    out->ham = mh;
    out->pos = mp;
    return SBE_OK;
}

void sbe_free_sampled_material(sbe_model *m)
{
    if (!m) return;
    std::free(const_cast<int32_t *>(m->rvec));
    std::free(const_cast<double *>(m->rweight));
    std::free(const_cast<double *>(m->ham));
    std::free(const_cast<double *>(m->pos));
    m->rvec = nullptr;
    m->rweight = m->ham = m->pos = nullptr;
    m->nrpts = 0;
}

}  // extern "C"
