// spectrum.cpp - the HHG spectrum of the k-summed current, as the
// reference's post-processing script computes it
// (src/GeneralModel/AlsisPythonShort/AnalysisReader_PyDir.py):
//   :163-167  w = linspace(-pi/dt, pi/dt - dw, Nt), dw = 2 pi / (Nt dt)
//   :249-253  a Blackman window over the whole series, per axis
//   :323-330  F_a = -i w fftshift(fft(window j_a) dt), inter and intra parts
//             transformed separately and added
//   :332-333  S_inter, S_intra = sum_a |F_a|^2 + hzero (hzero = 0.98e-25, :118)
//   :354      S = |F_x|^2 + |F_y|^2 + hzero (Dim = 2 there)
//   :367-382  the half w[Nt/2 : Nt-1] / w0 (harmonic order) written as
//             w/w0, Re F_a, Im F_a per axis, 10^log10(S), format %1.16e.
// Host-only. The discrete Fourier transform is an iterative radix-2 FFT for
// power-of-two lengths and Bluestein's chirp-z form (one radix-2 convolution
// of length >= 2 Nt - 1) for the others; twiddles and chirps are evaluated
// directly (the chirp index n^2 reduced mod 2 Nt in integers), so the
// transform keeps double precision at any length.
#include <cmath>
#include <complex>
#include <cstdio>
#include <string>
#include <vector>

#include "antelope/sbe_frontend.h"

namespace {

using cd = std::complex<double>;

thread_local std::string g_spec_error;

int spec_fail(int rc, const std::string &msg)
{
    g_spec_error = msg;
    return rc;
}

// In-place radix-2 DIT FFT of a power-of-two length, e^{-2 pi i m n / N}.
void fft_pow2(std::vector<cd> &a, bool inverse)
{
    const size_t n = a.size();
    for (size_t i = 1, j = 0; i < n; ++i) {  // bit reversal
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    const double sgn = inverse ? 1.0 : -1.0;
    std::vector<cd> tw(n / 2 > 0 ? n / 2 : 1);
    for (size_t len = 2; len <= n; len <<= 1) {
        const size_t half = len / 2;
        for (size_t k = 0; k < half; ++k) {
            const double ang = sgn * 2.0 * M_PI * static_cast<double>(k) / static_cast<double>(len);
            tw[k] = cd(std::cos(ang), std::sin(ang));
        }
        for (size_t i = 0; i < n; i += len)
            for (size_t k = 0; k < half; ++k) {
                const cd u = a[i + k], v = a[i + k + half] * tw[k];
                a[i + k] = u + v;
                a[i + k + half] = u - v;
            }
    }
}

// X[m] = sum_n x[n] e^{-2 pi i m n / N}, any N >= 1.
std::vector<cd> dft(const std::vector<double> &x)
{
    const size_t n = x.size();
    if ((n & (n - 1)) == 0) {
        std::vector<cd> a(x.begin(), x.end());
        fft_pow2(a, false);
        return a;
    }
    // Bluestein: m n = (m^2 + n^2 - (m - n)^2) / 2, so X[m] = c*[m] sum_n (x[n] c*[n]) c[m - n]
    // with the chirp c[n] = e^{i pi n^2 / N}
    size_t L = 1;
    while (L < 2 * n - 1) L <<= 1;
    std::vector<cd> chirp(n);
    const unsigned long long twoN = 2ull * n;
    for (size_t i = 0; i < n; ++i) {
        const unsigned long long q = (static_cast<unsigned long long>(i) * i) % twoN;
        const double ang = M_PI * static_cast<double>(q) / static_cast<double>(n);
        chirp[i] = cd(std::cos(ang), std::sin(ang));
    }
    std::vector<cd> a(L, cd(0, 0)), b(L, cd(0, 0));
    for (size_t i = 0; i < n; ++i) a[i] = x[i] * std::conj(chirp[i]);
    b[0] = chirp[0];
    for (size_t i = 1; i < n; ++i) b[i] = b[L - i] = chirp[i];
    fft_pow2(a, false);
    fft_pow2(b, false);
    for (size_t i = 0; i < L; ++i) a[i] *= b[i];
    fft_pow2(a, true);
    std::vector<cd> X(n);
    const double inv = 1.0 / static_cast<double>(L);
    for (size_t m = 0; m < n; ++m) X[m] = std::conj(chirp[m]) * a[m] * inv;
    return X;
}

// numpy.blackman(M): 0.42 - 0.5 cos(2 pi n / (M - 1)) + 0.08 cos(4 pi n / (M - 1))
std::vector<double> blackman(size_t M)
{
    std::vector<double> w(M, 1.0);
    if (M < 2) return w;
    for (size_t n = 0; n < M; ++n) {
        const double x = static_cast<double>(n) / static_cast<double>(M - 1);
        w[n] = 0.42 - 0.5 * std::cos(2.0 * M_PI * x) + 0.08 * std::cos(4.0 * M_PI * x);
    }
    return w;
}

// numpy.linspace(-pi/dt, pi/dt - dw, Nt) with dw = 2 pi / (Nt dt)
std::vector<double> freq_axis(size_t nt, double dt)
{
    const double wmax = M_PI / dt, dw = 2.0 * wmax / static_cast<double>(nt);
    const double start = -wmax, stop = wmax - dw;
    std::vector<double> w(nt, start);
    if (nt < 2) return w;
    const double step = (stop - start) / static_cast<double>(nt - 1);
    for (size_t i = 0; i < nt; ++i) w[i] = static_cast<double>(i) * step + start;
    w[nt - 1] = stop;
    return w;
}

constexpr double kHzero = 0.98e-25;

}  // namespace

extern "C" {

int64_t sbe_hhg_points(int64_t nt) { return nt >= 2 ? nt - nt / 2 - 1 : 0; }

int sbe_hhg_spectrum(int64_t nt, const double *j_inter, const double *j_intra, double dt, double w0, int32_t dim,
                     double *order, double *F, double *S, double *S_inter, double *S_intra)
{
    if (nt < 2) return spec_fail(SBE_ERR_RANGE, "sbe_hhg_spectrum: need at least 2 samples");
    if (dim < 1 || dim > 3) return spec_fail(SBE_ERR_RANGE, "sbe_hhg_spectrum: dim must be 1, 2 or 3");
    if (!(dt > 0.0) || !(w0 > 0.0)) return spec_fail(SBE_ERR_CONFIG, "sbe_hhg_spectrum: dt and w0 must be > 0");
    if (!j_inter || !j_intra) return spec_fail(SBE_ERR_CONFIG, "sbe_hhg_spectrum: null current series");
    const size_t n = static_cast<size_t>(nt);
    const size_t half = n / 2, nout = n - half - 1;
    const std::vector<double> win = blackman(n), w = freq_axis(n, dt);
    std::vector<double> s(nout, kHzero), si(nout, kHzero), sa(nout, kHzero), x(n);
    for (int ax = 0; ax < dim; ++ax) {
        std::vector<cd> Fp[2];
        const double *src[2] = {j_inter, j_intra};
        for (int part = 0; part < 2; ++part) {
            for (size_t i = 0; i < n; ++i) x[i] = win[i] * src[part][3 * i + ax];
            const std::vector<cd> X = dft(x);
            Fp[part].resize(nout);
            // fftshift: shifted[i] = X[(i - n/2) mod n]; only i in [n/2, n-1) is kept
            for (size_t o = 0; o < nout; ++o) {
                const size_t i = half + o;
                const cd v = X[(i + n - half) % n] * dt;
                Fp[part][o] = cd(0.0, -1.0) * w[i] * v;
            }
        }
        for (size_t o = 0; o < nout; ++o) {
            const cd f = Fp[0][o] + Fp[1][o];
            s[o] += std::norm(f);
            si[o] += std::norm(Fp[0][o]);
            sa[o] += std::norm(Fp[1][o]);
            if (F) {
                F[2 * (o * dim + ax)] = f.real();
                F[2 * (o * dim + ax) + 1] = f.imag();
            }
        }
    }
    for (size_t o = 0; o < nout; ++o) {
        if (order) order[o] = w[half + o] / w0;
        if (S) S[o] = s[o];
        if (S_inter) S_inter[o] = si[o];
        if (S_intra) S_intra[o] = sa[o];
    }
    return SBE_OK;
}

int sbe_write_hhg(const char *path, int64_t nt, const double *j_inter, const double *j_intra, double dt, double w0,
                  int32_t dim)
{
    if (!path) return spec_fail(SBE_ERR_CONFIG, "sbe_write_hhg: null path");
    const int64_t nout = sbe_hhg_points(nt);
    std::vector<double> order(nout), F(2 * nout * (dim > 0 ? dim : 1)), S(nout);
    int rc = sbe_hhg_spectrum(nt, j_inter, j_intra, dt, w0, dim, order.data(), F.data(), S.data(), nullptr, nullptr);
    if (rc != SBE_OK) return rc;
    FILE *f = std::fopen(path, "w");
    if (!f) return spec_fail(SBE_ERR_CONFIG, std::string("sbe_write_hhg: cannot open ") + path);
    for (int64_t o = 0; o < nout; ++o) {
        std::fprintf(f, "%1.16e", order[o]);
        for (int ax = 0; ax < dim; ++ax)
            std::fprintf(f, " %1.16e %1.16e", F[2 * (o * dim + ax)], F[2 * (o * dim + ax) + 1]);
        std::fprintf(f, " %1.16e\n", std::pow(10.0, std::log10(S[o])));  // the script's 10^log10(S)
    }
    const bool ok = std::fclose(f) == 0;
    return ok ? SBE_OK : spec_fail(SBE_ERR_CONFIG, std::string("sbe_write_hhg: write failed: ") + path);
}

const char *sbe_spectrum_last_error(void) { return g_spec_error.c_str(); }

}  // extern "C"
