// wannier_tb.cpp - Wannier90 seedname_tb.dat reader and writer (host only).
//
// Reads the file the way Materials/Wannier90.h:107-218 does: a header line,
// three lattice vectors (Angstrom), num_wann and nrpts, nrpts degeneracies
// (weight = 1/degeneracy), then for every R its three integer indices and
// num_wann^2 lines "m n Re Im" of <0m|H|Rn> (eV), then for every R again its
// indices and num_wann^2 lines "m n x_re x_im y_re y_im z_re z_im" of
// <0m|r|Rn> (Angstrom). Units become Hartree and bohr (constant.h:34-35).
// The reference combines the two parts of a position element as re * i*im
// (Wannier90.h:173); here they are re + i*im, the element Wannier90 wrote.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "antelope/sbe_frontend.h"

namespace {

constexpr double kAuAngstrom = 0.529177210903;  // constant.h:34
constexpr double kAuEV = 27.211386245988;       // constant.h:35

thread_local std::string g_tb_error;

int tb_fail(int code, const std::string &m)
{
    g_tb_error = m;
    return code;
}

}  // namespace

struct sbe_wannier_tb {
    int32_t nband = 0, nrpts = 0;
    double lattice[9] = {0};
    std::vector<int32_t> rvec;     // [R][3]
    std::vector<double> rweight;   // [R]
    std::vector<double> ham;       // [R][N][N][2] Hartree
    std::vector<double> pos;       // [R][3][N][N][2] bohr
    std::string header;
};

extern "C" {

const char *sbe_wannier_last_error(void) { return g_tb_error.c_str(); }

int sbe_wannier_read_tb(const char *path, sbe_wannier_tb **out)
{
    if (!path || !out) return tb_fail(SBE_ERR_CONFIG, "null argument");
    *out = nullptr;
    std::ifstream f(path);
    if (!f) return tb_fail(SBE_ERR_CONFIG, std::string("cannot open Wannier90 file ") + path);
    auto tb = new (std::nothrow) sbe_wannier_tb();
    if (!tb) return tb_fail(SBE_ERR_OOM, "host allocation failed");
    auto bad = [&](const std::string &what) {
        delete tb;
        return tb_fail(SBE_ERR_MODEL, std::string(path) + ": " + what);
    };
    std::getline(f, tb->header);
    for (int i = 0; i < 9; ++i)
        if (!(f >> tb->lattice[i])) return bad("lattice vectors");
    for (double &v : tb->lattice) v /= kAuAngstrom;
    if (!(f >> tb->nband >> tb->nrpts) || tb->nband < 1 || tb->nrpts < 1) return bad("num_wann / nrpts");
    const int N = tb->nband, nr = tb->nrpts;
    const size_t nn = static_cast<size_t>(N) * N;
    tb->rvec.assign(3 * static_cast<size_t>(nr), 0);
    tb->rweight.assign(nr, 0.0);
    tb->ham.assign(2 * nn * nr, 0.0);
    tb->pos.assign(6 * nn * nr, 0.0);
    for (int r = 0; r < nr; ++r) {
        double deg;
        if (!(f >> deg) || !(deg > 0)) return bad("degeneracies");
        tb->rweight[r] = 1.0 / deg;
    }
    std::vector<int> seen(nn);
    for (int r = 0; r < nr; ++r) {
        int32_t *R = &tb->rvec[3 * static_cast<size_t>(r)];
        if (!(f >> R[0] >> R[1] >> R[2])) return bad("R vector of block " + std::to_string(r));
        std::fill(seen.begin(), seen.end(), 0);
        for (size_t e = 0; e < nn; ++e) {
            int m, n;
            double re, im;
            if (!(f >> m >> n >> re >> im)) return bad("Hamiltonian element of block " + std::to_string(r));
            if (m < 1 || m > N || n < 1 || n > N) return bad("band index out of range");
            const size_t o = (static_cast<size_t>(r) * nn + static_cast<size_t>(m - 1) * N + (n - 1)) * 2;
            tb->ham[o] = re / kAuEV;
            tb->ham[o + 1] = im / kAuEV;
            seen[static_cast<size_t>(m - 1) * N + (n - 1)] = 1;
        }
        for (int v : seen)
            if (!v) return bad("missing Hamiltonian element in block " + std::to_string(r));
    }
    for (int r = 0; r < nr; ++r) {
        int32_t R[3];
        if (!(f >> R[0] >> R[1] >> R[2])) {
            if (r == 0 && f.eof()) {  // a file without position matrices: dipoles zero
                tb->pos.clear();
                break;
            }
            return bad("R vector of position block " + std::to_string(r));
        }
        const int32_t *H = &tb->rvec[3 * static_cast<size_t>(r)];
        if (R[0] != H[0] || R[1] != H[1] || R[2] != H[2]) return bad("position blocks not in Hamiltonian order");
        for (size_t e = 0; e < nn; ++e) {
            int m, n;
            if (!(f >> m >> n) || m < 1 || m > N || n < 1 || n > N) return bad("position element index");
            for (int ax = 0; ax < 3; ++ax) {
                double re, im;
                if (!(f >> re >> im)) return bad("position element of block " + std::to_string(r));
                const size_t o =
                    ((((static_cast<size_t>(r) * 3 + ax) * N) + (m - 1)) * N + (n - 1)) * 2;
                tb->pos[o] = re / kAuAngstrom;
                tb->pos[o + 1] = im / kAuAngstrom;
            }
        }
    }
    *out = tb;
    return SBE_OK;
}

int sbe_wannier_tb_model(const sbe_wannier_tb *tb, int32_t nval, sbe_model *m)
{
    if (!tb || !m) return tb_fail(SBE_ERR_CONFIG, "null argument");
    if (nval < 0 || nval > tb->nband) return tb_fail(SBE_ERR_MODEL, "Nval outside 0..num_wann");
    *m = sbe_model{};
    m->kind = SBE_MODEL_WANNIER;
    m->nband = tb->nband;
    m->nval = nval;
    m->is_dipole_zero = tb->pos.empty() ? 1 : 0;
    m->nrpts = tb->nrpts;
    for (int i = 0; i < 9; ++i) m->lattice[i] = tb->lattice[i];
    m->rvec = tb->rvec.data();
    m->rweight = tb->rweight.data();
    m->ham = tb->ham.data();
    m->pos = tb->pos.empty() ? nullptr : tb->pos.data();
    return SBE_OK;
}

void sbe_wannier_tb_free(sbe_wannier_tb *tb) { delete tb; }

int sbe_wannier_write_tb(const char *path, const sbe_model *m, const char *header)
{
    if (!path || !m || m->kind != SBE_MODEL_WANNIER || !m->rvec || !m->rweight || !m->ham)
        return tb_fail(SBE_ERR_CONFIG, "need a Wannier model with rvec, rweight and ham");
    FILE *f = std::fopen(path, "w");
    if (!f) return tb_fail(SBE_ERR_CONFIG, std::string("cannot write ") + path);
    const int N = m->nband, nr = m->nrpts;
    const size_t nn = static_cast<size_t>(N) * N;
    std::fprintf(f, "%s\n", header ? header : "written by antelope_amd");
    for (int i = 0; i < 3; ++i)
        std::fprintf(f, "%.16e %.16e %.16e\n", m->lattice[3 * i] * kAuAngstrom, m->lattice[3 * i + 1] * kAuAngstrom,
                     m->lattice[3 * i + 2] * kAuAngstrom);
    std::fprintf(f, "%d\n%d\n", N, nr);
    for (int r = 0; r < nr; ++r)
        std::fprintf(f, "%d%s", static_cast<int>(std::lround(1.0 / m->rweight[r])), (r % 15 == 14 || r == nr - 1) ? "\n" : " ");
    for (int r = 0; r < nr; ++r) {
        const int32_t *R = m->rvec + 3 * static_cast<size_t>(r);
        std::fprintf(f, "\n%d %d %d\n", R[0], R[1], R[2]);
        for (int n = 0; n < N; ++n)  // Wannier90's order: m fastest
            for (int mm = 0; mm < N; ++mm) {
                const double *h = m->ham + (static_cast<size_t>(r) * nn + static_cast<size_t>(mm) * N + n) * 2;
                std::fprintf(f, "%d %d %.16e %.16e\n", mm + 1, n + 1, h[0] * kAuEV, h[1] * kAuEV);
            }
    }
    if (m->pos) {
        for (int r = 0; r < nr; ++r) {
            const int32_t *R = m->rvec + 3 * static_cast<size_t>(r);
            std::fprintf(f, "\n%d %d %d\n", R[0], R[1], R[2]);
            for (int n = 0; n < N; ++n)
                for (int mm = 0; mm < N; ++mm) {
                    std::fprintf(f, "%d %d", mm + 1, n + 1);
                    for (int ax = 0; ax < 3; ++ax) {
                        const double *p = m->pos + ((((static_cast<size_t>(r) * 3 + ax) * N) + mm) * N + n) * 2;
                        std::fprintf(f, " %.16e %.16e", p[0] * kAuAngstrom, p[1] * kAuAngstrom);
                    }
                    std::fprintf(f, "\n");
                }
        }
    }
    const bool ok = std::fclose(f) == 0;
    return ok ? SBE_OK : tb_fail(SBE_ERR_CONFIG, std::string("write error on ") + path);
}

}  // extern "C"
