/* sbe_material.hpp - C++ adapter from a material class with the reference's
 * material interface (src/GeneralModel/BaseMaterial.h:22-78) to the GPU path.
 *
 * The reference evaluates a material's virtual GenHamiltonian / GenDipole /
 * GenJMatrix per k point inside its CPU loop. Here a material becomes data
 * once, on the host: antelope::SampledMaterial samples GenHamiltonian (and
 * GenDipole unless isDipoleZero) with sbe_sample_material into the
 * tight-binding blocks of an SBE_MODEL_WANNIER model, which sbe_create runs
 * on the GPU. The current J = -dH/dk + i[D, H] that GenJMatrix returns
 * follows from the blocks on the device (the Wannier path's current).
 *
 * Any class with the reference's members works, an antelope::BaseMaterial
 * subclass or one of the reference's own material classes unchanged:
 *   int Nband, Nval; bool isDipoleZero;
 *   void GenHamiltonian(std::complex<double> *h, std::array<double, 3> k);      // [N][N] row-major
 *   void GenDipole(std::complex<double> **d, std::array<double, 3> k);          // d[i] -> [N][N]
 *   std::tuple<std::array<double, 9>, std::array<double, 3>> GenBrillouinzone();  // (optional here)
 * The output buffers are zeroed before every call, so a class that writes
 * only its non-zero entries (or nothing, as a dipole-free class's GenDipole)
 * is read correctly.
 *
 *   Haldane mat(...);
 *   antelope::SampledMaterial s(mat, lattice, {1, 1, 0}, {{0, 0, 0}, {0, -a0, 0}});
 *   if (s.max_deviation() > 1e-12) ...            // hoppings reach further: raise nmax
 *   sbe_config cfg = ...;  antelope::apply_brillouin_zone(mat, cfg);
 *   sbe_create(&cfg, &s.model(), -1, -1, &ctx);
 *
 * The names avoid `complex` and `Ndim`, which the reference's constant.h
 * defines as a macro and a global constant. */
#ifndef ANTELOPE_SBE_MATERIAL_HPP
#define ANTELOPE_SBE_MATERIAL_HPP

#include <array>
#include <complex>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "sbe_c.h"

namespace antelope {

/* The reference's material interface, member for member (BaseMaterial.h:22-78;
 * GenInitialValue, the custom initial state, has no counterpart on the GPU
 * path and is left out). */
class BaseMaterial {
public:
    int Nband = 0;
    int Nval = 0;
    bool isDipoleZero = false;
    virtual ~BaseMaterial() {}
    virtual void GenHamiltonian(std::complex<double> *hstore, std::array<double, 3> kpoint) = 0;
    virtual void GenDipole(std::complex<double> **dstore, std::array<double, 3> kpoint) = 0;
    virtual void GenJMatrix(std::complex<double> **jstore, std::array<double, 3> kpoint) = 0;
    virtual std::tuple<std::array<double, 9>, std::array<double, 3>> GenBrillouinzone() = 0;
    virtual void PrintMaterialInformation() = 0;
};

namespace detail {

template <class Material>
struct Trampoline {
    Material *mat;
    int nband;
    std::exception_ptr error;

    static void hamiltonian(void *user, const double k[3], double *out)
    {
        auto *t = static_cast<Trampoline *>(user);
        const size_t nn = static_cast<size_t>(t->nband) * t->nband;
        std::memset(out, 0, sizeof(double) * 2 * nn);
        if (t->error) return;
        try {
            t->mat->GenHamiltonian(reinterpret_cast<std::complex<double> *>(out), std::array<double, 3>{k[0], k[1], k[2]});
        } catch (...) {
            t->error = std::current_exception();
        }
    }

    static void dipole(void *user, const double k[3], double *out)
    {
        auto *t = static_cast<Trampoline *>(user);
        const size_t nn = static_cast<size_t>(t->nband) * t->nband;
        std::memset(out, 0, sizeof(double) * 6 * nn);
        if (t->error) return;
        auto *c = reinterpret_cast<std::complex<double> *>(out);
        std::complex<double> *rows[3] = {c, c + nn, c + 2 * nn};
        try {
            t->mat->GenDipole(rows, std::array<double, 3>{k[0], k[1], k[2]});
        } catch (...) {
            t->error = std::current_exception();
        }
    }
};

}  // namespace detail

/* Tight-binding blocks sampled from a material; owns the sbe_model's arrays. */
class SampledMaterial {
public:
    /* lattice: rows a1, a2, a3 (bohr; a 2D material still gives an a3, e.g.
     * {0, 0, 10}); nmax: hopping range along each a_i; tau: orbital offsets
     * (bohr) of a dipole-free material whose H(k) carries the phases of the
     * bonds (Haldane.h:86-92), empty = 0. Throws std::runtime_error with the
     * library's message, or rethrows what the material threw. */
    template <class Material>
    SampledMaterial(Material &mat, const std::array<double, 9> &lattice, const std::array<int, 3> &nmax,
                    const std::vector<std::array<double, 3>> &tau = {})
    {
        detail::Trampoline<Material> t{&mat, mat.Nband, nullptr};
        if (!tau.empty() && static_cast<int>(tau.size()) != mat.Nband)
            throw std::runtime_error("sbe_material: tau needs one offset per band");
        std::vector<double> tv;
        for (const auto &p : tau) tv.insert(tv.end(), p.begin(), p.end());
        const int32_t nm[3] = {nmax[0], nmax[1], nmax[2]};
        const int rc = sbe_sample_material(mat.Nband, mat.Nval, mat.isDipoleZero ? 1 : 0, lattice.data(),
                                           tv.empty() ? nullptr : tv.data(), nm,
                                           &detail::Trampoline<Material>::hamiltonian,
                                           mat.isDipoleZero ? nullptr : &detail::Trampoline<Material>::dipole, &t,
                                           &model_, &max_dev_);
        if (t.error) {
            if (rc == SBE_OK) sbe_free_sampled_material(&model_);
            std::rethrow_exception(t.error);
        }
        if (rc != SBE_OK) throw std::runtime_error(std::string("sbe_sample_material: ") + sbe_last_error(nullptr));
    }
    ~SampledMaterial() { sbe_free_sampled_material(&model_); }
    SampledMaterial(const SampledMaterial &) = delete;
    SampledMaterial &operator=(const SampledMaterial &) = delete;
    SampledMaterial(SampledMaterial &&o) noexcept : model_(o.model_), max_dev_(o.max_dev_)
    {
        std::memset(&o.model_, 0, sizeof o.model_);
    }

    const sbe_model &model() const { return model_; }
    /* largest |H_sampled(k) - H(k)| (and |D_sampled - D|) at 16 k points off the
     * sampling grid: rounding for a material whose hoppings lie within nmax */
    double max_deviation() const { return max_dev_; }

private:
    sbe_model model_{};
    double max_dev_ = 0.0;
};

/* The material's Brillouin zone (GenBrillouinzone) as the run's k box,
 * each axis scaled by cfg.ksfactor, as SBEs.h:219-240 takes it; calc.BZaxes /
 * calc.BZorigin already given in cfg win. Without this call the sampled
 * model's zone is the reciprocal cell of its lattice. */
template <class Material>
void apply_brillouin_zone(Material &mat, sbe_config &cfg)
{
    const auto zone = mat.GenBrillouinzone();
    if (!cfg.has_bz_axes) {
        for (int i = 0; i < 9; ++i) cfg.bz_axes[i] = std::get<0>(zone)[i] * cfg.ksfactor[i / 3];
        cfg.has_bz_axes = 1;
    }
    if (!cfg.has_bz_origin) {
        for (int i = 0; i < 3; ++i) cfg.bz_origin[i] = std::get<1>(zone)[i];
        cfg.has_bz_origin = 1;
    }
}

}  // namespace antelope

#endif /* ANTELOPE_SBE_MATERIAL_HPP */
