// frontend.cpp - the caller side of the C ABI that the reference's driver
// provides: reading inputParam.cfg into sbe_config / sbe_model and writing
// the reference's output files. Host-only code (no HIP calls).
//
// Config reading restates antelope.cpp:90-138 (file, gauge, Npoints),
// SBEs.h:125-215 (target, initial value type), SBEs.h:432-561 (calc and
// laser groups, defaults, degree conversion) and Materials/Haldane.h:46-84
// (a0 in Angstrom, M0 or Mt2). The parser accepts the libconfig grammar the
// reference's files use: settings `name = value;` or `name : value;`,
// groups {...}, arrays [...], lists (...), numbers (int, hex, float),
// booleans, strings (adjacent literals concatenate) and #, //, /* */
// comments. Writers restate antelope.cpp:208-235 (outlaserdata.dat) and
// :514-548 (current, occupation and coherence files).
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "antelope/sbe_c.h"
#include "antelope/sbe_frontend.h"
#include "frontend_internal.hpp"
#include "sbe_internal.hpp"

namespace {

thread_local std::string g_fe_error;

int fe_fail(int code, const std::string &msg)
{
    g_fe_error = msg;
    return code;
}

// ------------------------------------------------------------ settings --
struct Setting {
    enum Kind { Int, Float, Bool, String, Array, List, Group } kind = Group;
    std::string name;
    long long i = 0;
    double f = 0.0;
    bool b = false;
    std::string s;
    std::vector<Setting> items;  // Array, List, Group

    const Setting *child(const std::string &n) const
    {
        if (kind != Group) return nullptr;
        for (const Setting &c : items)
            if (c.name == n) return &c;
        return nullptr;
    }
    bool number(double &out) const
    {
        if (kind == Float) { out = f; return true; }
        if (kind == Int) { out = static_cast<double>(i); return true; }
        return false;
    }
};

class Parser {
  public:
    explicit Parser(const std::string &text) : t_(text) {}
    bool parse(Setting &root, std::string &err)
    {
        root.kind = Setting::Group;
        if (!settings(root.items, '\0')) {
            err = "parse error at line " + std::to_string(line_) + ": " + err_;
            return false;
        }
        return true;
    }

  private:
    const std::string &t_;
    size_t p_ = 0;
    int line_ = 1;
    std::string err_;

    bool fail(const std::string &m)
    {
        if (err_.empty()) err_ = m;
        return false;
    }
    void ws()
    {
        for (;;) {
            while (p_ < t_.size() && std::isspace(static_cast<unsigned char>(t_[p_]))) {
                if (t_[p_] == '\n') ++line_;
                ++p_;
            }
            if (p_ < t_.size() && t_[p_] == '#') {
                while (p_ < t_.size() && t_[p_] != '\n') ++p_;
            } else if (t_.compare(p_, 2, "//") == 0) {
                while (p_ < t_.size() && t_[p_] != '\n') ++p_;
            } else if (t_.compare(p_, 2, "/*") == 0) {
                const size_t e = t_.find("*/", p_ + 2);
                for (size_t q = p_; q < (e == std::string::npos ? t_.size() : e); ++q)
                    if (t_[q] == '\n') ++line_;
                p_ = e == std::string::npos ? t_.size() : e + 2;
            } else {
                return;
            }
        }
    }
    bool settings(std::vector<Setting> &out, char close)
    {
        for (;;) {
            ws();
            if (p_ >= t_.size()) return close == '\0' ? true : fail("unexpected end of file");
            if (t_[p_] == close) return true;
            if (t_[p_] == '@') return fail("@include is not supported");
            Setting s;
            if (!name(s.name)) return false;
            ws();
            if (p_ >= t_.size() || (t_[p_] != '=' && t_[p_] != ':')) return fail("expected '=' or ':' after " + s.name);
            ++p_;
            if (!value(s)) return false;
            ws();
            if (p_ < t_.size() && (t_[p_] == ';' || t_[p_] == ',')) ++p_;
            out.push_back(std::move(s));
        }
    }
    bool name(std::string &n)
    {
        const size_t b = p_;
        if (p_ >= t_.size() || !(std::isalpha(static_cast<unsigned char>(t_[p_])) || t_[p_] == '*'))
            return fail("expected a setting name");
        while (p_ < t_.size() && (std::isalnum(static_cast<unsigned char>(t_[p_])) || t_[p_] == '_' ||
                                  t_[p_] == '-' || t_[p_] == '*'))
            ++p_;
        n = t_.substr(b, p_ - b);
        return true;
    }
    bool value(Setting &s)
    {
        ws();
        if (p_ >= t_.size()) return fail("missing value");
        const char c = t_[p_];
        if (c == '{') {
            ++p_;
            s.kind = Setting::Group;
            if (!settings(s.items, '}')) return false;
            ++p_;
            return true;
        }
        if (c == '[' || c == '(') {
            const char close = c == '[' ? ']' : ')';
            s.kind = c == '[' ? Setting::Array : Setting::List;
            ++p_;
            for (;;) {
                ws();
                if (p_ >= t_.size()) return fail("unterminated array/list");
                if (t_[p_] == close) {
                    ++p_;
                    return true;
                }
                Setting v;
                if (!value(v)) return false;
                s.items.push_back(std::move(v));
                ws();
                if (p_ < t_.size() && t_[p_] == ',') ++p_;
            }
        }
        if (c == '"') {
            s.kind = Setting::String;
            while (p_ < t_.size() && t_[p_] == '"') {
                ++p_;
                while (p_ < t_.size() && t_[p_] != '"') {
                    if (t_[p_] == '\\' && p_ + 1 < t_.size()) {
                        const char e = t_[++p_];
                        s.s += e == 'n' ? '\n' : e == 't' ? '\t' : e;
                    } else {
                        s.s += t_[p_];
                    }
                    ++p_;
                }
                if (p_ >= t_.size()) return fail("unterminated string");
                ++p_;
                ws();  // adjacent literals concatenate
            }
            return true;
        }
        // boolean or number
        size_t b = p_;
        while (p_ < t_.size() && (std::isalnum(static_cast<unsigned char>(t_[p_])) || t_[p_] == '.' || t_[p_] == '+' ||
                                  t_[p_] == '-'))
            ++p_;
        std::string tok = t_.substr(b, p_ - b);
        std::string low;
        for (char ch : tok) low += static_cast<char>(std::tolower(static_cast<unsigned char>(ch)));
        if (low == "true" || low == "false") {
            s.kind = Setting::Bool;
            s.b = low == "true";
            return true;
        }
        if (tok.empty()) return fail("bad value");
        std::string num = tok;
        while (!num.empty() && (num.back() == 'L' || num.back() == 'l')) num.pop_back();
        const bool hex = num.size() > 2 && num[0] == '0' && (num[1] == 'x' || num[1] == 'X');
        const bool is_float = !hex && num.find_first_of(".eE") != std::string::npos;
        char *end = nullptr;
        if (is_float) {
            s.kind = Setting::Float;
            s.f = std::strtod(num.c_str(), &end);
        } else {
            s.kind = Setting::Int;
            s.i = std::strtoll(num.c_str(), &end, hex ? 16 : 10);
        }
        if (!end || *end != '\0') return fail("bad number '" + tok + "'");
        return true;
    }
};

// -------------------------------------------------------------- lookups --
bool get_double(const Setting *g, const char *n, double &v)
{
    const Setting *s = g ? g->child(n) : nullptr;
    return s && s->number(v);
}
bool get_int(const Setting *g, const char *n, long long &v)
{
    const Setting *s = g ? g->child(n) : nullptr;
    if (!s || s->kind != Setting::Int) return false;
    v = s->i;
    return true;
}
bool get_bool(const Setting *g, const char *n, bool &v)
{
    const Setting *s = g ? g->child(n) : nullptr;
    if (!s || s->kind != Setting::Bool) return false;
    v = s->b;
    return true;
}
bool get_string(const Setting *g, const char *n, std::string &v)
{
    const Setting *s = g ? g->child(n) : nullptr;
    if (!s || s->kind != Setting::String) return false;
    v = s->s;
    return true;
}

constexpr double kPi = 3.14159265358979323846;
constexpr double kAuAngstrom = 0.529177210903;  // constant.h:34

}  // namespace


namespace {
constexpr double kAuEV = 27.211386245988;  // constant.h:35
thread_local std::string t_base_dir;      // directory of the config file being parsed

// GGA parameter sets of the three-band TMDC model (G.-B. Liu et al., Phys.
// Rev. B 88, 085433 (2013)) as TMDCs.h:69-115 hard-codes them: lattice
// constant (Angstrom), spin-orbit coupling and {e1, e2, t0, t1, t2, t11, t12,
// t22, r0, r1, r2, r11, r12, u0, u1, u2, u11, u12, u22} (eV).
struct TmdcRow {
    const char *name;
    double a_ang, soc_ev;
    double v[19];
};
const TmdcRow kTmdc[6] = {
    {"MoS2", 3.190, 0.073, {0.683, 1.707, -0.146, -0.114, 0.506, 0.085, 0.162, 0.073, 0.060, -0.236,
                            0.067, 0.016, 0.087, -0.038, 0.046, 0.001, 0.266, -0.176, -0.150}},
    {"WS2", 3.191, 0.211, {0.717, 1.916, -0.152, -0.097, 0.590, 0.047, 0.178, 0.016, 0.069, -0.261,
                           0.107, -0.003, 0.109, -0.054, 0.045, 0.002, 0.325, -0.206, -0.163}},
    {"MoSe2", 3.326, 0.091, {0.684, 1.546, -0.146, -0.130, 0.432, 0.144, 0.117, 0.075, 0.039, -0.209,
                             0.069, 0.052, 0.060, -0.042, 0.036, 0.008, 0.272, -0.172, -0.150}},
    {"WSe2", 3.325, 0.228, {0.728, 1.655, -0.146, -0.124, 0.507, 0.117, 0.127, 0.015, 0.036, -0.234,
                            0.107, 0.044, 0.075, -0.061, 0.032, 0.007, 0.329, -0.202, -0.164}},
    {"MoTe2", 3.557, 0.107, {0.588, 1.303, -0.226, -0.234, 0.036, 0.400, 0.098, 0.017, 0.003, -0.025,
                             -0.169, 0.082, 0.051, 0.057, 0.103, 0.187, -0.045, -0.141, 0.087}},
    {"WTe2", 3.560, 0.237, {0.697, 1.380, -0.109, -0.164, 0.368, 0.204, 0.093, 0.038, -0.015, -0.209,
                            0.107, 0.115, 0.009, -0.066, 0.011, -0.013, 0.312, -0.177, -0.132}},
};

// The analytic classes' constructors (keys, Angstrom conversion, band counts):
// WilsonMass.h:40-63, SquareSBand.h:39-59, KaneMele.h:47-75, XYHlattice.h:49-73,
// TMDCs.h:48-149 and the ones cited below.
int analytic_target(const std::string &name, const Setting *g, sbe_model &m)
{
    auto need = [&](const char *key, double &v) { return get_double(g, key, v); };
    m.is_dipole_zero = 1;
    double *p = m.params;
    if (name == "WilsonMass") {
        if (!(need("t", p[0]) && need("delta", p[1]) && need("mu", p[2]) && need("a0", p[3])))
            return fe_fail(SBE_ERR_MODEL, "Some WilsonMass paramters are missing");
        if (!need("a1", p[4])) p[4] = p[3];  // a.u., no Angstrom conversion in this class
        m.kind = SBE_MODEL_WILSON_MASS;
        m.nband = 2;
        m.nval = 1;
    } else if (name == "SquareSBand") {
        if (!(need("t1", p[0]) && need("M0", p[1]) && need("a0", p[2])))
            return fe_fail(SBE_ERR_MODEL, "Some SquareSBand paramters are missing");
        p[2] /= kAuAngstrom;
        m.kind = SBE_MODEL_SQUARE_SBAND;
        m.nband = 2;
        m.nval = 1;
    } else if (name == "KaneMele") {
        double mt2;
        if (!(need("t1", p[0]) && need("t2", p[1]) && need("tR", p[2]) && need("Mt2", mt2) && need("phi0", p[4]) &&
              need("a0", p[5])))
            return fe_fail(SBE_ERR_MODEL, "Some KaneMele paramters are missing");
        p[3] = mt2 * p[1];
        p[5] /= kAuAngstrom;
        m.kind = SBE_MODEL_KANE_MELE;
        m.nband = 4;
        m.nval = 2;
    } else if (name == "XYHlattice") {
        if (!(need("tl", p[0]) && need("tr", p[1]) && need("lambda", p[2]) && need("a0", p[3])))
            return fe_fail(SBE_ERR_MODEL, "Some XYHlattice paramters are missing");
        p[3] /= kAuAngstrom;
        m.kind = SBE_MODEL_XYH_LATTICE;
        m.nband = 4;
        m.nval = 2;
    } else if (name == "Haldane2L") {  // Haldane2L.h:52-148
        double *q = p;  // t1, t2, M0, phi0 (top), the same (bottom), a0, same, s0, s1, h0..h3
        double t11;
        std::string poly;
        if (!(need("t1", q[0]) && need("t2", q[1]) && need("phi0", q[3]) && need("a0", q[8]) && need("t11", t11) &&
              get_string(g, "polytype", poly)))
            return fe_fail(SBE_ERR_MODEL, "Some Haldane2L paramters are missing");
        q[8] /= kAuAngstrom;
        if (!need("M0", q[2])) {
            if (!need("Mt2", q[2])) return fe_fail(SBE_ERR_MODEL, "M0 param missing");
            q[2] *= q[1];
        }
        q[4] = q[0]; q[5] = q[1]; q[6] = q[2]; q[7] = q[3];
        bool same = true;
        // the class reads the first of these keys that exists and skips the rest
        // (a short-circuit ||, Haldane2L.h:83-88)
        if (need("t1_bottom", q[4]) || need("t2_bottom", q[5]) || need("phi0_bottom", q[7])) same = false;
        if (need("M0_bottom", q[6])) {
            same = false;
        } else if (need("Mt2_bottom", q[6])) {
            q[6] *= q[5];
            same = false;
        }
        q[9] = same ? 1.0 : 0.0;
        q[10] = q[11] = 1.0;
        double h[4] = {0, 0, 0, 0};
        if (poly == "AA") { h[0] = t11; h[3] = t11; }  // stackings of Phys. Rev. B 83, 235312
        else if (poly == "AB" || poly == "AB1") { h[1] = t11; }
        else if (poly == "AB2" || poly == "3R") { h[2] = t11; }
        else if (poly == "AA'" || poly == "2H") { h[1] = t11; h[2] = t11; q[11] = -1.0; }
        else if (poly == "A'B") { h[3] = t11; q[10] = -1.0; }
        else if (poly == "AB'") { h[0] = t11; q[11] = -1.0; }
        else return fe_fail(SBE_ERR_MODEL, "Undefined polytype");
        for (int i = 0; i < 4; ++i) q[12 + i] = h[i];
        m.kind = SBE_MODEL_HALDANE_2L;
        m.nband = 4;
        m.nval = 2;
    } else if (name == "Bi2Se3surf") {  // Bi2Se3surf.h:51-82: no parameters
        m.kind = SBE_MODEL_BI2SE3_SURF;
        m.nband = 2;
        m.nval = 1;
    } else if (name == "MultiWSM") {  // MultiWSM.h:47-74
        if (!(need("n", p[0]) && need("v", p[1]) && need("alphan", p[2]) && need("phi0", p[3]) && need("a0", p[4])))
            return fe_fail(SBE_ERR_MODEL, "Some MultiWSM paramters are missing");
        p[4] /= kAuAngstrom;
        m.kind = SBE_MODEL_MULTI_WSM;
        m.nband = 2;
        m.nval = 1;
    } else if (name == "WDSM_TaAs") {  // WDSM_TaAs.h:41-67
        if (!(need("my", p[2]) && need("mz", p[3]) && need("t", p[0]) && need("delta", p[1]) && need("a0", p[4])))
            return fe_fail(SBE_ERR_MODEL, "Some WDSM paramters are missing");
        p[4] /= kAuAngstrom;
        m.kind = SBE_MODEL_WDSM_TAAS;
        m.nband = 4;
        m.nval = 1;
    } else {  // TMDCs
        bool soc = false;
        get_bool(g, "isSOCEnabled", soc);
        std::string mat;
        get_string(g, "Material", mat);
        if (sbe_tmdc_params(mat.c_str(), soc ? 1 : 0, &m) != SBE_OK)
            return fe_fail(SBE_ERR_MODEL, "Unknowm material: " + mat);
    }
    return SBE_OK;
}
}  // namespace

extern "C" {

const char *sbe_frontend_last_error(void) { return g_fe_error.c_str(); }

int32_t sbe_spec_shot_number(const sbe_spec *spec) { return spec ? spec->shot_number : 0; }

int64_t sbe_snapshot_interval(const sbe_config *cfg, int32_t shot_number)
{
    // antelope.cpp:273-278: int(period0 / shotNumber / dt) with period0 = 2 pi / w0 of
    // pulse 0; no snapshots (Nt + 100) when shotNumber <= 0
    if (!cfg || shot_number <= 0 || cfg->npulses < 1 || !(cfg->pulses[0].w0 > 0) || !(cfg->dt > 0)) return 0;
    const double period0 = 2.0 * kPi / cfg->pulses[0].w0;
    return static_cast<int64_t>(period0 / shot_number / cfg->dt);
}

int sbe_write_snapshot(const char *dir, const sbe_config *cfg, const sbe_model *model, const double *rho_band,
                       int32_t append)
{
    if (!dir || !cfg || !model || !rho_band) return fe_fail(SBE_ERR_CONFIG, "null argument");
    sbe::Grid g;
    std::string err;
    int rc = g.build(*cfg, *model, -1, -1, err);
    if (rc) return fe_fail(rc, err);
    const int N = model->nband, nv = model->nval;
    if (nv < 1 || nv >= N) return fe_fail(SBE_ERR_MODEL, "snapshots need 1 <= Nval < N (conduction and coherence elements)");
    const std::string d(dir);
    const char *mode = append ? "a" : "w";
    const char *bmode = append ? "ab" : "wb";
    FILE *f1 = std::fopen((d + "/occupation_1d_integration.dat").c_str(), mode);
    FILE *fn = std::fopen((d + "/fisrt_conduction.dat").c_str(), bmode);
    FILE *fp = std::fopen((d + "/fisrt_coherence.dat").c_str(), bmode);
    FILE *fl = std::fopen((d + "/snapshot_log.dat").c_str(), mode);  // opened, never written (antelope.cpp:291)
    const bool ok = f1 && fn && fp && fl;
    if (ok) {
        const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
        const size_t nn = static_cast<size_t>(N) * N;
        std::vector<double> occ1d(2 * static_cast<size_t>(n0), 0.0), nc(g.ntotal);
        std::vector<double> pi(2 * static_cast<size_t>(g.ntotal));
        for (int64_t k = 0; k < g.ntotal; ++k) {
            const double *r = rho_band + 2 * nn * k;
            const double *cc = r + 2 * (static_cast<size_t>(nv) * N + nv);      // rho_H[Nval][Nval]
            const double *co = r + 2 * (static_cast<size_t>(nv) * N + nv - 1);  // rho_H[Nval][Nval-1]
            const int64_t i0 = k / (static_cast<int64_t>(n1) * n2);             // momaxis.h:87-91
            occ1d[2 * i0] += cc[0] * g.w[k];
            occ1d[2 * i0 + 1] += cc[1] * g.w[k];
            nc[k] = cc[0];
            pi[2 * k] = co[0];
            pi[2 * k + 1] = co[1];
        }
        for (int m = 0; m < n0; ++m) std::fprintf(f1, "%16e    ", occ1d[2 * m] / static_cast<double>(n1));
        std::fprintf(f1, "\n");
        std::fwrite(nc.data(), sizeof(double), nc.size(), fn);
        std::fwrite(pi.data(), sizeof(double), pi.size(), fp);
    }
    for (FILE *f : {f1, fn, fp, fl})
        if (f) std::fclose(f);
    return ok ? SBE_OK : fe_fail(SBE_ERR_CONFIG, "cannot write snapshot files in " + d);
}

int sbe_tmdc_params(const char *material, int32_t soc, sbe_model *m)
{
    if (!material || !m) return SBE_ERR_CONFIG;
    for (const TmdcRow &r : kTmdc) {
        if (std::string(material) != r.name) continue;
        m->kind = SBE_MODEL_TMDC;
        m->nband = soc ? 6 : 3;
        m->nval = soc ? 2 : 1;
        m->is_dipole_zero = 1;
        m->params[0] = r.a_ang / kAuAngstrom;
        m->params[1] = r.soc_ev / kAuEV;
        for (int i = 0; i < 19; ++i) m->params[2 + i] = r.v[i] / kAuEV;
        return SBE_OK;
    }
    return SBE_ERR_MODEL;
}

int sbe_spec_parse_string(const char *text, sbe_spec **out)
{
    if (!text || !out) return fe_fail(SBE_ERR_CONFIG, "null argument");
    *out = nullptr;
    const std::string src(text);
    Setting root;
    std::string err;
    if (!Parser(src).parse(root, err)) return fe_fail(SBE_ERR_CONFIG, err);
    auto spec = std::make_unique<sbe_spec>();
    sbe_config &c = spec->cfg;
    c.abi_version = SBE_ABI_VERSION;

    // gauge (antelope.cpp:111-129); default LengthWannier with a warning
    std::string gauge;
    get_string(&root, "gauge", gauge);
    if (gauge == "Wannier" || gauge == "LengthWannier") c.gauge = SBE_GAUGE_LENGTH_WANNIER;
    else if (gauge == "Hamiltonian" || gauge == "LengthHamiltonian") c.gauge = SBE_GAUGE_LENGTH_HAMILTONIAN;
    else if (gauge == "Velocity" || gauge == "VelocityHamiltonian") c.gauge = SBE_GAUGE_VELOCITY_HAMILTONIAN;
    else {
        c.gauge = SBE_GAUGE_LENGTH_WANNIER;
        spec->warnings += "No gauge type is specified. Choose LengthWannier for default gauge.\n";
    }

    // calc group (SBEs.h:432-509)
    const Setting *calc = root.child("calc");
    if (!calc || calc->kind != Setting::Group) return fe_fail(SBE_ERR_CONFIG, "Setting not found while search calc group");
    const Setting *np = calc->child("Npoints");
    if (!np || (np->kind != Setting::Array && np->kind != Setting::List) || np->items.empty() || np->items.size() > 3)
        return fe_fail(SBE_ERR_CONFIG, "calc.Npoints must be an array of 1 to 3 integers");
    for (int i = 0; i < 3; ++i) c.npoints[i] = 1;
    for (size_t i = 0; i < np->items.size(); ++i) {
        if (np->items[i].kind != Setting::Int) return fe_fail(SBE_ERR_CONFIG, "Type exception in calc group (Npoints)");
        c.npoints[i] = static_cast<int32_t>(np->items[i].i);
    }
    c.dt = 0.1;
    get_double(calc, "dt", c.dt);
    c.T1 = -1.0;
    c.T2 = -1.0;
    get_double(calc, "T1", c.T1);
    get_double(calc, "T2", c.T2);
    long long rk = 5;  // SBEs.h:438
    get_int(calc, "RKorder", rk);
    c.rk_order = static_cast<int32_t>(rk);
    std::string method;
    if (get_string(calc, "method", method) && (method == "cash_karp" || method == "CashKarp")) c.rk_order = SBE_RK_CASH_KARP;
    bool ii = false;  // SBEs.h:442
    get_bool(calc, "InterIntra", ii);
    c.inter_intra = ii;
    for (int i = 0; i < 3; ++i) c.ksfactor[i] = 1.0;
    if (const Setting *ks = calc->child("ksfactor")) {
        for (size_t i = 0; i < ks->items.size() && i < 3; ++i)
            if (!ks->items[i].number(c.ksfactor[i])) return fe_fail(SBE_ERR_CONFIG, "Type exception in calc group (ksfactor)");
    } else {
        spec->warnings += "calc.ksfactor missing (the reference requires it): using [1, 1, 1]\n";
    }
    if (const Setting *ax = calc->child("BZaxes")) {
        c.has_bz_axes = 1;
        for (size_t i = 0; i < ax->items.size() && i < 9; ++i) ax->items[i].number(c.bz_axes[i]);
    }
    if (const Setting *og = calc->child("BZorigin")) {
        c.has_bz_origin = 1;
        for (size_t i = 0; i < og->items.size() && i < 3; ++i) og->items[i].number(c.bz_origin[i]);
    }
    std::string s;
    if (get_string(calc, "grid", s)) c.grid = (s == "mp" || s == "MP") ? SBE_GRID_MP : SBE_GRID_REFERENCE;
    if (get_string(calc, "precision", s))
        c.precision = s == "fp32" ? SBE_PREC_FP32 : s == "split16" ? SBE_PREC_SPLIT16 : SBE_PREC_FP64;
    c.rtol = 1e-8;
    c.atol = 1e-12;
    {
        long long shots = 0;  // SBEs.h:441,461
        if (get_int(calc, "shotNumber", shots)) spec->shot_number = static_cast<int32_t>(shots);
    }
    get_double(calc, "rtol", c.rtol);
    get_double(calc, "atol", c.atol);

    // laser group (SBEs.h:511-561)
    const Setting *laser = root.child("laser");
    const Setting *pulses = laser ? laser->child("pulses") : nullptr;
    if (!pulses || pulses->items.empty()) return fe_fail(SBE_ERR_CONFIG, "No laser, check config file");
    for (const Setting &p : pulses->items) {
        sbe_pulse q{};
        std::string env;
        double cep, phix;
        if (!(get_double(&p, "E0", q.E0) && get_double(&p, "ellip", q.ellip) && get_double(&p, "w0", q.w0) &&
              get_double(&p, "ncycles", q.ncycles) && get_double(&p, "cep", cep) && get_double(&p, "t0", q.t0) &&
              get_double(&p, "phix", phix) && get_string(&p, "env_name", env)))
            return fe_fail(SBE_ERR_CONFIG, "Error while initilizng laser pulses. Check config file");
        double thetaz = 0.0, phiz = 0.0, ncc = 0.0;
        get_double(&p, "thetaz", thetaz);
        get_double(&p, "phiz", phiz);
        get_double(&p, "ncycles_clipped", ncc);
        const double d = kPi / 180.0;
        q.cep = cep * d;
        q.phix = phix * d;
        q.thetaz = thetaz * d;
        q.phiz = phiz * d;
        q.ncycles_clipped = ncc;
        if (env == "cos4") q.envelope = SBE_ENV_COS4;
        else if (env == "gauss") q.envelope = SBE_ENV_GAUSS;
        else if (env == "gauss_clipped") q.envelope = SBE_ENV_GAUSS_CLIPPED;
        else if (env == "cos4_clipped") q.envelope = SBE_ENV_COS4_CLIPPED;
        else if (env == "sin2") q.envelope = SBE_ENV_SIN2;
        else return fe_fail(SBE_ERR_CONFIG, "invalid envelope type " + env);
        spec->pulses.push_back(q);
    }
    if (const Setting *off = laser->child("offset")) {
        if (off->items.size() >= 2) {
            off->items[0].number(c.offset_before);
            off->items[1].number(c.offset_after);
        }
    }
    if (get_string(laser, "efield", s)) c.efield_mode = s == "consistent" ? SBE_EFIELD_CONSISTENT : SBE_EFIELD_REFERENCE;

    // target and model (SBEs.h:130-215, Haldane.h:46-84)
    if (!get_string(&root, "target", spec->target)) return fe_fail(SBE_ERR_CONFIG, "No 'target' in input file");
    const Setting *mg = root.child(spec->target);
    bool custom = false;
    if (mg) get_bool(mg, "UsingCustomInitial", custom);
    if (custom) return fe_fail(SBE_ERR_UNSUPPORTED, "UsingCustomInitial is not supported");
    if (mg && get_double(mg, "Ef", c.fermi_energy)) c.init_type = SBE_INIT_FERMI_DIRAC;
    sbe_model &m = spec->model;
    if (spec->target == "Haldane") {
        double a0;
        if (!(get_double(mg, "t1", m.t1) && get_double(mg, "t2", m.t2) && get_double(mg, "phi0", m.phi0) &&
              get_double(mg, "a0", a0)))
            return fe_fail(SBE_ERR_MODEL, "Some Haldane paramters are missing");
        m.a0 = a0 / kAuAngstrom;
        if (!get_double(mg, "M0", m.M0)) {
            double mt2;
            if (!get_double(mg, "Mt2", mt2)) return fe_fail(SBE_ERR_MODEL, "M0 param missing");
            m.M0 = mt2 * m.t2;
        }
        m.kind = SBE_MODEL_HALDANE;
        m.nband = 2;
        m.nval = 1;
        m.is_dipole_zero = 1;
    } else if (spec->target == "WilsonMass" || spec->target == "SquareSBand" || spec->target == "KaneMele" ||
               spec->target == "XYHlattice" || spec->target == "TMDCs" || spec->target == "Haldane2L" ||
               spec->target == "Bi2Se3surf" || spec->target == "MultiWSM" || spec->target == "WDSM_TaAs") {
        int rc = analytic_target(spec->target, mg, m);
        if (rc) return rc;
    } else if (mg && mg->child("dataName")) {  // Wannier90.h:107-125
        std::string path;
        if (!get_string(mg, "dataName", path)) return fe_fail(SBE_ERR_CONFIG, "dataName must be a string");
        if (!path.empty() && path[0] != '/' && !t_base_dir.empty()) path = t_base_dir + "/" + path;
        long long nval = 0;
        double fermi_ev;
        if (get_double(mg, "FermiE", fermi_ev)) {  // if FermiE exists, the valence band number is ignored
            c.init_type = SBE_INIT_FERMI_DIRAC;
            c.fermi_energy = fermi_ev / kAuEV;
        } else if (!get_int(mg, "Nval", nval)) {
            return fe_fail(SBE_ERR_MODEL, "Wannier90 target needs Nval or FermiE");
        }
        int rc = sbe_wannier_read_tb(path.c_str(), &spec->tb);
        if (rc) return fe_fail(rc, sbe_wannier_last_error());
        rc = sbe_wannier_tb_model(spec->tb, static_cast<int32_t>(nval), &m);
        if (rc) return fe_fail(rc, sbe_wannier_last_error());
    } else {
        return fe_fail(SBE_ERR_UNSUPPORTED, "target '" + spec->target +
                                                "' is not a model class (Haldane, Haldane2L, KaneMele, WilsonMass, "
                                                "Bi2Se3surf, TMDCs, XYHlattice, SquareSBand, WDSM_TaAs, MultiWSM, "
                                                "or a Wannier90 group with dataName)");
    }
    c.npulses = static_cast<int32_t>(spec->pulses.size());
    c.pulses = spec->pulses.data();
    *out = spec.release();
    return SBE_OK;
}

int sbe_spec_parse_file(const char *path, sbe_spec **out)
{
    if (!path) return fe_fail(SBE_ERR_CONFIG, "null path");
    std::ifstream f(path);
    if (!f) return fe_fail(SBE_ERR_CONFIG, std::string("I/O error while reading file ") + path);
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string p(path);
    const size_t slash = p.find_last_of('/');
    t_base_dir = slash == std::string::npos ? std::string() : p.substr(0, slash);
    const int rc = sbe_spec_parse_string(ss.str().c_str(), out);
    t_base_dir.clear();
    return rc;
}

int sbe_spec_get(const sbe_spec *spec, const sbe_config **cfg, const sbe_model **model)
{
    if (!spec) return fe_fail(SBE_ERR_CONFIG, "null spec");
    if (cfg) *cfg = &spec->cfg;
    if (model) *model = &spec->model;
    return SBE_OK;
}

const char *sbe_spec_warnings(const sbe_spec *spec) { return spec ? spec->warnings.c_str() : ""; }

void sbe_spec_free(sbe_spec *spec) { delete spec; }

int sbe_write_outputs(const char *dir, const sbe_config *cfg, const sbe_model *model, const sbe_info *info,
                      const double *j_inter, const double *j_intra, const double *rho_bar, const double *band_pop)
{
    if (!dir || !cfg || !model || !info || !j_inter || !j_intra || !rho_bar)
        return fe_fail(SBE_ERR_CONFIG, "null argument");
    const std::string d(dir);
    auto open = [&](const char *name) { return std::fopen((d + "/" + name).c_str(), "w"); };
    sbe::Laser L;
    std::string err;
    int rc = L.init(*cfg, err);
    if (rc) return fe_fail(rc, err);
    const int64_t nt = info->nt;
    const int N = info->nband;
    const double nk = static_cast<double>(info->nk_total);
    const double dV = info->dV;
    // outlaserdata.dat, antelope.cpp:208-235
    FILE *f = open("outlaserdata.dat");
    if (!f) return fe_fail(SBE_ERR_CONFIG, "cannot open outlaserdata.dat in " + d);
    for (int64_t n = 0; n < nt; ++n) {
        const double at = L.atmin + L.dt * static_cast<double>(n);
        double A[3], E[3];
        L.eval(at, A, E);
        std::fprintf(f, "%e    ", at);
        for (int i = 0; i < 3; ++i) std::fprintf(f, "%e      ", E[i]);
        for (int i = 0; i < 3; ++i) std::fprintf(f, "%e      ", A[i]);
        std::fprintf(f, "\n");
    }
    std::fclose(f);
    // laserParameters.dat is opened and never written by the reference (antelope.cpp:179)
    f = open("laserParameters.dat");
    if (f) std::fclose(f);
    // currents x dV, occupations / Ntotal, coherence / Ntotal (antelope.cpp:514-548)
    FILE *fi = open("interband_dipole_full_evol.dat");
    FILE *fa = open("intraband_current_full_evol.dat");
    FILE *fo = open("occupation__full__evol.dat");
    FILE *fc = open("coherence__full__evol.dat");
    FILE *fb = band_pop ? open("band_population__full__evol.dat") : nullptr;
    if (!fi || !fa || !fo || !fc || (band_pop && !fb)) {
        for (FILE *x : {fi, fa, fo, fc, fb})
            if (x) std::fclose(x);
        return fe_fail(SBE_ERR_CONFIG, "cannot open output files in " + d);
    }
    const int v = model->nval > 0 ? model->nval - 1 : 0;
    const int cnd = model->nval < N ? model->nval : N - 1;
    for (int64_t n = 0; n < nt; ++n) {
        for (int a = 0; a < 3; ++a) {
            std::fprintf(fi, "%.16e    ", j_inter[3 * n + a] * dV);
            std::fprintf(fa, "%.16e    ", j_intra[3 * n + a] * dV);
        }
        std::fprintf(fi, "\n");
        std::fprintf(fa, "\n");
        const double *r = rho_bar + n * 2 * N * N;
        double tot = 0.0;
        for (int m = 0; m < N; ++m) {
            std::fprintf(fo, "%.16e    ", r[2 * (m * N + m)] / nk);
            tot += r[2 * (m * N + m)];
        }
        std::fprintf(fo, "%.16e\n", tot / nk);
        // the reference indexes Nval*N + Nval + 1, outside the N x N block
        // (reads the next row); the intended valence-conduction element is written
        const int e = v * N + cnd;
        std::fprintf(fc, "%.16e    ", r[2 * e] / nk);
        std::fprintf(fc, "%.16e\n", r[2 * e + 1] / nk);
        if (fb) {  // band populations normalised by the summed k weights
            double s = 0.0;
            for (int m = 0; m < N; ++m) {
                std::fprintf(fb, "%.16e    ", band_pop[n * N + m] / info->total_weight);
                s += band_pop[n * N + m];
            }
            std::fprintf(fb, "%.16e\n", s / info->total_weight);
        }
    }
    for (FILE *x : {fi, fa, fo, fc, fb})
        if (x) std::fclose(x);
    return SBE_OK;
}

}  // extern "C"
