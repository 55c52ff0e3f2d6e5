// atl.cpp - reader of the .atl input format (src/inputFile.atl), mapped onto
// the same sbe_config / sbe_model as the inputParam.cfg front end.
//
// The reference ships src/inputFile.atl but nothing parses it. Its layout:
// BEGIN <SECTION> ... END <SECTION> blocks of `name = value` lines with `#`
// comments. GLOBAL_VARIABLES values are arithmetic expressions over earlier
// variables (+ - * / ^, parentheses, acos/asin/atan/sin/cos/tan/exp/log/sqrt/
// abs) with an optional unit word after the expression (au, rad, ang, bohr,
// nm, microm/um, fs, as, eV, deg). What maps onto the solver:
//   SYSTEM   E(k).model = MODEL_HALDANE_BANDS(la0, t1, t2, phi0, M0)
//            U_f(t).model = LASER_SIN2 | LASER_COS4 | LASER_GAUSS
//                           (E0, omega, ncycles [, ellip [, cep [, t0]]])
//   SETUP    tgrid.step (dt); kgrid.size or kgrid.n1/n2/n3 (Npoints);
//            T1, T2, InterIntra; offset.before/after
//   SOLVER   tsolver = RUNGE_KUTTA_4() | RUNGE_KUTTA_5() | CASH_KARP(rtol, atol)
// The file's other directives belong to solvers this path does not have and
// are errors that name them: LASER_TOPHAT_SIN (a flat-top pulse), a k-space
// solver (ksolver = CRANK_NICOLSON()), a k box (kbox.*); tgrid.min/max are
// set by the pulse envelope here and are checked only for being numbers.
// PROPERTIES switches are accepted with a warning (the outputs are fixed).
#include <cctype>
#include <cmath>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "antelope/sbe_frontend.h"
#include "frontend_internal.hpp"

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kAuAngstrom = 0.529177210903;  // constant.h:34
constexpr double kAuEV = 27.211386245988;       // constant.h:35
constexpr double kAuFs = 0.02418884326585747;   // atomic unit of time in fs

thread_local std::string g_atl_error;

struct AtlError {
    int code;
    std::string msg;
};

[[noreturn]] void fail(int code, const std::string &m) { throw AtlError{code, m}; }

bool unit_factor(const std::string &u, double &f)
{
    static const std::map<std::string, double> units = {
        {"au", 1.0}, {"rad", 1.0}, {"bohr", 1.0}, {"ang", 1.0 / kAuAngstrom}, {"nm", 10.0 / kAuAngstrom},
        {"microm", 1e4 / kAuAngstrom}, {"um", 1e4 / kAuAngstrom}, {"fs", 1.0 / kAuFs}, {"as", 1e-3 / kAuFs},
        {"eV", 1.0 / kAuEV}, {"deg", kPi / 180.0}};
    auto it = units.find(u);
    if (it == units.end()) return false;
    f = it->second;
    return true;
}

// Recursive-descent evaluator for one expression.
class Expr {
public:
    Expr(const std::string &s, const std::map<std::string, double> &vars, int line)
        : s_(s), vars_(vars), line_(line) {}

    double value()
    {
        double v = sum();
        skip();
        if (pos_ < s_.size()) {  // a trailing unit word
            size_t b = pos_;
            while (pos_ < s_.size() && (std::isalnum(static_cast<unsigned char>(s_[pos_])) || s_[pos_] == '_')) ++pos_;
            const std::string u = s_.substr(b, pos_ - b);
            double f;
            if (u.empty() || !unit_factor(u, f)) err("unknown unit or trailing text '" + s_.substr(b) + "'");
            v *= f;
            skip();
            if (pos_ < s_.size()) err("trailing text '" + s_.substr(pos_) + "'");
        }
        return v;
    }

private:
    const std::string &s_;
    const std::map<std::string, double> &vars_;
    int line_;
    size_t pos_ = 0;

    [[noreturn]] void err(const std::string &m) const
    {
        fail(SBE_ERR_CONFIG, "line " + std::to_string(line_) + ": " + m);
    }
    void skip()
    {
        while (pos_ < s_.size() && std::isspace(static_cast<unsigned char>(s_[pos_]))) ++pos_;
    }
    bool eat(char c)
    {
        skip();
        if (pos_ < s_.size() && s_[pos_] == c) {
            ++pos_;
            return true;
        }
        return false;
    }
    double sum()
    {
        double v = product();
        for (;;) {
            if (eat('+')) v += product();
            else if (eat('-')) v -= product();
            else return v;
        }
    }
    double product()
    {
        double v = unary();
        for (;;) {
            skip();
            if (pos_ + 1 < s_.size() && s_[pos_] == '*' && s_[pos_ + 1] == '*') return v;  // handled in power
            if (eat('*')) v *= unary();
            else if (eat('/')) v /= unary();
            else return v;
        }
    }
    double unary()
    {
        if (eat('-')) return -unary();
        if (eat('+')) return unary();
        return power();
    }
    double power()
    {
        double b = primary();
        skip();
        if (eat('^')) return std::pow(b, unary());
        if (pos_ + 1 < s_.size() && s_[pos_] == '*' && s_[pos_ + 1] == '*') {
            pos_ += 2;
            return std::pow(b, unary());
        }
        return b;
    }
    double primary()
    {
        skip();
        if (eat('(')) {
            double v = sum();
            if (!eat(')')) err("missing ')'");
            return v;
        }
        if (pos_ < s_.size() && (std::isdigit(static_cast<unsigned char>(s_[pos_])) || s_[pos_] == '.')) {
            const char *b = s_.c_str() + pos_;
            char *e = nullptr;
            const double v = std::strtod(b, &e);
            if (e == b) err("bad number");
            pos_ += static_cast<size_t>(e - b);
            return v;
        }
        size_t b = pos_;
        while (pos_ < s_.size() && (std::isalnum(static_cast<unsigned char>(s_[pos_])) || s_[pos_] == '_')) ++pos_;
        const std::string id = s_.substr(b, pos_ - b);
        if (id.empty()) err("expected a value at '" + s_.substr(b) + "'");
        if (eat('(')) {
            double a = sum();
            if (!eat(')')) err("missing ')' after " + id);
            static const std::map<std::string, double (*)(double)> fn = {
                {"acos", [](double x) { return std::acos(x); }}, {"asin", [](double x) { return std::asin(x); }},
                {"atan", [](double x) { return std::atan(x); }}, {"sin", [](double x) { return std::sin(x); }},
                {"cos", [](double x) { return std::cos(x); }},   {"tan", [](double x) { return std::tan(x); }},
                {"exp", [](double x) { return std::exp(x); }},   {"log", [](double x) { return std::log(x); }},
                {"sqrt", [](double x) { return std::sqrt(x); }}, {"abs", [](double x) { return std::fabs(x); }}};
            auto it = fn.find(id);
            if (it == fn.end()) err("unknown function " + id);
            return it->second(a);
        }
        auto it = vars_.find(id);
        if (it == vars_.end()) {
            double f;
            if (unit_factor(id, f)) err("unit '" + id + "' without a value");
            err("undefined variable " + id);
        }
        return it->second;
    }
};

// NAME( a, b, ... ) -> NAME and the argument texts
bool directive(const std::string &v, std::string &name, std::vector<std::string> &args)
{
    const size_t p = v.find('(');
    if (p == std::string::npos || v.back() != ')') return false;
    name = v.substr(0, p);
    while (!name.empty() && std::isspace(static_cast<unsigned char>(name.back()))) name.pop_back();
    for (char c : name)
        if (!(std::isupper(static_cast<unsigned char>(c)) || std::isdigit(static_cast<unsigned char>(c)) || c == '_'))
            return false;
    args.clear();
    std::string cur;
    int depth = 0;
    for (size_t i = p + 1; i + 1 < v.size(); ++i) {
        const char c = v[i];
        if (c == '(') ++depth;
        if (c == ')') --depth;
        if (c == ',' && depth == 0) {
            args.push_back(cur);
            cur.clear();
        } else {
            cur += c;
        }
    }
    if (cur.find_first_not_of(" \t") != std::string::npos) args.push_back(cur);
    return true;
}

std::string trim(const std::string &s)
{
    const size_t b = s.find_first_not_of(" \t\r");
    if (b == std::string::npos) return "";
    const size_t e = s.find_last_not_of(" \t\r");
    return s.substr(b, e - b + 1);
}

}  // namespace

extern "C" {

const char *sbe_atl_last_error(void) { return g_atl_error.c_str(); }

int sbe_spec_parse_atl_string(const char *text, sbe_spec **out)
{
    if (!text || !out) {
        g_atl_error = "null argument";
        return SBE_ERR_CONFIG;
    }
    *out = nullptr;
    try {
        std::map<std::string, double> vars;
        std::istringstream in(text);
        std::string raw, section;
        int line = 0;
        sbe_config c{};
        c.abi_version = SBE_ABI_VERSION;
        c.gauge = SBE_GAUGE_LENGTH_WANNIER;
        c.dt = 0.1;
        c.T1 = c.T2 = -1.0;
        c.rk_order = SBE_RK4;
        c.rtol = 1e-8;
        c.atol = 1e-12;
        for (double &k : c.ksfactor) k = 1.0;
        int npoints[3] = {0, 1, 1};
        sbe_model m{};
        bool have_model = false;
        std::vector<sbe_pulse> pulses;
        std::string warnings;
        auto num = [&](const std::string &e) { return Expr(e, vars, line).value(); };
        while (std::getline(in, raw)) {
            ++line;
            const size_t hash = raw.find('#');
            const std::string l = trim(hash == std::string::npos ? raw : raw.substr(0, hash));
            if (l.empty()) continue;
            if (l.rfind("BEGIN ", 0) == 0) {
                if (!section.empty()) fail(SBE_ERR_CONFIG, "line " + std::to_string(line) + ": BEGIN inside " + section);
                section = trim(l.substr(6));
                continue;
            }
            if (l.rfind("END", 0) == 0) {
                if (trim(l.substr(3)) != section) fail(SBE_ERR_CONFIG, "line " + std::to_string(line) + ": END does not close " + section);
                section.clear();
                continue;
            }
            const size_t eq = l.find('=');
            if (eq == std::string::npos) fail(SBE_ERR_CONFIG, "line " + std::to_string(line) + ": expected name = value");
            const std::string name = trim(l.substr(0, eq)), val = trim(l.substr(eq + 1));
            std::string dname;
            std::vector<std::string> args;
            if (section == "GLOBAL_VARIABLES") {
                vars[name] = num(val);
            } else if (section == "SIMULATION_SETUP") {
                if (name.rfind("kbox.", 0) == 0)
                    fail(SBE_ERR_UNSUPPORTED, "line " + std::to_string(line) + ": kbox is not supported: the k grid spans the "
                                              "model's Brillouin zone; give kgrid.size (or kgrid.n1/n2/n3)");
                if (name == "tgrid.step") c.dt = num(val);
                else if (name == "tgrid.min" || name == "tgrid.max") (void)num(val);
                else if (name == "kgrid.size") npoints[0] = npoints[1] = static_cast<int>(std::lround(num(val)));
                else if (name == "kgrid.n1") npoints[0] = static_cast<int>(std::lround(num(val)));
                else if (name == "kgrid.n2") npoints[1] = static_cast<int>(std::lround(num(val)));
                else if (name == "kgrid.n3") npoints[2] = static_cast<int>(std::lround(num(val)));
                else if (name == "T1") c.T1 = num(val);
                else if (name == "T2") c.T2 = num(val);
                else if (name == "InterIntra") c.inter_intra = (val == "TRUE" || val == "true") ? 1 : 0;
                else if (name == "offset.before") c.offset_before = num(val);
                else if (name == "offset.after") c.offset_after = num(val);
                else fail(SBE_ERR_UNSUPPORTED, "line " + std::to_string(line) + ": unknown setup key " + name);
            } else if (section == "PROPERTIES") {
                warnings += "PROPERTIES." + name + " ignored: the output files are fixed\n";
            } else if (section == "SYSTEM") {
                if (!directive(val, dname, args))
                    fail(SBE_ERR_CONFIG, "line " + std::to_string(line) + ": expected a directive NAME(...) for " + name);
                if (name == "E(k).model") {
                    if (dname != "MODEL_HALDANE_BANDS")
                        fail(SBE_ERR_UNSUPPORTED, "line " + std::to_string(line) + ": band model " + dname + " is not supported");
                    if (args.size() != 5)
                        fail(SBE_ERR_CONFIG, "line " + std::to_string(line) + ": MODEL_HALDANE_BANDS(la0, t1, t2, phi0, M0)");
                    m.kind = SBE_MODEL_HALDANE;
                    m.nband = 2;
                    m.nval = 1;
                    m.is_dipole_zero = 1;
                    m.a0 = num(args[0]);  // already in bohr when written with a unit (la0 = 1.0 ang)
                    m.t1 = num(args[1]);
                    m.t2 = num(args[2]);
                    m.phi0 = num(args[3]);
                    m.M0 = num(args[4]);
                    have_model = true;
                } else if (name.rfind("U_f(t).model", 0) == 0) {
                    sbe_pulse p{};
                    if (dname == "LASER_SIN2") p.envelope = SBE_ENV_SIN2;
                    else if (dname == "LASER_COS4") p.envelope = SBE_ENV_COS4;
                    else if (dname == "LASER_GAUSS") p.envelope = SBE_ENV_GAUSS;
                    else
                        fail(SBE_ERR_UNSUPPORTED, "line " + std::to_string(line) + ": laser " + dname +
                                                      " is not supported (LASER_SIN2, LASER_COS4, LASER_GAUSS)");
                    if (args.size() < 3 || args.size() > 6)
                        fail(SBE_ERR_CONFIG, "line " + std::to_string(line) + ": " + dname +
                                                 "(E0, omega, ncycles [, ellip [, cep [, t0]]])");
                    p.E0 = num(args[0]);
                    p.w0 = num(args[1]);
                    p.ncycles = num(args[2]);
                    if (args.size() > 3) p.ellip = num(args[3]);
                    if (args.size() > 4) p.cep = num(args[4]);
                    if (args.size() > 5) p.t0 = num(args[5]);
                    pulses.push_back(p);
                } else {
                    fail(SBE_ERR_UNSUPPORTED, "line " + std::to_string(line) + ": unknown system entry " + name);
                }
            } else if (section == "SOLVER") {
                if (!directive(val, dname, args))
                    fail(SBE_ERR_CONFIG, "line " + std::to_string(line) + ": expected a directive NAME(...) for " + name);
                if (name == "tsolver") {
                    if (dname == "RUNGE_KUTTA_4") c.rk_order = SBE_RK4;
                    else if (dname == "RUNGE_KUTTA_5") c.rk_order = SBE_RK5_FEHLBERG;
                    else if (dname == "CASH_KARP") {
                        c.rk_order = SBE_RK_CASH_KARP;
                        if (args.size() > 0) c.rtol = num(args[0]);
                        if (args.size() > 1) c.atol = num(args[1]);
                    } else {
                        fail(SBE_ERR_UNSUPPORTED, "line " + std::to_string(line) + ": time solver " + dname + " is not supported");
                    }
                } else if (name == "ksolver") {
                    fail(SBE_ERR_UNSUPPORTED, "line " + std::to_string(line) + ": ksolver = " + dname +
                                                  " is not supported: k points are independent here (no k-space solver)");
                } else {
                    fail(SBE_ERR_UNSUPPORTED, "line " + std::to_string(line) + ": unknown solver entry " + name);
                }
            } else {
                fail(SBE_ERR_CONFIG, "line " + std::to_string(line) + ": entry outside a known BEGIN/END section");
            }
        }
        if (!section.empty()) fail(SBE_ERR_CONFIG, "missing END " + section);
        if (!have_model) fail(SBE_ERR_MODEL, "no E(k).model in SYSTEM");
        if (pulses.empty()) fail(SBE_ERR_CONFIG, "no U_f(t).model laser in SYSTEM");
        if (npoints[0] < 1) fail(SBE_ERR_CONFIG, "no kgrid.size (or kgrid.n1/n2/n3) in SIMULATION_SETUP");
        for (int i = 0; i < 3; ++i) c.npoints[i] = npoints[i];
        auto spec = std::make_unique<sbe_spec>();
        spec->cfg = c;
        spec->model = m;
        spec->pulses = pulses;
        spec->cfg.npulses = static_cast<int32_t>(spec->pulses.size());
        spec->cfg.pulses = spec->pulses.data();
        spec->target = "Haldane";
        spec->warnings = warnings;
        *out = spec.release();
        return SBE_OK;
    } catch (const AtlError &e) {
        g_atl_error = e.msg;
        return e.code;
    }
}

int sbe_spec_parse_atl_file(const char *path, sbe_spec **out)
{
    if (!path) {
        g_atl_error = "null path";
        return SBE_ERR_CONFIG;
    }
    std::ifstream f(path);
    if (!f) {
        g_atl_error = std::string("I/O error while reading file ") + path;
        return SBE_ERR_CONFIG;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    return sbe_spec_parse_atl_string(ss.str().c_str(), out);
}

}  // extern "C"
