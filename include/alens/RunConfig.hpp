// RunConfig.hpp -- the reference's run configuration file (RunConfig.yaml) for the rod engine.
//
// The reference reads RunConfig.yaml into SimToolbox's SylinderConfig (`rodSystem.runConfig`) when
// TubuleSystem starts (SRC/main.cpp:26, SRC/TubuleSystem.cpp:79 and :38); the keys and their meaning
// are listed in docs/source/quickstart.md:316-358 and used by the Examples/*/RunConfig.yaml files.
// readRunConfig() parses the YAML subset those files use: one `key: value` per line at column 0,
// `#` comments, scalars (numbers, true/false, quoted or bare strings) and flow sequences `[a, b, c]`.
// A key with an empty value opens a block (the `boundaries:` list); its indented lines are kept as
// text and not interpreted (wall boundaries are outside this engine).  Errors throw alens::Error with
// ALENS_E_ARG naming the file, line and key.
#pragma once
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../alens_capi.h"

namespace alens {

#ifndef ALENS_ERROR_DEFINED
#define ALENS_ERROR_DEFINED
struct Error : std::runtime_error {
    alens_status status;
    Error(alens_status s, const std::string &m) : std::runtime_error(m), status(s) {}
};
#endif

// SylinderConfig fields (names as in the YAML file).  Defaults are those the reference documents
// for optional keys (docs/source/quickstart.md:316-358): monolayer, initCircularX, sylinderFixed
// false; initPreSteps 100; timerLevel 0.  Keys without a default are required when the engine needs
// them (see toEngineConfig).
struct RunConfig {
    unsigned rngSeed = 0;
    int logLevel = 2, timerLevel = 0;
    double simBoxLow[3] = {0, 0, 0}, simBoxHigh[3] = {0, 0, 0};
    bool simBoxPBC[3] = {false, false, false};
    bool monolayer = false;
    double initBoxLow[3] = {0, 0, 0}, initBoxHigh[3] = {0, 0, 0};
    double initOrient[3] = {2, 2, 2};  // components < -1 or > 1 are drawn at random
    bool initCircularX = false;
    int initPreSteps = 100;
    double viscosity = 0, KBT = 0, linkKappa = 0, linkGap = 0;
    bool sylinderFixed = false;
    int sylinderNumber = 0;
    double sylinderLength = 0, sylinderLengthSigma = 0, sylinderDiameter = 0;
    double sylinderDiameterColRatio = 1.0, sylinderLengthColRatio = 1.0, sylinderColBuf = 0;
    double dt = 0, timeTotal = 0, timeSnap = 0;
    double conResTol = 0;
    int conMaxIte = 0, conSolverChoice = 0;
    // every key read from the file (scalars and sequences as written), and the raw block of any key
    // whose value is an indented block (e.g. "boundaries")
    std::map<std::string, std::string> values;
    std::map<std::string, std::string> blocks;
    bool has(const std::string &key) const { return values.count(key) != 0; }

    // The engine configuration of these keys: simBox*, monolayer, sylinderColBuf (absolute, um),
    // viscosity, KBT (<= 0: no Brownian motion), dt, conResTol, conMaxIte, conSolverChoice (0 =
    // BBPGD, the only solver here), the collision ratios and rngSeed.  The engine-only settings
    // (precision form, warm start, device, cell edge) are not reference keys; they come from `base`.
    alens_config toEngineConfig(const alens_config &base) const {
        for (const char *k : {"simBoxLow", "simBoxHigh", "simBoxPBC", "viscosity", "KBT", "dt", "sylinderColBuf",
                              "conResTol", "conMaxIte"})
            if (!has(k)) throw Error(ALENS_E_ARG, std::string("RunConfig: required key '") + k + "' is missing");
        alens_config c = base;
        for (int k = 0; k < 3; ++k) {
            c.box_lo[k] = simBoxLow[k];
            c.box_hi[k] = simBoxHigh[k];
            c.pbc[k] = simBoxPBC[k] ? 1 : 0;
        }
        c.monolayer = monolayer ? 1 : 0;
        c.col_buf = sylinderColBuf;
        c.viscosity = viscosity;
        c.kT = KBT;
        c.dt = dt;
        c.tol = conResTol;
        c.diam_col_ratio = sylinderDiameterColRatio;
        c.len_col_ratio = sylinderLengthColRatio;
        c.max_ite = conMaxIte;
        c.solver = conSolverChoice;
        c.seed = rngSeed;
        return c;
    }
};

namespace yaml_subset {

inline std::string trim(const std::string &s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

// text before a '#' that is not inside quotes
inline std::string strip_comment(const std::string &s) {
    char q = 0;
    for (size_t i = 0; i < s.size(); ++i) {
        const char c = s[i];
        if (q) {
            if (c == q) q = 0;
        } else if (c == '"' || c == '\'') {
            q = c;
        } else if (c == '#' && (i == 0 || std::isspace((unsigned char)s[i - 1]))) {
            return s.substr(0, i);
        }
    }
    return s;
}

inline std::string unquote(const std::string &s) {
    if (s.size() >= 2 && (s.front() == '"' || s.front() == '\'') && s.back() == s.front())
        return s.substr(1, s.size() - 2);
    return s;
}

[[noreturn]] inline void bad(const std::string &where, const std::string &what) {
    throw Error(ALENS_E_ARG, where + ": " + what);
}

inline double to_double(const std::string &v, const std::string &where) {
    const std::string s = unquote(trim(v));
    if (s.empty()) bad(where, "empty number");
    const char *p = s.c_str();
    char *end = nullptr;
    errno = 0;
    const double x = std::strtod(p, &end);
    if (end == p || *end != '\0' || errno == ERANGE) bad(where, "not a number: '" + s + "'");
    return x;
}

inline long long to_int(const std::string &v, const std::string &where) {
    const double x = to_double(v, where);
    if (x != std::floor(x) || std::fabs(x) > 9.0e15) bad(where, "not an integer: '" + trim(v) + "'");
    return (long long)x;
}

// YAML 1.1 booleans as yaml-cpp reads them (true/false, yes/no, on/off, y/n in the usual cases)
inline bool to_bool(const std::string &v, const std::string &where) {
    const std::string s = unquote(trim(v));
    static const char *yes[] = {"true", "True", "TRUE", "yes", "Yes", "YES", "on", "On", "ON", "y", "Y"};
    static const char *no[] = {"false", "False", "FALSE", "no", "No", "NO", "off", "Off", "OFF", "n", "N"};
    for (const char *t : yes)
        if (s == t) return true;
    for (const char *t : no)
        if (s == t) return false;
    bad(where, "not a boolean: '" + s + "'");
}

// items of a flow sequence "[a, b, c]"
inline std::vector<std::string> seq(const std::string &v, const std::string &where) {
    const std::string s = trim(v);
    if (s.size() < 2 || s.front() != '[' || s.back() != ']') bad(where, "expected a sequence [a, b, c]");
    std::vector<std::string> out;
    std::string cur;
    const std::string body = s.substr(1, s.size() - 2);
    if (trim(body).empty()) return out;
    std::istringstream in(body);
    while (std::getline(in, cur, ',')) out.push_back(trim(cur));
    return out;
}

}  // namespace yaml_subset

// Parse a RunConfig.yaml file.
inline RunConfig readRunConfig(const std::string &path) {
    using namespace yaml_subset;
    std::ifstream in(path);
    if (!in) throw Error(ALENS_E_ARG, "cannot open run configuration " + path);
    RunConfig rc;
    std::string line, block_key;
    int ln = 0;
    while (std::getline(in, line)) {
        ++ln;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        const std::string body = strip_comment(line);
        if (trim(body).empty()) continue;
        const bool indented = std::isspace((unsigned char)body[0]) || body[0] == '-';
        if (indented) {
            if (block_key.empty()) bad(path + ":" + std::to_string(ln), "indented line outside a block");
            rc.blocks[block_key] += body + "\n";
            continue;
        }
        const size_t colon = body.find(':');
        if (colon == std::string::npos) bad(path + ":" + std::to_string(ln), "expected 'key: value'");
        const std::string key = trim(body.substr(0, colon));
        const std::string val = trim(body.substr(colon + 1));
        if (key.empty()) bad(path + ":" + std::to_string(ln), "empty key");
        if (rc.values.count(key) || rc.blocks.count(key)) bad(path + ":" + std::to_string(ln), "duplicate key '" + key + "'");
        if (val.empty()) {
            block_key = key;
            rc.blocks[key];
            continue;
        }
        block_key.clear();
        rc.values[key] = val;
    }
    auto where = [&](const std::string &k) { return path + ": key '" + k + "'"; };
    auto get_d = [&](const char *k, double &x) {
        if (rc.has(k)) x = to_double(rc.values[k], where(k));
    };
    auto get_i = [&](const char *k, int &x) {
        if (rc.has(k)) x = (int)to_int(rc.values[k], where(k));
    };
    auto get_b = [&](const char *k, bool &x) {
        if (rc.has(k)) x = to_bool(rc.values[k], where(k));
    };
    auto get_d3 = [&](const char *k, double *x) {
        if (!rc.has(k)) return;
        const auto it = seq(rc.values[k], where(k));
        if (it.size() != 3) bad(where(k), "expected 3 items");
        for (int a = 0; a < 3; ++a) x[a] = to_double(it[a], where(k));
    };
    auto get_b3 = [&](const char *k, bool *x) {
        if (!rc.has(k)) return;
        const auto it = seq(rc.values[k], where(k));
        if (it.size() != 3) bad(where(k), "expected 3 items");
        for (int a = 0; a < 3; ++a) x[a] = to_bool(it[a], where(k));
    };
    if (rc.has("rngSeed")) {
        const long long s = to_int(rc.values["rngSeed"], where("rngSeed"));
        if (s < 0 || s > 0xffffffffLL) bad(where("rngSeed"), "out of range");
        rc.rngSeed = (unsigned)s;
    }
    get_i("logLevel", rc.logLevel);
    get_i("timerLevel", rc.timerLevel);
    get_d3("simBoxLow", rc.simBoxLow);
    get_d3("simBoxHigh", rc.simBoxHigh);
    get_b3("simBoxPBC", rc.simBoxPBC);
    get_b("monolayer", rc.monolayer);
    get_d3("initBoxLow", rc.initBoxLow);
    get_d3("initBoxHigh", rc.initBoxHigh);
    get_d3("initOrient", rc.initOrient);
    get_b("initCircularX", rc.initCircularX);
    get_i("initPreSteps", rc.initPreSteps);
    get_d("viscosity", rc.viscosity);
    get_d("KBT", rc.KBT);
    get_d("linkKappa", rc.linkKappa);
    get_d("linkGap", rc.linkGap);
    get_b("sylinderFixed", rc.sylinderFixed);
    get_i("sylinderNumber", rc.sylinderNumber);
    get_d("sylinderLength", rc.sylinderLength);
    get_d("sylinderLengthSigma", rc.sylinderLengthSigma);
    get_d("sylinderDiameter", rc.sylinderDiameter);
    get_d("sylinderDiameterColRatio", rc.sylinderDiameterColRatio);
    get_d("sylinderLengthColRatio", rc.sylinderLengthColRatio);
    get_d("sylinderColBuf", rc.sylinderColBuf);
    get_d("dt", rc.dt);
    get_d("timeTotal", rc.timeTotal);
    get_d("timeSnap", rc.timeSnap);
    get_d("conResTol", rc.conResTol);
    get_i("conMaxIte", rc.conMaxIte);
    get_i("conSolverChoice", rc.conSolverChoice);
    return rc;
}

}  // namespace alens
