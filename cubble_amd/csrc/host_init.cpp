// host_init.cpp -- configuration defaults, a small JSON reader for the
// reference's input_parameters.json schema, and the synthetic initial foam.
//
// The reference draws its initial state with cuRAND MTGP32 and places
// bubbles on a jittered grid before a long stabilisation
// (src/simulator.cu:686-956).  This build instead draws radii from the
// configured distribution and positions uniformly in a box sized so that the
// gas fraction is exactly phi, on the host, with its own RNG: SplitMix64
// (Steele, Lea and Flood, OOPSLA 2014) seeding xoshiro256** (Blackman and
// Vigna, 2018), Marsaglia polar normals.  Compiled with -ffp-contract=off.
#include "host_init.hpp"

#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>

namespace cubble_amd {

namespace {

struct Rng {
    uint64_t s[4];
    bool have = false;
    double spare = 0.0;
    static uint64_t splitmix(uint64_t &x) {
        uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
    }
    static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (auto &w : s) w = splitmix(x);
    }
    uint64_t next() {
        const uint64_t out = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return out;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    // Marsaglia polar method (Marsaglia and Bray, 1964): no trigonometric
    // function, so only log and correctly rounded operations are involved.
    double normal() {
        if (have) {
            have = false;
            return spare;
        }
        double u, v, s;
        do {
            u = 2.0 * uniform() - 1.0;
            v = 2.0 * uniform() - 1.0;
            s = u * u + v * v;
        } while (s >= 1.0 || s == 0.0);
        const double m = std::sqrt(-2.0 * std::log(s) / s);
        spare = v * m;
        have = true;
        return u * m;
    }
};

// ---------------------------------------------------------------- JSON
struct JVal {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
    double num = 0.0;
    bool b = false;
    std::string str;
    std::vector<JVal> arr;
    std::map<std::string, JVal> obj;
    const JVal *get(const char *key) const {
        if (kind != Obj) return nullptr;
        auto it = obj.find(key);
        return it == obj.end() ? nullptr : &it->second;
    }
};

struct JParser {
    const std::string &t;
    size_t p = 0;
    std::string err;
    explicit JParser(const std::string &text) : t(text) {}
    void ws() {
        while (p < t.size() && std::isspace((unsigned char)t[p])) ++p;
    }
    bool fail(const char *m) {
        if (err.empty()) {
            std::ostringstream o;
            o << m << " at offset " << p;
            err = o.str();
        }
        return false;
    }
    bool str(std::string &out) {
        if (t[p] != '"') return fail("expected string");
        ++p;
        while (p < t.size() && t[p] != '"') {
            if (t[p] == '\\' && p + 1 < t.size()) {
                ++p;
                const char e = t[p];
                out.push_back(e == 'n' ? '\n' : e == 't' ? '\t' : e);
            } else {
                out.push_back(t[p]);
            }
            ++p;
        }
        if (p >= t.size()) return fail("unterminated string");
        ++p;
        return true;
    }
    bool val(JVal &v) {
        ws();
        if (p >= t.size()) return fail("unexpected end");
        const char ch = t[p];
        if (ch == '{') {
            v.kind = JVal::Obj;
            ++p;
            ws();
            if (p < t.size() && t[p] == '}') {
                ++p;
                return true;
            }
            for (;;) {
                ws();
                std::string k;
                if (!str(k)) return false;
                ws();
                if (p >= t.size() || t[p] != ':') return fail("expected ':'");
                ++p;
                JVal sub;
                if (!val(sub)) return false;
                v.obj[k] = std::move(sub);
                ws();
                if (p < t.size() && t[p] == ',') {
                    ++p;
                    continue;
                }
                if (p < t.size() && t[p] == '}') {
                    ++p;
                    return true;
                }
                return fail("expected ',' or '}'");
            }
        }
        if (ch == '[') {
            v.kind = JVal::Arr;
            ++p;
            ws();
            if (p < t.size() && t[p] == ']') {
                ++p;
                return true;
            }
            for (;;) {
                JVal sub;
                if (!val(sub)) return false;
                v.arr.push_back(std::move(sub));
                ws();
                if (p < t.size() && t[p] == ',') {
                    ++p;
                    continue;
                }
                if (p < t.size() && t[p] == ']') {
                    ++p;
                    return true;
                }
                return fail("expected ',' or ']'");
            }
        }
        if (ch == '"') {
            v.kind = JVal::Str;
            return str(v.str);
        }
        if (t.compare(p, 4, "true") == 0) {
            v.kind = JVal::Bool;
            v.b = true;
            p += 4;
            return true;
        }
        if (t.compare(p, 5, "false") == 0) {
            v.kind = JVal::Bool;
            p += 5;
            return true;
        }
        if (t.compare(p, 4, "null") == 0) {
            p += 4;
            return true;
        }
        const char *b = t.c_str() + p;
        char *e = nullptr;
        const double d = std::strtod(b, &e);
        if (e == b) return fail("bad value");
        v.kind = JVal::Num;
        v.num = d;
        p += (size_t)(e - b);
        return true;
    }
};

bool jnum(const JVal &root, std::initializer_list<const char *> path, double &out) {
    const JVal *v = &root;
    for (const char *k : path) {
        v = v->get(k);
        if (!v) return false;
    }
    if (v->kind == JVal::Num) {
        out = v->num;
        return true;
    }
    if (v->kind == JVal::Bool) {
        out = v->b ? 1.0 : 0.0;
        return true;
    }
    return false;
}

bool jstr(const JVal &root, std::initializer_list<const char *> path, std::string &out) {
    const JVal *v = &root;
    for (const char *k : path) {
        v = v->get(k);
        if (!v) return false;
    }
    if (v->kind != JVal::Str) return false;
    out = v->str;
    return true;
}

}  // namespace

void config_default(cubble_config &cfg) {
    std::memset(&cfg, 0, sizeof(cfg));
    // values of the reference's input_parameters.json
    cfg.dim = 2;
    cfg.num_start = 100000;
    cfg.num_end = 1000;
    cfg.radius_dist = 0;
    cfg.r_mean = 1.0;
    cfg.r_std = 0.21;
    cfg.r_lo = 0.5;
    cfg.r_hi = 1.5;
    cfg.rel_dim[0] = cfg.rel_dim[1] = cfg.rel_dim[2] = 1.0;
    cfg.wall_drag = 0.1;
    cfg.phi = 0.9;
    cfg.mu = 1.0;
    cfg.sigma = 1.0;
    cfg.kappa = 0.1;
    cfg.k_param = 0.001;
    cfg.error_tolerance = 1e-5;
    cfg.dt0 = 1e-4;  // reference src/data_definitions.h:247
    cfg.seed = 426;
    cfg.precision = CUBBLE_PREC_FP64;
    cfg.nbr_cap = 0;
    cfg.cell_cap_factor = 2.0;
    cfg.flow = 0;
    cfg.flow_vel[0] = 0.01;
    cfg.flow_lbb[1] = 0.475;
    cfg.flow_tfr[0] = 0.4;
    cfg.flow_tfr[1] = 0.525;
}

void host_cell_dims(const Consts &c, double max_r, long long cap, int cd[3]) {
    const double edge = 2 * (max_r + c.skin);
    for (int k = 0; k < 3; ++k) {
        double q = k < c.dim ? std::floor(c.interval[k] / edge) : 1.0;
        if (!(q >= 1.0)) q = 1.0;
        if (q > 1.0e6) q = 1.0e6;
        cd[k] = (int)q;
    }
    if (2 == c.dim) cd[2] = 1;
    while ((long long)cd[0] * cd[1] * cd[2] > cap) {
        int big = 0;
        for (int k = 1; k < 3; ++k)
            if (cd[k] > cd[big]) big = k;
        cd[big] -= 1;
    }
}

void lattice_dims(const cubble_config &cfg, int32_t bpd[3]) {
    float n = (float)cfg.num_start;
    const float rx = (float)cfg.rel_dim[0], ry = (float)cfg.rel_dim[1], rz = (float)cfg.rel_dim[2];
    if (3 == cfg.dim) {
        n = std::cbrt(n);
        const float a = std::cbrt(rx / ry), b = std::cbrt(rx / rz), c = std::cbrt(ry / rz);
        bpd[0] = (int32_t)std::ceil(n * a * b);
        bpd[1] = (int32_t)std::ceil(n * c / a);
        bpd[2] = (int32_t)std::ceil(n / (b * c));
    } else {
        n = std::sqrt(n);
        const float a = std::sqrt(rx / ry);
        bpd[0] = (int32_t)std::ceil(n * a);
        bpd[1] = (int32_t)std::ceil(n / a);
        bpd[2] = 1;
    }
}

int generate_foam(const cubble_config &cfg, std::vector<double> &x4, InitOut &out,
                  std::string &err) {
    const int D = cfg.dim;
    const int n = cfg.num_start;
    if (D != 2 && D != 3) {
        err = "box.dimensionality must be 2 or 3";
        return CUBBLE_E_INVALID;
    }
    if (n < 1) {
        err = "bubbles.numStart must be >= 1";
        return CUBBLE_E_INVALID;
    }
    int32_t bpd[3] = {1, 1, 1};
    if (1 == cfg.placement || 2 == cfg.placement) {
        lattice_dims(cfg, bpd);
        const long long lat = (long long)bpd[0] * bpd[1] * bpd[2];
        if (lat != n) {
            err = "init.placement \"grid\": num_start must be the lattice count " +
                  std::to_string(lat);
            return CUBBLE_E_INVALID;
        }
    } else if (0 != cfg.placement) {
        err = "placement must be 0 (uniform), 1 (grid) or 2 (grid in the lattice box)";
        return CUBBLE_E_INVALID;
    }
    if (cfg.radius_dist < 0 || cfg.radius_dist > 2) {
        err = "radius_dist must be 0 (clipped normal), 1 (log-normal) or 2 (the reference's |normal|)";
        return CUBBLE_E_INVALID;
    }
    if (!(cfg.phi > 0.0 && cfg.phi < 1.0) || !(cfg.r_mean > 0.0) || !(cfg.mu > 0.0)) {
        err = "phi must be in (0, 1) and radius.mean, mu positive";
        return CUBBLE_E_INVALID;
    }
    Rng rng(cfg.seed);
    x4.assign((size_t)4 * n, 0.0);
    double *r = x4.data() + (size_t)3 * n;
    for (int i = 0; i < n; ++i) {
        const double z = rng.normal();
        double v;
        if (1 == cfg.radius_dist) {
            v = cfg.r_mean * std::exp(cfg.r_std * z);
        } else if (2 == cfg.radius_dist) {
            // the reference's rule (assignDataToBubbles, src/kernels.cu:1051-1053):
            // |N(mean, std)|, plus minRad (0.1 mean) when at or below minRad
            const double min_rad = 0.1 * cfg.r_mean;
            v = std::fabs(cfg.r_mean + cfg.r_std * z);
            v = v > min_rad ? v : v + min_rad;
        } else {
            v = cfg.r_mean + cfg.r_std * z;
            v = v < cfg.r_lo ? cfg.r_lo : (v > cfg.r_hi ? cfg.r_hi : v);
        }
        r[i] = v;
    }
    // total volume: reference src/kernels.cu:1016-1020
    double vtot = 0.0;
    for (int i = 0; i < n; ++i) {
        double vol = r[i] * r[i] * kPi;
        if (3 == D) vol *= r[i] * 1.33333333333333333333333333;
        vtot += vol;
    }
    double rel = cfg.rel_dim[0] * cfg.rel_dim[1];
    if (3 == D) rel *= cfg.rel_dim[2];
    double t = vtot / (cfg.phi * rel);
    t = 3 == D ? std::cbrt(t) : std::sqrt(t);
    Consts &c = out.c;
    std::memset(&c, 0, sizeof(c));
    c.dim = D;
    for (int k = 0; k < 3; ++k) {
        c.wall[k] = k < D ? cfg.wall[k] : 0;
        c.lbb[k] = 0.0;
        c.tfr[k] = k < D ? t * cfg.rel_dim[k] : 0.0;
        c.interval[k] = c.tfr[k] - c.lbb[k];
    }
    // constants: reference src/simulator.cu:546-556, :820-826
    c.wall_drag = cfg.wall_drag;
    c.skin = 0.3 * cfg.r_mean;
    c.min_rad = 0.1 * cfg.r_mean;
    c.f0 = (float)cfg.sigma * cfg.r_mean / cfg.mu;
    c.k_param = cfg.k_param;
    c.kappa = cfg.kappa;
    c.tol = cfg.error_tolerance;
    // imposed flow: reference src/simulator.cu:561-565, :812-817
    for (int k = 0; k < 3; ++k) {
        c.flow_lbb[k] = c.interval[k] * cfg.flow_lbb[k] + c.lbb[k];
        c.flow_tfr[k] = c.interval[k] * cfg.flow_tfr[k] + c.lbb[k];
        c.flow_vel[k] = cfg.flow_vel[k] * c.f0;
    }
    double vbox = c.interval[0] * c.interval[1];
    if (3 == D) vbox *= c.interval[2];
    const double mult = cfg.phi * vbox / kPi;
    c.vol_mul = 3 == D ? std::cbrt(0.75 * mult) : std::sqrt(mult);
    // mean initial area: reference src/kernels.cu:1065-1069, simulator.cu:708-716
    double ain = 0.0, maxr = 0.0;
    for (int i = 0; i < n; ++i) {
        double area = 2.0 * kPi * r[i];
        if (3 == D) area *= 2.0 * r[i];
        ain += area / n;
        maxr = std::fmax(maxr, r[i]);
    }
    c.avg_area_in = ain;
    if (2 == cfg.placement) {
        // the lattice's own box, edge (float)(2 R) per site (reference
        // src/simulator.cu:577, :603-606); cubble_rescale_box takes it to the
        // target gas fraction later.  The volume multiplier does not depend on
        // the box (phi V_box = the bubbles' volume), and the flow box is kept
        // relative to the box.
        const double d = (double)(float)(2.0 * cfg.r_mean);
        for (int k = 0; k < D; ++k) {
            c.tfr[k] = c.lbb[k] + d * bpd[k];
            c.interval[k] = c.tfr[k] - c.lbb[k];
            c.flow_lbb[k] = c.interval[k] * cfg.flow_lbb[k] + c.lbb[k];
            c.flow_tfr[k] = c.interval[k] * cfg.flow_tfr[k] + c.lbb[k];
        }
    }
    if (1 == cfg.placement || 2 == cfg.placement) {
        // the reference's lattice (assignDataToBubbles, src/kernels.cu:1034-1056):
        // site i at ((i mod bx) / bx, ...) of the box, plus an offset of length
        // R w along the normalised vector of D uniforms (so in the positive
        // orthant), folded back into the box; then, as the reference's rescale
        // (src/simulator.cu:797-817), the relative positions are kept in the
        // box sized for the target gas fraction
        for (int i = 0; i < n; ++i) {
            const double site[3] = {(double)(i % bpd[0]) / bpd[0],
                                    (double)((i / bpd[0]) % bpd[1]) / bpd[1],
                                    (double)(i / (bpd[0] * bpd[1])) / bpd[2]};
            double u[3] = {rng.uniform(), rng.uniform(), 3 == D ? rng.uniform() : 0.0};
            const double w = rng.uniform();
            const double len = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
            for (int k = 0; k < D; ++k) {
                // lattice box: edge 2 R per site
                const double lat = 2.0 * cfg.r_mean * bpd[k];
                double p = site[k] * lat + (len > 0.0 ? u[k] / len : 0.0) * cfg.r_mean * w;
                p = p > 0.0 ? (p < lat ? p : p - lat) : p + lat;
                x4[(size_t)k * n + i] = c.lbb[k] + c.interval[k] * (p / lat);
            }
        }
    } else {
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < D; ++k)
                x4[(size_t)k * n + i] = c.lbb[k] + c.interval[k] * rng.uniform();
    }
    out.max_r = maxr;
    int cd[3];
    host_cell_dims(c, maxr, 1LL << 30, cd);
    const double fac = cfg.cell_cap_factor > 0.0 ? cfg.cell_cap_factor : 2.0;
    double cap = fac * (double)cd[0] * cd[1] * cd[2];
    if (cap < 64.0) cap = 64.0;
    out.cell_cap = (int32_t)std::fmin(cap, 2.0e9);
    return CUBBLE_OK;
}

int parse_config_json(const char *path, cubble_config &cfg, std::string &err) {
    std::ifstream f(path);
    if (!f.is_open()) {
        err = std::string("couldn't open input file ") + (path ? path : "(null)");
        return CUBBLE_E_IO;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    JParser jp(text);
    JVal root;
    if (!jp.val(root) || root.kind != JVal::Obj) {
        err = "malformed JSON: " + jp.err;
        return CUBBLE_E_IO;
    }
    config_default(cfg);
    double v = 0.0;
    // reference keys: src/simulator.cu:538-575, :690, :702
    if (jnum(root, {"bubbles", "numStart"}, v)) cfg.num_start = (int32_t)v;
    if (jnum(root, {"bubbles", "numEnd"}, v)) cfg.num_end = (int32_t)v;
    if (jnum(root, {"bubbles", "radius", "mean"}, v)) cfg.r_mean = v;
    if (jnum(root, {"bubbles", "radius", "std"}, v)) cfg.r_std = v;
    if (jnum(root, {"box", "dimensionality"}, v)) cfg.dim = (int32_t)v;
    if (jnum(root, {"box", "relativeDimensions", "x"}, v)) cfg.rel_dim[0] = v;
    if (jnum(root, {"box", "relativeDimensions", "y"}, v)) cfg.rel_dim[1] = v;
    if (jnum(root, {"box", "relativeDimensions", "z"}, v)) cfg.rel_dim[2] = v;
    if (jnum(root, {"box", "wall", "drag"}, v)) cfg.wall_drag = v;
    if (jnum(root, {"box", "wall", "x"}, v)) cfg.wall[0] = v == 1.0;
    if (jnum(root, {"box", "wall", "y"}, v)) cfg.wall[1] = v == 1.0;
    if (jnum(root, {"box", "wall", "z"}, v)) cfg.wall[2] = v == 1.0;
    if (jnum(root, {"constants", "phi", "value"}, v)) cfg.phi = v;
    if (jnum(root, {"constants", "mu", "value"}, v)) cfg.mu = v;
    if (jnum(root, {"constants", "sigma", "value"}, v)) cfg.sigma = v;
    if (jnum(root, {"constants", "kappa", "value"}, v)) cfg.kappa = v;
    if (jnum(root, {"constants", "K", "value"}, v)) cfg.k_param = v;
    if (jnum(root, {"errorTolerance", "value"}, v)) cfg.error_tolerance = v;
    if (jnum(root, {"rngSeed", "value"}, v)) cfg.seed = (uint64_t)v;
    if (jnum(root, {"flow", "impose"}, v)) cfg.flow = v == 1.0;
    const char *ax[3] = {"x", "y", "z"};
    for (int k = 0; k < 3; ++k) {
        if (jnum(root, {"flow", "velocity", ax[k]}, v)) cfg.flow_vel[k] = v;
        if (jnum(root, {"flow", "lbb", ax[k]}, v)) cfg.flow_lbb[k] = v;
        if (jnum(root, {"flow", "tfr", ax[k]}, v)) cfg.flow_tfr[k] = v;
    }
    cfg.r_lo = 0.5 * cfg.r_mean;
    cfg.r_hi = 1.5 * cfg.r_mean;
    // keys this build adds (ignored by the reference, README.md:66-69).  A
    // file without an "init" object is the reference's own input: its init
    // places the bubbles on the lattice and takes the radii as |N(mean, std)|
    // raised by minRad at or below minRad (src/simulator.cu:576-606,
    // src/kernels.cu:1034-1053).  With an "init" object the defaults are
    // uniform placement and the normal clipped to radiusClip.
    if (!root.get("init")) {
        cfg.placement = 1;
        cfg.radius_dist = 2;
    }
    std::string s;
    if (jstr(root, {"init", "radiusDistribution"}, s)) {
        if (s == "lognormal") cfg.radius_dist = 1;
        else if (s == "normal") cfg.radius_dist = 0;
        else if (s == "reference") cfg.radius_dist = 2;
        else {
            err = "init.radiusDistribution must be \"normal\", \"lognormal\" or \"reference\"";
            return CUBBLE_E_INVALID;
        }
    }
    const JVal *clip = root.get("init") ? root.get("init")->get("radiusClip") : nullptr;
    if (clip && clip->kind == JVal::Arr && clip->arr.size() == 2) {
        cfg.r_lo = clip->arr[0].num;
        cfg.r_hi = clip->arr[1].num;
    }
    if (jnum(root, {"init", "dt0"}, v)) cfg.dt0 = v;
    if (jnum(root, {"init", "neighbourCapacity"}, v)) cfg.nbr_cap = (int32_t)v;
    if (jnum(root, {"init", "cellCapacityFactor"}, v)) cfg.cell_cap_factor = v;
    if (jnum(root, {"run", "maxSteps"}, v)) cfg.max_steps = (int64_t)v;
    if (jstr(root, {"init", "placement"}, s)) {
        if (s == "grid") cfg.placement = 1;
        else if (s == "uniform") cfg.placement = 0;
        else {
            err = "init.placement must be \"uniform\" or \"grid\"";
            return CUBBLE_E_INVALID;
        }
    }
    if (jstr(root, {"init", "precision"}, s)) {
        if (s == "mixed") cfg.precision = CUBBLE_PREC_MIXED;
        else if (s == "fp64") cfg.precision = CUBBLE_PREC_FP64;
        else {
            err = "init.precision must be \"fp64\" or \"mixed\"";
            return CUBBLE_E_INVALID;
        }
    }
    // the reference places numStart rounded up to a full lattice
    // (src/simulator.cu:576-600)
    if (1 == cfg.placement) {
        int32_t bpd[3];
        lattice_dims(cfg, bpd);
        cfg.num_start = bpd[0] * bpd[1] * bpd[2];
    }
    return CUBBLE_OK;
}

int parse_run_json(const char *path, RunSettings &rs, std::string &err) {
    rs = RunSettings();
    std::ifstream f(path);
    if (!f.is_open()) {
        err = std::string("couldn't open input file ") + (path ? path : "(null)");
        return CUBBLE_E_IO;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    JParser jp(text);
    JVal root;
    if (!jp.val(root) || root.kind != JVal::Obj) {
        err = "malformed JSON: " + jp.err;
        return CUBBLE_E_IO;
    }
    double v = 0.0;
    if (jnum(root, {"snapShot", "frequency"}, v)) rs.snapshot_frequency = v;
    jstr(root, {"snapShot", "filename"}, rs.snapshot_name);
    if (jnum(root, {"stabilization", "steps"}, v)) rs.stabilization_steps = (int64_t)v;
    if (jnum(root, {"stabilization", "maxDeltaEnergy"}, v)) rs.max_delta_energy = v;
    if (rs.stabilization_steps < 0) {
        err = "stabilization.steps must be >= 0";
        return CUBBLE_E_INVALID;
    }
    return CUBBLE_OK;
}

namespace {

bool read_cache(const char *path, JVal &root) {
    std::ifstream f(path);
    if (!f.is_open()) return false;
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    JParser jp(text);
    return jp.val(root) && root.kind == JVal::Obj;
}

}  // namespace

bool tuning_cache_lookup(const char *path, const std::string &key, int shape[3]) {
    JVal root;
    if (!path || !read_cache(path, root)) return false;
    const JVal *e = root.get(key.c_str());
    if (!e) return false;
    double v[3];
    if (!jnum(*e, {"bubbles_per_thread"}, v[0]) || !jnum(*e, {"search_tile"}, v[1]) ||
        !jnum(*e, {"graph_steps"}, v[2]))
        return false;
    for (int k = 0; k < 3; ++k) shape[k] = (int)v[k];
    return true;
}

int tuning_cache_store(const char *path, const std::string &key, const int shape[3], double rate,
                       std::string &err) {
    // keep the other entries; an unreadable file is replaced
    JVal root;
    std::map<std::string, std::string> entries;
    if (read_cache(path, root)) {
        for (const auto &kv : root.obj) {
            double v[4] = {0, 0, 0, 0};
            if (!jnum(kv.second, {"bubbles_per_thread"}, v[0]) ||
                !jnum(kv.second, {"search_tile"}, v[1]) || !jnum(kv.second, {"graph_steps"}, v[2]))
                continue;
            jnum(kv.second, {"rate"}, v[3]);
            char buf[256];
            std::snprintf(buf, sizeof(buf),
                          "{\"bubbles_per_thread\": %d, \"search_tile\": %d, \"graph_steps\": %d, "
                          "\"rate\": %.6g}",
                          (int)v[0], (int)v[1], (int)v[2], v[3]);
            entries[kv.first] = buf;
        }
    }
    char buf[256];
    std::snprintf(buf, sizeof(buf),
                  "{\"bubbles_per_thread\": %d, \"search_tile\": %d, \"graph_steps\": %d, "
                  "\"rate\": %.6g}",
                  shape[0], shape[1], shape[2], rate);
    entries[key] = buf;
    const std::string tmp = std::string(path) + ".tmp";
    std::ofstream f(tmp);
    if (!f.is_open()) {
        err = std::string("couldn't write tuning cache ") + path;
        return CUBBLE_E_IO;
    }
    f << "{\n";
    size_t k = 0;
    for (const auto &kv : entries)
        f << "  \"" << kv.first << "\": " << kv.second << (++k < entries.size() ? ",\n" : "\n");
    f << "}\n";
    f.close();
    if (!f || std::rename(tmp.c_str(), path) != 0) {
        err = std::string("couldn't write tuning cache ") + path;
        return CUBBLE_E_IO;
    }
    return CUBBLE_OK;
}

}  // namespace cubble_amd
