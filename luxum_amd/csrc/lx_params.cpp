// lx_params.cpp -- parameters.yaml reader and parameter checks.
//
// Replaces class reader (lib/io/reader.cc:68-162).  The reference reads five
// consecutive YAML flow mappings with the legacy yaml-cpp API
// (Parser::GetNextDocument, reader.cc:76-127); yaml-cpp is not available, so
// this is a small tolerant parser for exactly that shape: an optional UTF-8
// BOM, '#' comments, quoted keys, trailing commas, and an optional sixth
// mapping with the extension keys.  Keys are looked up by their exact names
// in the same document order as the reference.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/luxum.h"
#include "lx_util.hpp"

namespace {

using Doc = std::map<std::string, std::string>;

struct Parser {
    std::string s;
    size_t p = 0;
    std::string err;

    void skip_ws() {
        for (;;) {
            while (p < s.size() && std::isspace((unsigned char)s[p])) ++p;
            if (p < s.size() && s[p] == '#') {
                while (p < s.size() && s[p] != '\n') ++p;
                continue;
            }
            if (p + 2 < s.size() && s.compare(p, 3, "---") == 0) {   // document separator
                while (p < s.size() && s[p] != '\n') ++p;
                continue;
            }
            if (p + 2 < s.size() && s.compare(p, 3, "...") == 0) {
                while (p < s.size() && s[p] != '\n') ++p;
                continue;
            }
            return;
        }
    }

    bool quoted(std::string &out) {
        char q = s[p];
        ++p;
        out.clear();
        while (p < s.size() && s[p] != q) {
            if (s[p] == '\\' && q == '"' && p + 1 < s.size()) ++p;
            out.push_back(s[p++]);
        }
        if (p >= s.size()) { err = "unterminated string"; return false; }
        ++p;
        return true;
    }

    bool scalar(std::string &out) {
        skip_ws();
        if (p >= s.size()) { err = "unexpected end of file"; return false; }
        if (s[p] == '"' || s[p] == '\'') return quoted(out);
        size_t b = p;
        while (p < s.size() && s[p] != ',' && s[p] != '}' && s[p] != '#' && s[p] != '\n' && s[p] != ':') ++p;
        out = s.substr(b, p - b);
        while (!out.empty() && std::isspace((unsigned char)out.back())) out.pop_back();
        if (out.empty()) { err = "empty value"; return false; }
        return true;
    }

    // Returns 1 on a document, 0 at end of input, -1 on error.
    int doc(Doc &d) {
        d.clear();
        skip_ws();
        if (p >= s.size()) return 0;
        if (s[p] != '{') { err = "expected '{' at offset " + std::to_string(p); return -1; }
        ++p;
        for (;;) {
            skip_ws();
            if (p >= s.size()) { err = "unterminated mapping"; return -1; }
            if (s[p] == '}') { ++p; return 1; }
            if (s[p] == ',') { ++p; continue; }
            std::string k, v;
            if (!scalar(k)) return -1;
            skip_ws();
            if (p >= s.size() || s[p] != ':') { err = "expected ':' after key \"" + k + "\""; return -1; }
            ++p;
            if (!scalar(v)) { err += " (key \"" + k + "\")"; return -1; }
            d[k] = v;
        }
    }
};

bool get_str(const Doc &d, const char *key, std::string &out) {
    auto it = d.find(key);
    if (it == d.end()) return false;
    out = it->second;
    return true;
}

lx_status need_double(const Doc &d, const char *key, double *out) {
    std::string v;
    if (!get_str(d, key, v)) return lx::fail(LX_EINVAL, "parameters: missing key \"%s\"", key);
    char *end = nullptr;
    double x = std::strtod(v.c_str(), &end);
    if (end == v.c_str() || *end) return lx::fail(LX_EINVAL, "parameters: \"%s\" is not a number: %s", key, v.c_str());
    *out = x;
    return LX_OK;
}

lx_status need_int(const Doc &d, const char *key, int32_t *out) {
    std::string v;
    if (!get_str(d, key, v)) return lx::fail(LX_EINVAL, "parameters: missing key \"%s\"", key);
    char *end = nullptr;
    long x = std::strtol(v.c_str(), &end, 10);
    if (end == v.c_str() || *end) return lx::fail(LX_EINVAL, "parameters: \"%s\" is not an integer: %s", key, v.c_str());
    *out = (int32_t)x;
    return LX_OK;
}

lx_status need_bool(const Doc &d, const char *key, int32_t *out) {
    std::string v;
    if (!get_str(d, key, v)) return lx::fail(LX_EINVAL, "parameters: missing key \"%s\"", key);
    std::string l;
    for (char c : v) l.push_back((char)std::tolower((unsigned char)c));
    if (l == "true" || l == "yes" || l == "on" || l == "y") { *out = 1; return LX_OK; }
    if (l == "false" || l == "no" || l == "off" || l == "n") { *out = 0; return LX_OK; }
    return lx::fail(LX_EINVAL, "parameters: \"%s\" is not a boolean: %s", key, v.c_str());
}

#define TRY(x)                          \
    do {                                \
        lx_status _s = (x);             \
        if (_s != LX_OK) return _s;     \
    } while (0)

lx_status opt_double(const Doc &d, const char *key, double *out) {
    if (!d.count(key)) return LX_OK;
    return need_double(d, key, out);
}
lx_status opt_int(const Doc &d, const char *key, int32_t *out) {
    if (!d.count(key)) return LX_OK;
    return need_int(d, key, out);
}
lx_status opt_enum(const Doc &d, const char *key, int32_t *out, const std::vector<std::string> &names) {
    std::string v;
    if (!get_str(d, key, v)) return LX_OK;
    for (size_t i = 0; i < names.size(); ++i)
        if (v == names[i]) { *out = (int32_t)i; return LX_OK; }
    return lx::fail(LX_EINVAL, "parameters: bad value for \"%s\": %s", key, v.c_str());
}

}  // namespace

extern "C" lx_status lx_params_default(lx_params *p) {
    if (!p) return lx::fail(LX_EINVAL, "lx_params_default: null pointer");
    std::memset(p, 0, sizeof(*p));
    p->nx = p->ny = p->nz = 32;
    p->dx = p->dy = p->dz = 5e-8;
    p->courant = 0.5;
    p->num_timesteps = 100;
    p->npx = p->npy = p->n_threads = 1;
    p->k_pw[0] = 1.0;
    p->e_pw[2] = 1.0;
    p->sigma_ratio = 1.0;
    p->sigma_m = 3.0;
    p->kappa_max = 5.0;
    p->a_max = 0.05;
    p->a_m = 1.0;
    p->precision = LX_FP32;
    p->storage_scale = 0.0;
    p->incident = LX_INC_ANALYTIC;
    p->pw_sigma = std::pow(10, -4) / 0.5;   // maxwell.cc:71
    p->pw_wave = LX_WAVE_GAUSSIAN_SINE;
    p->pw_lambda0 = 1e-6;
    p->pw_fwhm = 3.0;
    p->pw_delay = 4.0;
    p->src_kind = LX_SRC_HARD_CW;
    p->src_comp = LX_EZ;
    p->src_wave = LX_WAVE_GAUSSIAN_SINE;
    p->src_lambda0 = 1e-6;
    p->src_fwhm = 3.0;
    p->src_delay = 4.0;
    return LX_OK;
}

extern "C" lx_status lx_params_load_yaml(const char *path, lx_params *p) {
    if (!path || !p) return lx::fail(LX_EINVAL, "lx_params_load_yaml: null pointer");
    std::ifstream in(path, std::ios::binary);
    if (!in) return lx::fail(LX_EIO, "cannot open %s", path);
    std::stringstream ss;
    ss << in.rdbuf();
    Parser ps;
    ps.s = ss.str();
    if (ps.s.size() >= 3 && (unsigned char)ps.s[0] == 0xEF && (unsigned char)ps.s[1] == 0xBB &&
        (unsigned char)ps.s[2] == 0xBF)
        ps.p = 3;
    std::vector<Doc> docs;
    for (;;) {
        Doc d;
        int r = ps.doc(d);
        if (r < 0) return lx::fail(LX_EINVAL, "%s: %s", path, ps.err.c_str());
        if (r == 0) break;
        docs.push_back(d);
    }
    if (docs.size() < 5) return lx::fail(LX_EINVAL, "%s: expected 5 documents, found %zu", path, docs.size());
    lx_params_default(p);
    const Doc &d1 = docs[0], &d2 = docs[1], &d3 = docs[2], &d4 = docs[3], &d5 = docs[4];
    TRY(need_bool(d1, "isPlanar", &p->is_planar));
    TRY(need_int(d2, "X Length", &p->nx));
    TRY(need_int(d2, "Y Length", &p->ny));
    TRY(need_int(d2, "Z Length", &p->nz));
    TRY(need_double(d2, "X Resolution", &p->dx));
    TRY(need_double(d2, "Y Resolution", &p->dy));
    TRY(need_double(d2, "Z Resolution", &p->dz));
    TRY(need_double(d2, "Courant factor", &p->courant));
    TRY(need_int(d2, "Number of timesteps", &p->num_timesteps));
    TRY(need_int(d3, "X Number of Procs", &p->npx));
    TRY(need_int(d3, "Y Number of Procs", &p->npy));
    TRY(need_int(d3, "Number of Threads per node", &p->n_threads));
    TRY(need_bool(d4, "usePointSource", &p->use_ps));
    TRY(need_int(d4, "Source X Pos", &p->src_pos[0]));
    TRY(need_int(d4, "Source Y Pos", &p->src_pos[1]));
    TRY(need_int(d4, "Source Z Pos", &p->src_pos[2]));
    TRY(need_double(d4, "Source Amplitude", &p->src_amp));
    TRY(need_double(d4, "Source Omega", &p->src_omega));
    TRY(need_bool(d4, "usePW", &p->use_pw));
    TRY(need_double(d4, "kx_pw", &p->k_pw[0]));
    TRY(need_double(d4, "ky_pw", &p->k_pw[1]));
    TRY(need_double(d4, "kz_pw", &p->k_pw[2]));
    TRY(need_double(d4, "Ex_pw", &p->e_pw[0]));
    TRY(need_double(d4, "Ey_pw", &p->e_pw[1]));
    TRY(need_double(d4, "Ez_pw", &p->e_pw[2]));
    TRY(need_double(d4, "off_pw", &p->off_pw));
    TRY(need_int(d4, "x1_pw", &p->tf_lo[0]));
    TRY(need_int(d4, "y1_pw", &p->tf_lo[1]));
    TRY(need_int(d4, "z1_pw", &p->tf_lo[2]));
    TRY(need_int(d4, "x2_pw", &p->tf_hi[0]));
    TRY(need_int(d4, "y2_pw", &p->tf_hi[1]));
    TRY(need_int(d4, "z2_pw", &p->tf_hi[2]));
    TRY(need_bool(d5, "usePML", &p->use_pml));
    TRY(need_double(d5, "Thickness of PML", &p->pml_thickness));
    TRY(need_double(d5, "sigma_max/sigma_optimal", &p->sigma_ratio));
    TRY(need_double(d5, "m value of sigma", &p->sigma_m));
    TRY(need_double(d5, "kappa max", &p->kappa_max));
    TRY(need_double(d5, "max value of a", &p->a_max));
    TRY(need_double(d5, "m value for a", &p->a_m));
    p->src_comp = p->is_planar ? LX_EY : LX_EZ;
    if (docs.size() >= 6) {
        const Doc &d6 = docs[5];
        TRY(opt_enum(d6, "precision", &p->precision, {"fp32", "fp16"}));
        TRY(opt_enum(d6, "incident", &p->incident, {"analytic", "aux1d"}));
        TRY(opt_double(d6, "pw sigma", &p->pw_sigma));
        TRY(opt_enum(d6, "pw waveform", &p->pw_wave, {"gaussian", "gaussian_sine"}));
        TRY(opt_double(d6, "pw lambda0", &p->pw_lambda0));
        TRY(opt_double(d6, "pw fwhm periods", &p->pw_fwhm));
        TRY(opt_double(d6, "pw delay", &p->pw_delay));
        TRY(opt_enum(d6, "source kind", &p->src_kind, {"hard_cw", "soft"}));
        TRY(opt_enum(d6, "source component", &p->src_comp, {"Ex", "Ey", "Ez"}));
        TRY(opt_enum(d6, "source waveform", &p->src_wave, {"gaussian", "gaussian_sine"}));
        TRY(opt_double(d6, "source lambda0", &p->src_lambda0));
        TRY(opt_double(d6, "source fwhm periods", &p->src_fwhm));
        TRY(opt_double(d6, "source delay", &p->src_delay));
        TRY(opt_int(d6, "snapshot every", &p->snapshot_every));
        TRY(opt_int(d6, "monitor every", &p->monitor_every));
        TRY(opt_double(d6, "storage scale", &p->storage_scale));
    }
    return lx_params_check(p);
}

extern "C" lx_status lx_params_check(const lx_params *p) {
    if (!p) return lx::fail(LX_EINVAL, "lx_params_check: null pointer");
    // reader::checkData (reader.cc:141-162)
    if (p->npx < 1 || p->npy < 1) return lx::fail(LX_EINVAL, "number of procs must be >= 1");
    if ((p->nx - 2) % p->npx != 0)
        return lx::fail(LX_EINVAL, "(Number of points - 2) mod Number of processors is not satisfied in X-direction");
    if ((p->ny - 2) % p->npy != 0)
        return lx::fail(LX_EINVAL, "(Number of points - 2) mod Number of processors is not satisfied in Y-direction");
    if (p->is_planar && p->ny != 1)
        return lx::fail(LX_EINVAL, "put number of grid points in Y-direction as 1 for planar solutions");
    if (p->is_planar && p->npy != 1)
        return lx::fail(LX_EINVAL, "put number of processors in Y-direction as 1 for planar solutions");
    // solver preconditions (one device holds the whole domain)
    const int N[3] = {p->nx, p->is_planar ? 1 : p->ny, p->nz};
    const double d[3] = {p->dx, p->dy, p->dz};
    for (int a = 0; a < 3; ++a) {
        if (p->is_planar && a == 1) continue;
        if (N[a] < 3) return lx::fail(LX_EINVAL, "grid needs at least 3 nodes along each axis");
        if (!(d[a] > 0)) return lx::fail(LX_EINVAL, "resolution must be positive");
    }
    if ((long long)N[0] * N[1] * N[2] >= (1LL << 31)) return lx::fail(LX_EINVAL, "grid exceeds 2^31 nodes");
    if (!(p->courant > 0)) return lx::fail(LX_EINVAL, "Courant factor must be positive");
    if (p->num_timesteps < 0) return lx::fail(LX_EINVAL, "Number of timesteps must be >= 0");
    if (p->precision != LX_FP32 && p->precision != LX_FP16) return lx::fail(LX_EINVAL, "bad precision");
    if (p->storage_scale != 0.0) {
        int ex = 0;
        if (!(p->storage_scale > 0) || !std::isfinite(p->storage_scale) || std::frexp(p->storage_scale, &ex) != 0.5 ||
            ex < -60 || ex > 62)
            return lx::fail(LX_EINVAL, "storage scale must be 0 (automatic) or a power of two in [2^-61, 2^61]");
    }
    if (p->use_pml) {
        for (int a = 0; a < 3; ++a) {
            if (p->is_planar && a == 1) continue;
            double dd = p->pml_thickness / d[a];
            if (!(dd >= 2.0)) return lx::fail(LX_EINVAL, "CPML must be at least 2 cells thick");
            if (2.0 * dd >= N[a] - 1) return lx::fail(LX_EINVAL, "CPML slabs overlap: thickness too large for the grid");
        }
        if (!(p->sigma_m >= 0) || !(p->a_m >= 0) || !(p->kappa_max >= 1))
            return lx::fail(LX_EINVAL, "bad CPML grading parameters");
    }
    if (p->use_pw) {
        const double *k = p->k_pw, *e = p->e_pw;
        double hx = k[1] * e[2] - k[2] * e[1], hy = k[2] * e[0] - k[0] * e[2], hz = k[0] * e[1] - k[1] * e[0];
        if (hx * hx + hy * hy + hz * hz == 0.0) return lx::fail(LX_EINVAL, "plane wave: k x E is zero");
        for (int a = 0; a < 3; ++a) {
            if (p->is_planar && a == 1) continue;
            if (p->tf_lo[a] < 1 || p->tf_hi[a] > N[a] - 2 || p->tf_lo[a] >= p->tf_hi[a])
                return lx::fail(LX_EINVAL, "TF/SF box must satisfy 1 <= lo < hi <= N-2 on every axis");
        }
        if (p->incident == LX_INC_AUX1D) {
            int ax = -1, n = 0;
            for (int a = 0; a < 3; ++a)
                if (k[a] != 0.0) { ax = a; ++n; }
            if (n != 1 || k[ax] != 1.0 || e[ax] != 0.0)
                return lx::fail(LX_EINVAL, "aux1d incident line needs k along +x, +y or +z and E normal to k");
            if (p->is_planar && ax == 1) return lx::fail(LX_EINVAL, "aux1d: planar runs propagate in x or z");
            if (!(p->pw_lambda0 > 0) || !(p->pw_fwhm > 0)) return lx::fail(LX_EINVAL, "aux1d: bad waveform");
        } else if (p->incident == LX_INC_ANALYTIC) {
            if (!(p->pw_sigma > 0)) return lx::fail(LX_EINVAL, "plane wave sigma must be positive");
        } else {
            return lx::fail(LX_EINVAL, "bad incident mode");
        }
    }
    if (p->use_ps) {
        int c = p->src_comp;
        if (c < 0 || c > 2) return lx::fail(LX_EINVAL, "point source component must be Ex, Ey or Ez");
        if (p->is_planar && c != LX_EY) return lx::fail(LX_EINVAL, "planar point source drives Ey");
        int ext[3];
        lx::comp_extent(N, p->is_planar, c, ext);
        int pos[3] = {p->src_pos[0], p->is_planar ? 0 : p->src_pos[1], p->src_pos[2]};
        for (int a = 0; a < 3; ++a)
            if (pos[a] < 0 || pos[a] >= ext[a]) return lx::fail(LX_EINVAL, "point source outside the grid");
        if (p->src_kind != LX_SRC_HARD_CW && p->src_kind != LX_SRC_SOFT) return lx::fail(LX_EINVAL, "bad source kind");
        if (p->src_kind == LX_SRC_SOFT && (!(p->src_lambda0 > 0) || !(p->src_fwhm > 0)))
            return lx::fail(LX_EINVAL, "bad source waveform");
    }
    return LX_OK;
}
