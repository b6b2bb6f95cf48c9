// topology_json.cpp -- PrismPrepTopology JSON ingest and export (host only).
//
// Restates the reader and validator of the reference's on-disk topology,
// PrismPrepTopology (crates/prism-nhs/src/input.rs:1027-1182): load
// (:1970-1993), validate (:1995-2300), the harmonic lowering to the
// U = k*delta^2 convention (lower_harmonic_parameters, :467-538, and
// HarmonicParameterContract::validate_for_topology, :300-323, whose digest is
// the SHA-256 of the contract's canonical JSON, python_canonical_json :44-110),
// the exact-exception digests (:1934-1968) and hydrogen mass repartitioning
// (apply_hmr_with_historical_repeat, :2494-2565).  The JSON parser, SHA-256
// and writer are this file's own.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "prism_topology.h"

namespace {

thread_local std::string g_error = "no error";

int fail(int rc, const std::string &msg) {
    g_error = msg;
    return rc;
}

// ---------------------------------------------------------------------------
// SHA-256 (FIPS 180-4)
// ---------------------------------------------------------------------------
class Sha256 {
  public:
    Sha256() { reset(); }
    void reset() {
        static const uint32_t h0[8] = {0x6a09e667u, 0xbb67ae85u, 0x3c6ef372u, 0xa54ff53au,
                                       0x510e527fu, 0x9b05688cu, 0x1f83d9abu, 0x5be0cd19u};
        std::memcpy(h_, h0, sizeof(h_));
        len_ = 0;
        fill_ = 0;
    }
    void update(const void *data, size_t n) {
        const uint8_t *p = static_cast<const uint8_t *>(data);
        len_ += n;
        while (n > 0) {
            const size_t take = std::min(n, (size_t)64 - fill_);
            std::memcpy(buf_ + fill_, p, take);
            fill_ += take;
            p += take;
            n -= take;
            if (fill_ == 64) {
                block(buf_);
                fill_ = 0;
            }
        }
    }
    void update(const std::string &s) { update(s.data(), s.size()); }
    template <class T> void update_le(T v) {  // little-endian bytes of an integer or IEEE value
        uint8_t b[sizeof(T)];
        std::memcpy(b, &v, sizeof(T));  // x86-64 / gfx hosts are little-endian
        update(b, sizeof(T));
    }
    std::string hex() {
        const uint64_t bits = len_ * 8;
        const uint8_t one = 0x80, zero = 0;
        update(&one, 1);
        while (fill_ != 56) update(&zero, 1);
        uint8_t lenb[8];
        for (int i = 0; i < 8; ++i) lenb[i] = (uint8_t)(bits >> (56 - 8 * i));
        update(lenb, 8);
        static const char *dig = "0123456789abcdef";
        std::string out;
        for (int i = 0; i < 8; ++i)
            for (int k = 28; k >= 0; k -= 4) out.push_back(dig[(h_[i] >> k) & 15u]);
        reset();
        return out;
    }

  private:
    static uint32_t ror(uint32_t x, int r) { return (x >> r) | (x << (32 - r)); }
    void block(const uint8_t *p) {
        static const uint32_t K[64] = {
            0x428a2f98u, 0x71374491u, 0xb5c0fbcfu, 0xe9b5dba5u, 0x3956c25bu, 0x59f111f1u, 0x923f82a4u, 0xab1c5ed5u,
            0xd807aa98u, 0x12835b01u, 0x243185beu, 0x550c7dc3u, 0x72be5d74u, 0x80deb1feu, 0x9bdc06a7u, 0xc19bf174u,
            0xe49b69c1u, 0xefbe4786u, 0x0fc19dc6u, 0x240ca1ccu, 0x2de92c6fu, 0x4a7484aau, 0x5cb0a9dcu, 0x76f988dau,
            0x983e5152u, 0xa831c66du, 0xb00327c8u, 0xbf597fc7u, 0xc6e00bf3u, 0xd5a79147u, 0x06ca6351u, 0x14292967u,
            0x27b70a85u, 0x2e1b2138u, 0x4d2c6dfcu, 0x53380d13u, 0x650a7354u, 0x766a0abbu, 0x81c2c92eu, 0x92722c85u,
            0xa2bfe8a1u, 0xa81a664bu, 0xc24b8b70u, 0xc76c51a3u, 0xd192e819u, 0xd6990624u, 0xf40e3585u, 0x106aa070u,
            0x19a4c116u, 0x1e376c08u, 0x2748774cu, 0x34b0bcb5u, 0x391c0cb3u, 0x4ed8aa4au, 0x5b9cca4fu, 0x682e6ff3u,
            0x748f82eeu, 0x78a5636fu, 0x84c87814u, 0x8cc70208u, 0x90befffau, 0xa4506cebu, 0xbef9a3f7u, 0xc67178f2u};
        uint32_t w[64];
        for (int i = 0; i < 16; ++i)
            w[i] = (uint32_t)p[4 * i] << 24 | (uint32_t)p[4 * i + 1] << 16 | (uint32_t)p[4 * i + 2] << 8 | p[4 * i + 3];
        for (int i = 16; i < 64; ++i) {
            const uint32_t s0 = ror(w[i - 15], 7) ^ ror(w[i - 15], 18) ^ (w[i - 15] >> 3);
            const uint32_t s1 = ror(w[i - 2], 17) ^ ror(w[i - 2], 19) ^ (w[i - 2] >> 10);
            w[i] = w[i - 16] + s0 + w[i - 7] + s1;
        }
        uint32_t a = h_[0], b = h_[1], c = h_[2], d = h_[3], e = h_[4], f = h_[5], g = h_[6], h = h_[7];
        for (int i = 0; i < 64; ++i) {
            const uint32_t S1 = ror(e, 6) ^ ror(e, 11) ^ ror(e, 25);
            const uint32_t ch = (e & f) ^ (~e & g);
            const uint32_t t1 = h + S1 + ch + K[i] + w[i];
            const uint32_t S0 = ror(a, 2) ^ ror(a, 13) ^ ror(a, 22);
            const uint32_t mj = (a & b) ^ (a & c) ^ (b & c);
            const uint32_t t2 = S0 + mj;
            h = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
        }
        h_[0] += a; h_[1] += b; h_[2] += c; h_[3] += d; h_[4] += e; h_[5] += f; h_[6] += g; h_[7] += h;
    }
    uint32_t h_[8];
    uint64_t len_;
    uint8_t buf_[64];
    size_t fill_;
};

// ---------------------------------------------------------------------------
// JSON document model and recursive-descent parser (RFC 8259)
// ---------------------------------------------------------------------------
struct Json {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    bool integral = false;       // number written without fraction or exponent
    long long ival = 0;          // its value when integral (and in range)
    std::string str;             // String, or the number's source text
    std::vector<Json> arr;
    std::vector<std::pair<std::string, Json>> obj;

    const Json *get(const char *key) const {
        for (const auto &kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
};

class Parser {
  public:
    Parser(const char *p, size_t n) : p_(p), end_(p + n) {}
    bool parse(Json &out, std::string &err) {
        ws();
        if (!value(out, 0)) {
            err = err_.empty() ? "malformed JSON" : err_;
            err += " at byte " + std::to_string(pos());
            return false;
        }
        ws();
        if (p_ != end_) {
            err = "trailing data after the JSON document at byte " + std::to_string(pos());
            return false;
        }
        return true;
    }

  private:
    size_t pos() const { return (size_t)(p_ - begin_()); }
    const char *begin_() const { return end_ - total_; }
    void ws() {
        while (p_ < end_ && (*p_ == ' ' || *p_ == '\t' || *p_ == '\n' || *p_ == '\r')) ++p_;
    }
    bool lit(const char *s) {
        const size_t n = std::strlen(s);
        if ((size_t)(end_ - p_) < n || std::memcmp(p_, s, n) != 0) return false;
        p_ += n;
        return true;
    }
    bool value(Json &v, int depth) {
        if (depth > 64) {
            err_ = "nesting deeper than 64";
            return false;
        }
        if (p_ >= end_) return false;
        switch (*p_) {
        case '{': return object(v, depth);
        case '[': return array(v, depth);
        case '"': v.kind = Json::String; return string(v.str);
        case 't': v.kind = Json::Bool; v.b = true; return lit("true");
        case 'f': v.kind = Json::Bool; v.b = false; return lit("false");
        case 'n': v.kind = Json::Null; return lit("null");
        default: return number(v);
        }
    }
    bool object(Json &v, int depth) {
        v.kind = Json::Object;
        ++p_;
        ws();
        if (p_ < end_ && *p_ == '}') {
            ++p_;
            return true;
        }
        for (;;) {
            ws();
            std::string key;
            if (p_ >= end_ || *p_ != '"' || !string(key)) return false;
            ws();
            if (p_ >= end_ || *p_ != ':') return false;
            ++p_;
            ws();
            v.obj.emplace_back(std::move(key), Json());
            if (!value(v.obj.back().second, depth + 1)) return false;
            ws();
            if (p_ < end_ && *p_ == ',') {
                ++p_;
                continue;
            }
            if (p_ < end_ && *p_ == '}') {
                ++p_;
                return true;
            }
            return false;
        }
    }
    bool array(Json &v, int depth) {
        v.kind = Json::Array;
        ++p_;
        ws();
        if (p_ < end_ && *p_ == ']') {
            ++p_;
            return true;
        }
        for (;;) {
            ws();
            v.arr.emplace_back();
            if (!value(v.arr.back(), depth + 1)) return false;
            ws();
            if (p_ < end_ && *p_ == ',') {
                ++p_;
                continue;
            }
            if (p_ < end_ && *p_ == ']') {
                ++p_;
                return true;
            }
            return false;
        }
    }
    static void utf8(std::string &s, unsigned cp) {
        if (cp < 0x80) {
            s.push_back((char)cp);
        } else if (cp < 0x800) {
            s.push_back((char)(0xC0 | (cp >> 6)));
            s.push_back((char)(0x80 | (cp & 0x3F)));
        } else if (cp < 0x10000) {
            s.push_back((char)(0xE0 | (cp >> 12)));
            s.push_back((char)(0x80 | ((cp >> 6) & 0x3F)));
            s.push_back((char)(0x80 | (cp & 0x3F)));
        } else {
            s.push_back((char)(0xF0 | (cp >> 18)));
            s.push_back((char)(0x80 | ((cp >> 12) & 0x3F)));
            s.push_back((char)(0x80 | ((cp >> 6) & 0x3F)));
            s.push_back((char)(0x80 | (cp & 0x3F)));
        }
    }
    bool hex4(unsigned &cp) {
        if (end_ - p_ < 4) return false;
        cp = 0;
        for (int i = 0; i < 4; ++i) {
            const char c = *p_++;
            cp <<= 4;
            if (c >= '0' && c <= '9') cp |= (unsigned)(c - '0');
            else if (c >= 'a' && c <= 'f') cp |= (unsigned)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') cp |= (unsigned)(c - 'A' + 10);
            else return false;
        }
        return true;
    }
    bool string(std::string &s) {
        ++p_;  // opening quote
        while (p_ < end_) {
            const char c = *p_++;
            if (c == '"') return true;
            if ((unsigned char)c < 0x20) {
                err_ = "control character in string";
                return false;
            }
            if (c != '\\') {
                s.push_back(c);
                continue;
            }
            if (p_ >= end_) return false;
            const char e = *p_++;
            switch (e) {
            case '"': s.push_back('"'); break;
            case '\\': s.push_back('\\'); break;
            case '/': s.push_back('/'); break;
            case 'b': s.push_back('\b'); break;
            case 'f': s.push_back('\f'); break;
            case 'n': s.push_back('\n'); break;
            case 'r': s.push_back('\r'); break;
            case 't': s.push_back('\t'); break;
            case 'u': {
                unsigned cp;
                if (!hex4(cp)) return false;
                if (cp >= 0xD800 && cp < 0xDC00) {  // surrogate pair
                    unsigned lo;
                    if (end_ - p_ < 6 || p_[0] != '\\' || p_[1] != 'u') return false;
                    p_ += 2;
                    if (!hex4(lo) || lo < 0xDC00 || lo >= 0xE000) return false;
                    cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                }
                utf8(s, cp);
                break;
            }
            default: return false;
            }
        }
        return false;
    }
    bool number(Json &v) {
        const char *s = p_;
        if (p_ < end_ && *p_ == '-') ++p_;
        if (p_ >= end_ || !(*p_ >= '0' && *p_ <= '9')) return false;
        if (*p_ == '0') ++p_;
        else while (p_ < end_ && *p_ >= '0' && *p_ <= '9') ++p_;
        bool integral = true;
        if (p_ < end_ && *p_ == '.') {
            integral = false;
            ++p_;
            if (p_ >= end_ || !(*p_ >= '0' && *p_ <= '9')) return false;
            while (p_ < end_ && *p_ >= '0' && *p_ <= '9') ++p_;
        }
        if (p_ < end_ && (*p_ == 'e' || *p_ == 'E')) {
            integral = false;
            ++p_;
            if (p_ < end_ && (*p_ == '+' || *p_ == '-')) ++p_;
            if (p_ >= end_ || !(*p_ >= '0' && *p_ <= '9')) return false;
            while (p_ < end_ && *p_ >= '0' && *p_ <= '9') ++p_;
        }
        v.kind = Json::Number;
        v.str.assign(s, p_);
        v.num = std::strtod(v.str.c_str(), nullptr);  // correctly rounded binary64
        v.integral = integral && v.str.size() <= 18;
        if (v.integral) v.ival = std::strtoll(v.str.c_str(), nullptr, 10);
        return true;
    }

    const char *p_;
    const char *end_;
    size_t total_ = (size_t)(end_ - p_);
    std::string err_;
};

// Canonical JSON of a document with no floating-point numbers, byte for byte
// what Python's json.dumps(v, sort_keys=True, separators=(",", ":")) writes
// (the reference's python_canonical_json, input.rs:44-110, for the harmonic
// contract, whose fields are integers, ASCII strings and booleans).
bool canonical(const Json &v, std::string &out) {
    switch (v.kind) {
    case Json::Null: out += "null"; return true;
    case Json::Bool: out += v.b ? "true" : "false"; return true;
    case Json::Number:
        if (!v.integral) return false;
        out += std::to_string(v.ival);
        return true;
    case Json::String: {
        out.push_back('"');
        for (size_t i = 0; i < v.str.size(); ++i) {
            const unsigned char c = (unsigned char)v.str[i];
            switch (c) {
            case '"': out += "\\\""; break;
            case '\\': out += "\\\\"; break;
            case '\b': out += "\\b"; break;
            case '\f': out += "\\f"; break;
            case '\n': out += "\\n"; break;
            case '\r': out += "\\r"; break;
            case '\t': out += "\\t"; break;
            default:
                if (c < 0x20 || c >= 0x80) return false;  // ASCII only
                out.push_back((char)c);
            }
        }
        out.push_back('"');
        return true;
    }
    case Json::Array:
        out.push_back('[');
        for (size_t i = 0; i < v.arr.size(); ++i) {
            if (i) out.push_back(',');
            if (!canonical(v.arr[i], out)) return false;
        }
        out.push_back(']');
        return true;
    case Json::Object: {
        std::vector<const std::pair<std::string, Json> *> kv;
        for (const auto &e : v.obj) kv.push_back(&e);
        std::sort(kv.begin(), kv.end(), [](auto *a, auto *b) { return a->first < b->first; });
        out.push_back('{');
        for (size_t i = 0; i < kv.size(); ++i) {
            if (i) out.push_back(',');
            Json k;
            k.kind = Json::String;
            k.str = kv[i]->first;
            if (!canonical(k, out)) return false;
            out.push_back(':');
            if (!canonical(kv[i]->second, out)) return false;
        }
        out.push_back('}');
        return true;
    }
    }
    return false;
}

// ---------------------------------------------------------------------------
// Field access with the reference's error behaviour
// ---------------------------------------------------------------------------
struct Reader {
    std::string err;
    bool ok = true;

    bool bad(const std::string &m) {
        if (ok) err = m;
        ok = false;
        return false;
    }
    const Json *field(const Json &o, const char *k, bool required = true) {
        const Json *v = o.get(k);
        if (!v && required) bad(std::string("missing field `") + k + "`");
        return v;
    }
    bool num(const Json &o, const char *k, double &out, bool required = true, double dflt = 0.0) {
        const Json *v = field(o, k, required);
        if (!v) {
            out = dflt;
            return ok;
        }
        if (v->kind != Json::Number) return bad(std::string("field `") + k + "` is not a number");
        out = v->num;
        return true;
    }
    bool index(const Json &o, const char *k, long long &out, bool required = true, long long dflt = 0) {
        const Json *v = field(o, k, required);
        if (!v) {
            out = dflt;
            return ok;
        }
        if (v->kind != Json::Number || !v->integral)
            return bad(std::string("field `") + k + "` is not an integer");
        out = v->ival;
        return true;
    }
    const Json *array(const Json &o, const char *k, bool required = true) {
        const Json *v = field(o, k, required);
        if (v && v->kind != Json::Array) {
            bad(std::string("field `") + k + "` is not an array");
            return nullptr;
        }
        return v;
    }
};

}  // namespace

// ---------------------------------------------------------------------------
// Owned topology
// ---------------------------------------------------------------------------
struct PrismTopoHandle {
    PrismMdTopologyV1 view{};
    PrismTopoInfoV1 info{};
    std::vector<float> pos, mass, charge, sigma, eps;
    std::vector<int32_t> bond_atoms, angle_atoms, dih_atoms, dih_n, ex_off, ex_atoms, exc_atoms;
    std::vector<float> bond_r0, bond_k, angle_t0, angle_k, dih_phase, dih_k, exc_qq, exc_sig, exc_eps;
    std::vector<uint8_t> dih_improper;
    std::vector<int32_t> cl_central, cl_h, cl_nh;
    std::vector<float> cl_d, cl_imc, cl_imh;

    void bind() {
        PrismMdTopologyV1 &t = view;
        t.n_atoms = (int32_t)mass.size();
        t.positions = pos.data();
        t.masses = mass.data();
        t.charges = charge.data();
        t.sigma = sigma.data();
        t.epsilon = eps.data();
        t.n_bonds = (int32_t)bond_r0.size();
        t.bond_atoms = bond_atoms.data();
        t.bond_r0 = bond_r0.data();
        t.bond_k = bond_k.data();
        t.n_angles = (int32_t)angle_t0.size();
        t.angle_atoms = angle_atoms.data();
        t.angle_theta0 = angle_t0.data();
        t.angle_k = angle_k.data();
        t.n_dihedrals = (int32_t)dih_k.size();
        t.dihedral_atoms = dih_atoms.data();
        t.dihedral_periodicity = dih_n.data();
        t.dihedral_phase = dih_phase.data();
        t.dihedral_k = dih_k.data();
        t.dihedral_improper = dih_improper.data();
        t.exclusion_offsets = ex_off.data();
        t.exclusion_atoms = ex_atoms.data();
        t.n_exceptions = (int32_t)exc_qq.size();
        t.exception_atoms = exc_atoms.data();
        t.exception_charge_product = exc_qq.data();
        t.exception_sigma = exc_sig.data();
        t.exception_epsilon = exc_eps.data();
        t.n_clusters = (int32_t)cl_central.size();
        t.cluster_central = cl_central.data();
        t.cluster_hydrogens = cl_h.data();
        t.cluster_bond_lengths = cl_d.data();
        t.cluster_n_hydrogens = cl_nh.data();
        t.cluster_inv_mass_central = cl_imc.data();
        t.cluster_inv_mass_h = cl_imh.data();
    }
};

namespace {

bool string_is(const Json *v, const char *s) { return v && v->kind == Json::String && v->str == s; }
bool int_is(const Json *v, long long x) { return v && v->kind == Json::Number && v->integral && v->ival == x; }
bool bool_is(const Json *v, bool x) { return v && v->kind == Json::Bool && v->b == x; }

// f32 narrowing rule of the exact exceptions (input.rs:2148-2170).
bool f32_safe(double value) {
    const float narrowed = (float)value;
    if (!std::isfinite(narrowed) || (value != 0.0 && narrowed == 0.0f) || (value == 0.0 && std::signbit(value)))
        return false;
    return std::fabs((double)narrowed - value) <= std::max(1.0e-12, std::fabs(value) * 1.0e-6);
}

struct ExceptionRow {
    long long i, j;
    double qq, sigma, eps;
};

std::string exception_digest(std::vector<ExceptionRow> rows) {
    std::stable_sort(rows.begin(), rows.end(),
                     [](const ExceptionRow &a, const ExceptionRow &b) { return a.i != b.i ? a.i < b.i : a.j < b.j; });
    Sha256 h;
    h.update("PRISMOS_NBEX_ACTIVE_V1", 23);  // including the terminating NUL
    h.update_le<uint64_t>(rows.size());
    for (const auto &r : rows) {
        h.update_le<uint64_t>((uint64_t)r.i);
        h.update_le<uint64_t>((uint64_t)r.j);
        h.update_le<double>(r.qq);
        h.update_le<double>(r.sigma);
        h.update_le<double>(r.eps);
    }
    return h.hex();
}

std::string strict_digest(const std::set<std::pair<long long, long long>> &pairs) {
    Sha256 h;
    h.update("PRISMOS_NBEX_STRICT_V1", 23);
    h.update_le<uint64_t>(pairs.size());
    for (const auto &p : pairs) {
        h.update_le<uint64_t>((uint64_t)p.first);
        h.update_le<uint64_t>((uint64_t)p.second);
    }
    return h.hex();
}

int convention_of(const Json *v) {
    if (string_is(v, "AMBER_NO_HALF_ENERGY")) return PRISM_TOPO_CONVENTION_AMBER_NO_HALF;
    if (string_is(v, "OPENMM_FULL_K_HALF_ENERGY")) return PRISM_TOPO_CONVENTION_OPENMM_HALF;
    return -1;
}

// Energy prefactor p of U = p k delta^2 (HarmonicEnergyConvention::energy_prefactor).
double prefactor(int conv) { return conv == PRISM_TOPO_CONVENTION_OPENMM_HALF ? 0.5 : 1.0; }

int load(const Json &doc, const PrismTopoLoadOptionsV1 &opt, std::unique_ptr<PrismTopoHandle> &out) {
    if (doc.kind != Json::Object) return fail(PRISM_MD_E_INVALID, "topology JSON is not an object");
    Reader R;
    auto h = std::make_unique<PrismTopoHandle>();
    long long n = 0, n_res = 0, n_chains = 0;
    R.index(doc, "n_atoms", n);
    R.index(doc, "n_residues", n_res, false, 0);
    R.index(doc, "n_chains", n_chains, false, 0);
    if (!R.ok) return fail(PRISM_MD_E_INVALID, R.err);
    if (n <= 0 || n > (1ll << 30)) return fail(PRISM_MD_E_INVALID, "n_atoms out of range");

    auto floats = [&](const char *k, std::vector<float> &dst, size_t expect) -> bool {
        const Json *a = R.array(doc, k);
        if (!a) return false;
        if (a->arr.size() != expect)
            return R.bad(std::string(k) + " length " + std::to_string(a->arr.size()) + " != expected " +
                         std::to_string(expect));
        dst.resize(expect);
        for (size_t i = 0; i < expect; ++i) {
            if (a->arr[i].kind != Json::Number) return R.bad(std::string(k) + " holds a non-number");
            dst[i] = (float)a->arr[i].num;
        }
        return true;
    };
    // positions, charges, masses (validate, input.rs:2002-2052)
    if (!floats("positions", h->pos, 3 * (size_t)n) || !floats("charges", h->charge, (size_t)n) ||
        !floats("masses", h->mass, (size_t)n))
        return fail(PRISM_MD_E_INVALID, R.err);
    for (size_t i = 0; i < h->pos.size(); ++i)
        if (!std::isfinite(h->pos[i])) return fail(PRISM_MD_E_INVALID, "position value " + std::to_string(i) + " is non-finite");
    for (long long i = 0; i < n; ++i) {
        if (!std::isfinite(h->charge[i])) return fail(PRISM_MD_E_INVALID, "charge value " + std::to_string(i) + " is non-finite");
        if (!std::isfinite(h->mass[i]) || !(h->mass[i] > 0.0f))
            return fail(PRISM_MD_E_INVALID, "mass value " + std::to_string(i) + " must be finite and positive");
    }
    // LJ parameters
    const Json *lj = R.array(doc, "lj_params");
    if (!lj) return fail(PRISM_MD_E_INVALID, R.err);
    if ((long long)lj->arr.size() != n)
        return fail(PRISM_MD_E_INVALID, "lj_params length " + std::to_string(lj->arr.size()) + " != expected " + std::to_string(n));
    h->sigma.resize(n);
    h->eps.resize(n);
    for (long long i = 0; i < n; ++i) {
        double s, e;
        R.num(lj->arr[i], "sigma", s);
        R.num(lj->arr[i], "epsilon", e);
        if (!R.ok) return fail(PRISM_MD_E_INVALID, "lj_params[" + std::to_string(i) + "]: " + R.err);
        if (!std::isfinite(s) || !std::isfinite(e) || s < 0.0 || e < 0.0)
            return fail(PRISM_MD_E_INVALID, "invalid LJ parameters for atom " + std::to_string(i));
        h->sigma[i] = (float)s;
        h->eps[i] = (float)e;
    }

    // bonds and angles, raw binary64 (input.rs:2068-2094)
    const Json *bonds = R.array(doc, "bonds");
    const Json *angles = R.array(doc, "angles");
    if (!bonds || !angles) return fail(PRISM_MD_E_INVALID, R.err);
    std::vector<double> bond_k_raw, angle_k_raw;
    for (size_t b = 0; b < bonds->arr.size(); ++b) {
        long long i, j;
        double r0, k;
        R.index(bonds->arr[b], "i", i);
        R.index(bonds->arr[b], "j", j);
        R.num(bonds->arr[b], "r0", r0);
        R.num(bonds->arr[b], "k", k);
        if (!R.ok) return fail(PRISM_MD_E_INVALID, "bonds[" + std::to_string(b) + "]: " + R.err);
        if (i < 0 || j < 0 || i >= n || j >= n || i == j || !std::isfinite(r0) || !std::isfinite(k) || r0 <= 0.0 || k < 0.0)
            return fail(PRISM_MD_E_INVALID, "invalid bond term at index " + std::to_string(b));
        h->bond_atoms.push_back((int32_t)i);
        h->bond_atoms.push_back((int32_t)j);
        h->bond_r0.push_back((float)r0);
        bond_k_raw.push_back(k);
    }
    for (size_t a = 0; a < angles->arr.size(); ++a) {
        long long i, j, k;
        double t0, fk;
        R.index(angles->arr[a], "i", i);
        R.index(angles->arr[a], "j", j);
        R.index(angles->arr[a], "k_idx", k);
        R.num(angles->arr[a], "theta0", t0);
        R.num(angles->arr[a], "force_k", fk);
        if (!R.ok) return fail(PRISM_MD_E_INVALID, "angles[" + std::to_string(a) + "]: " + R.err);
        if (i < 0 || j < 0 || k < 0 || i >= n || j >= n || k >= n || i == j || j == k || i == k || !std::isfinite(t0) ||
            !std::isfinite(fk) || t0 <= 0.0 || t0 > M_PI || fk < 0.0)
            return fail(PRISM_MD_E_INVALID, "invalid angle term at index " + std::to_string(a));
        for (long long x : {i, j, k}) h->angle_atoms.push_back((int32_t)x);
        h->angle_t0.push_back((float)t0);
        angle_k_raw.push_back(fk);
    }

    // harmonic lowering (input.rs:300-323, :467-538)
    int conv = opt.harmonic_convention;
    const Json *hc = doc.get("harmonic_parameter_contract");
    if (hc && hc->kind == Json::Null) hc = nullptr;
    if (hc) {
        if (hc->kind != Json::Object) return fail(PRISM_MD_E_INVALID, "harmonic_parameter_contract is not an object");
        static const char *known[] = {"schema_kind", "schema_version", "source_provider", "bond_convention",
                                      "angle_convention", "bond_coefficient_unit", "angle_coefficient_unit",
                                      "bond_equilibrium_unit", "angle_equilibrium_unit",
                                      "bond_uses_periodic_boundary_conditions",
                                      "angle_uses_periodic_boundary_conditions", "bond_count", "angle_count",
                                      "contract_sha256"};
        for (const auto &kv : hc->obj)  // deny_unknown_fields
            if (std::find_if(std::begin(known), std::end(known), [&](const char *s) { return kv.first == s; }) ==
                std::end(known))
                return fail(PRISM_MD_E_INVALID, "harmonic_parameter_contract: unknown field `" + kv.first + "`");
        const Json *prov = hc->get("source_provider");
        const int bc = convention_of(hc->get("bond_convention")), ac = convention_of(hc->get("angle_convention"));
        bool prov_ok = prov && prov->kind == Json::String;
        if (prov_ok) {
            prov_ok = false;
            for (char c : prov->str)
                if (!std::isspace((unsigned char)c)) prov_ok = true;
        }
        if (!string_is(hc->get("schema_kind"), "prismos_harmonic_parameter_contract") || !int_is(hc->get("schema_version"), 1) ||
            !prov_ok || bc < 0 || ac < 0 || !string_is(hc->get("bond_coefficient_unit"), "kcal/mol/angstrom^2") ||
            !string_is(hc->get("angle_coefficient_unit"), "kcal/mol/radian^2") ||
            !string_is(hc->get("bond_equilibrium_unit"), "angstrom") ||
            !string_is(hc->get("angle_equilibrium_unit"), "radian") ||
            !bool_is(hc->get("bond_uses_periodic_boundary_conditions"), false) ||
            !bool_is(hc->get("angle_uses_periodic_boundary_conditions"), false) ||
            !int_is(hc->get("bond_count"), (long long)bond_k_raw.size()) ||
            !int_is(hc->get("angle_count"), (long long)angle_k_raw.size()))
            return fail(PRISM_MD_E_INVALID,
                        "harmonic source contract has unsupported schema, units, provider or term counts");
        Json body = *hc;
        body.obj.erase(std::remove_if(body.obj.begin(), body.obj.end(),
                                      [](const std::pair<std::string, Json> &kv) { return kv.first == "contract_sha256"; }),
                       body.obj.end());
        std::string text;
        if (!canonical(body, text)) return fail(PRISM_MD_E_INVALID, "harmonic contract is not canonicalisable");
        Sha256 sh;
        sh.update(text);
        if (!string_is(hc->get("contract_sha256"), sh.hex().c_str()))
            return fail(PRISM_MD_E_INVALID, "harmonic source contract canonical digest mismatch");
        if (bc != ac) return fail(PRISM_MD_E_UNSUPPORTED, "bond and angle conventions differ");
        if (conv != PRISM_TOPO_CONVENTION_FROM_CONTRACT && conv != bc)
            return fail(PRISM_MD_E_INVALID, "requested harmonic convention disagrees with the file's contract");
        conv = bc;
    } else if (conv == PRISM_TOPO_CONVENTION_FROM_CONTRACT) {
        return fail(PRISM_MD_E_INVALID,
                    "harmonic source convention is absent: refuse production lowering (pass an explicit convention)");
    } else if (conv != PRISM_TOPO_CONVENTION_AMBER_NO_HALF && conv != PRISM_TOPO_CONVENTION_OPENMM_HALF) {
        return fail(PRISM_MD_E_INVALID, "unknown harmonic convention option");
    }
    // lower_harmonic_parameters: r0/theta0 representable, k * (p_src / p_target) as f32
    const double scale = prefactor(conv) / prefactor(PRISM_TOPO_CONVENTION_AMBER_NO_HALF);
    for (size_t b = 0; b < bond_k_raw.size(); ++b) {
        const double mapped = bond_k_raw[b] * scale;
        const float dev = (float)mapped;
        if (!(h->bond_r0[b] > 0.0f) || !std::isfinite(h->bond_r0[b]) || !std::isfinite(dev) ||
            (bond_k_raw[b] > 0.0 && dev == 0.0f))
            return fail(PRISM_MD_E_INVALID, "bond coefficient " + std::to_string(b) + " is not representable in float32");
        h->bond_k.push_back(dev);
    }
    for (size_t a = 0; a < angle_k_raw.size(); ++a) {
        const double mapped = angle_k_raw[a] * scale;
        const float dev = (float)mapped;
        if (!(h->angle_t0[a] > 0.0f) || !std::isfinite(dev) || (angle_k_raw[a] > 0.0 && dev == 0.0f))
            return fail(PRISM_MD_E_INVALID, "angle coefficient " + std::to_string(a) + " is not representable in float32");
        h->angle_k.push_back(dev);
    }

    // dihedrals (input.rs:2095-2106); periodicity, phase, force_k default 0
    const Json *dih = R.array(doc, "dihedrals");
    if (!dih) return fail(PRISM_MD_E_INVALID, R.err);
    for (size_t d = 0; d < dih->arr.size(); ++d) {
        long long a[4], per;
        double phase, fk;
        R.index(dih->arr[d], "i", a[0]);
        R.index(dih->arr[d], "j", a[1]);
        R.index(dih->arr[d], "k_idx", a[2]);
        R.index(dih->arr[d], "l", a[3]);
        R.index(dih->arr[d], "periodicity", per, false, 0);
        R.num(dih->arr[d], "phase", phase, false, 0.0);
        R.num(dih->arr[d], "force_k", fk, false, 0.0);
        const Json *imp = dih->arr[d].get("improper");
        if (!R.ok) return fail(PRISM_MD_E_INVALID, "dihedrals[" + std::to_string(d) + "]: " + R.err);
        bool bad = per <= 0 || !std::isfinite(phase) || !std::isfinite(fk) || (imp && imp->kind != Json::Bool);
        for (long long x : a) bad |= x < 0 || x >= n;
        if (bad) return fail(PRISM_MD_E_INVALID, "invalid dihedral term at index " + std::to_string(d));
        for (long long x : a) h->dih_atoms.push_back((int32_t)x);
        h->dih_n.push_back((int32_t)per);
        h->dih_phase.push_back((float)phase);
        h->dih_k.push_back((float)fk);
        h->dih_improper.push_back(imp && imp->b ? 1 : 0);
    }

    // strict exclusions: per atom, no self, no duplicate, symmetric (input.rs:2107-2122)
    const Json *ex = R.array(doc, "exclusions");
    if (!ex) return fail(PRISM_MD_E_INVALID, R.err);
    if ((long long)ex->arr.size() != n)
        return fail(PRISM_MD_E_INVALID, "exclusions length " + std::to_string(ex->arr.size()) + " != expected " + std::to_string(n));
    std::vector<std::vector<int32_t>> lists((size_t)n);
    for (long long i = 0; i < n; ++i) {
        const Json &row = ex->arr[i];
        if (row.kind != Json::Array) return fail(PRISM_MD_E_INVALID, "exclusions[" + std::to_string(i) + "] is not an array");
        for (const Json &p : row.arr) {
            if (p.kind != Json::Number || !p.integral) return fail(PRISM_MD_E_INVALID, "exclusion partner is not an integer");
            lists[i].push_back((int32_t)std::max<long long>(-1, std::min<long long>(p.ival, n)));
            const long long q = p.ival;
            if (q < 0 || q >= n || q == i)
                return fail(PRISM_MD_E_INVALID, "invalid, self, or duplicate strict exclusion (" + std::to_string(i) + ", " +
                                                    std::to_string(q) + ")");
        }
        std::vector<int32_t> s = lists[i];
        std::sort(s.begin(), s.end());
        if (std::adjacent_find(s.begin(), s.end()) != s.end())
            return fail(PRISM_MD_E_INVALID, "invalid, self, or duplicate strict exclusion (" + std::to_string(i) + ", " +
                                                std::to_string(*std::adjacent_find(s.begin(), s.end())) + ")");
    }
    std::vector<std::vector<int32_t>> sorted(lists);
    for (auto &s : sorted) std::sort(s.begin(), s.end());
    std::set<std::pair<long long, long long>> strict;
    for (long long i = 0; i < n; ++i)
        for (int32_t q : lists[i]) {
            if (!std::binary_search(sorted[q].begin(), sorted[q].end(), (int32_t)i))
                return fail(PRISM_MD_E_INVALID, "asymmetric strict exclusion (" + std::to_string(i) + ", " + std::to_string(q) + ")");
            strict.insert({std::min<long long>(i, q), std::max<long long>(i, q)});
        }
    h->ex_off.assign(1, 0);
    for (long long i = 0; i < n; ++i) {
        for (int32_t q : lists[i]) h->ex_atoms.push_back(q);
        h->ex_off.push_back((int32_t)h->ex_atoms.size());
    }

    // exact exceptions and their contract (input.rs:2123-2290)
    const Json *exc = R.array(doc, "nonbonded_exceptions", false);
    if (!R.ok) return fail(PRISM_MD_E_INVALID, R.err);
    std::vector<ExceptionRow> rows;
    std::set<std::pair<long long, long long>> seen;
    if (exc) {
        for (size_t e = 0; e < exc->arr.size(); ++e) {
            ExceptionRow r{};
            R.index(exc->arr[e], "i", r.i);
            R.index(exc->arr[e], "j", r.j);
            R.num(exc->arr[e], "charge_product_e2", r.qq);
            R.num(exc->arr[e], "sigma_A", r.sigma);
            R.num(exc->arr[e], "epsilon_kcal_mol", r.eps);
            if (!R.ok) return fail(PRISM_MD_E_INVALID, "nonbonded_exceptions[" + std::to_string(e) + "]: " + R.err);
            const std::string pr = "(" + std::to_string(r.i) + ", " + std::to_string(r.j) + ")";
            if (r.i < 0 || r.j < 0 || r.i >= n || r.j >= n)
                return fail(PRISM_MD_E_INVALID, "nonbonded exception index out of range: " + pr);
            if (r.i >= r.j) return fail(PRISM_MD_E_INVALID, "nonbonded exception pair must be canonical i < j: " + pr);
            if (!std::isfinite(r.qq) || !std::isfinite(r.sigma) || !std::isfinite(r.eps) || r.sigma < 0.0 || r.eps < 0.0)
                return fail(PRISM_MD_E_INVALID, "invalid nonbonded exception parameters for pair " + pr);
            if (!f32_safe(r.qq) || !f32_safe(r.sigma) || !f32_safe(r.eps))
                return fail(PRISM_MD_E_INVALID, "nonbonded exception parameter is not safely representable as f32 for pair " + pr);
            if (!seen.insert({r.i, r.j}).second) return fail(PRISM_MD_E_INVALID, "duplicate nonbonded exception pair: " + pr);
            if (strict.count({r.i, r.j})) return fail(PRISM_MD_E_INVALID, "nonbonded exception overlaps strict exclusion: " + pr);
            rows.push_back(r);
        }
    }
    const Json *xc = doc.get("nonbonded_exception_contract");
    if (xc && xc->kind == Json::Null) xc = nullptr;
    if (rows.empty() && xc)
        return fail(PRISM_MD_E_INVALID, "nonbonded exception contract is present but the exact pair vector is empty");
    if (!rows.empty() && !xc)
        return fail(PRISM_MD_E_INVALID, "exact nonbonded exceptions require a versioned completeness/digest contract");
    if (xc) {
        if (xc->kind != Json::Object) return fail(PRISM_MD_E_INVALID, "nonbonded_exception_contract is not an object");
        if (!string_is(xc->get("schema_kind"), "prismos_exact_nonbonded_exceptions_v2") || !int_is(xc->get("schema_version"), 2))
            return fail(PRISM_MD_E_INVALID, "unsupported nonbonded exception contract");
        if (!string_is(xc->get("semantics"), "OPENMM_NONBONDEDFORCE_ABSOLUTE_EXCEPTION_PARAMETERS_V1") ||
            !string_is(xc->get("dynamic_charge_scaling_semantics"),
                       "EXCITED_STATE_MULTIPLIES_GROUND_STATE_CHARGE_PRODUCT_BY_PER_ATOM_SCALE_PRODUCT_V1"))
            return fail(PRISM_MD_E_INVALID, "unsupported exact nonbonded exception semantics");
        if (!string_is(xc->get("pair_digest_algorithm"), "SHA256_PRISMOS_NBEX_ACTIVE_V1_LE_U64_U64_F64_F64_F64") ||
            !string_is(xc->get("strict_exclusion_digest_algorithm"), "SHA256_PRISMOS_NBEX_STRICT_V1_LE_U64_U64"))
            return fail(PRISM_MD_E_INVALID, "unsupported nonbonded exception digest algorithms");
        if (!bool_is(xc->get("complete"), true) || !bool_is(xc->get("fallback_used"), false) ||
            !bool_is(xc->get("hardcoded_uniform_scaling_allowed"), false))
            return fail(PRISM_MD_E_INVALID, "exact nonbonded exception contract is incomplete or records fallback");
        if (!int_is(xc->get("exact_scaled_pair_count"), (long long)rows.size()))
            return fail(PRISM_MD_E_INVALID, "nonbonded exception count mismatch");
        if (!int_is(xc->get("strict_exclusion_pair_count"), (long long)strict.size()))
            return fail(PRISM_MD_E_INVALID, "strict exclusion count mismatch");
        if (!int_is(xc->get("total_openmm_exception_count"), (long long)(rows.size() + strict.size())))
            return fail(PRISM_MD_E_INVALID, "OpenMM exception completeness mismatch");
        if (!string_is(xc->get("pair_digest_sha256"), exception_digest(rows).c_str()))
            return fail(PRISM_MD_E_INVALID, "nonbonded exception digest mismatch");
        if (!string_is(xc->get("strict_exclusion_digest_sha256"), strict_digest(strict).c_str()))
            return fail(PRISM_MD_E_INVALID, "strict exclusion digest mismatch");
        const Json *prm = doc.get("generated_prmtop_sha256");
        const Json *src = xc->get("source_prmtop_sha256");
        bool hex_ok = prm && prm->kind == Json::String && prm->str.size() == 64;
        if (hex_ok)
            for (char c : prm->str) hex_ok &= std::isxdigit((unsigned char)c) != 0;
        if (!hex_ok || !src || src->kind != Json::String || src->str != prm->str)
            return fail(PRISM_MD_E_INVALID, "exact exception source prmtop hash is absent, malformed, or mismatched");
    }
    for (const auto &r : rows) {
        h->exc_atoms.push_back((int32_t)r.i);
        h->exc_atoms.push_back((int32_t)r.j);
        h->exc_qq.push_back((float)r.qq);
        h->exc_sig.push_back((float)r.sigma);
        h->exc_eps.push_back((float)r.eps);
    }

    // hydrogen clusters (HCluster, input.rs:625-642)
    const Json *hcl = R.array(doc, "h_clusters", false);
    if (!R.ok) return fail(PRISM_MD_E_INVALID, R.err);
    struct Cl {
        long long c;
        long long hs[3];
        double d[3];
        long long nh;
        double imc, imh;
    };
    std::vector<Cl> cls;
    if (hcl) {
        for (size_t c = 0; c < hcl->arr.size(); ++c) {
            const Json &o = hcl->arr[c];
            Cl k{};
            long long type;
            R.index(o, "type", type);
            R.index(o, "central_atom", k.c);
            R.index(o, "n_hydrogens", k.nh);
            R.num(o, "inv_mass_central", k.imc);
            R.num(o, "inv_mass_h", k.imh);
            const Json *ha = R.array(o, "hydrogen_atoms"), *bl = R.array(o, "bond_lengths");
            if (!R.ok) return fail(PRISM_MD_E_INVALID, "h_clusters[" + std::to_string(c) + "]: " + R.err);
            const std::string where = "h_clusters[" + std::to_string(c) + "]";
            if (k.nh < 1 || k.nh > 3 || k.c < 0 || k.c >= n || (long long)ha->arr.size() < k.nh ||
                (long long)bl->arr.size() < k.nh || !(k.imc > 0.0) || !(k.imh > 0.0) || !std::isfinite(k.imc) ||
                !std::isfinite(k.imh))
                return fail(PRISM_MD_E_INVALID, "invalid hydrogen cluster " + where);
            for (int s = 0; s < 3; ++s) {
                k.hs[s] = -1;
                k.d[s] = 0.0;
            }
            for (long long s = 0; s < k.nh; ++s) {
                const Json &a = ha->arr[s], &d = bl->arr[s];
                if (a.kind != Json::Number || !a.integral || a.ival < 0 || a.ival >= n || d.kind != Json::Number ||
                    !(d.num > 0.0) || !std::isfinite(d.num))
                    return fail(PRISM_MD_E_INVALID, "invalid hydrogen cluster " + where);
                k.hs[s] = a.ival;
                k.d[s] = d.num;
            }
            cls.push_back(k);
        }
    }
    // HMR (apply_hmr_with_historical_repeat, input.rs:2494-2565): only a
    // recorded `hmr_applied: true` suppresses the repartition.
    const Json *hm = doc.get("hmr_applied");
    h->info.hmr_in_file = (!hm || hm->kind == Json::Null) ? -1 : (hm->kind == Json::Bool ? (hm->b ? 1 : 0) : -2);
    if (h->info.hmr_in_file == -2) return fail(PRISM_MD_E_INVALID, "hmr_applied is not a boolean");
    double rec_fs = 0.0;
    R.num(doc, "recommended_timestep_fs", rec_fs, false, 0.0);
    if (!R.ok) {
        const Json *r = doc.get("recommended_timestep_fs");
        if (!(r && r->kind == Json::Null)) return fail(PRISM_MD_E_INVALID, R.err);
        rec_fs = 0.0;
    }
    if (opt.apply_hmr && h->info.hmr_in_file != 1) {
        if (!(opt.hmr_factor > 0.0) || !std::isfinite(opt.hmr_factor)) return fail(PRISM_MD_E_INVALID, "hmr_factor");
        for (auto &k : cls) {
            const double m_h = 1.0 / k.imh, m_c = 1.0 / k.imc;
            const double m_h_new = m_h * opt.hmr_factor;
            const double m_c_new = m_c - (double)k.nh * (m_h_new - m_h);
            if (m_c_new < 1.0) {
                ++h->info.hmr_skipped;
                continue;
            }
            k.imh = 1.0 / m_h_new;
            k.imc = 1.0 / m_c_new;
            h->mass[k.c] = (float)m_c_new;
            for (long long s = 0; s < k.nh; ++s) h->mass[k.hs[s]] = (float)m_h_new;
            ++h->info.hmr_applied_here;
        }
        rec_fs = 4.0;
    }
    for (const auto &k : cls) {
        h->cl_central.push_back((int32_t)k.c);
        for (int s = 0; s < 3; ++s) {
            h->cl_h.push_back((int32_t)k.hs[s]);
            h->cl_d.push_back((float)k.d[s]);
        }
        h->cl_nh.push_back((int32_t)k.nh);
        h->cl_imc.push_back((float)k.imc);
        h->cl_imh.push_back((float)k.imh);
    }

    h->info.n_atoms = (int32_t)n;
    h->info.n_residues = (int32_t)n_res;
    h->info.n_chains = (int32_t)n_chains;
    h->info.harmonic_convention = conv;
    h->info.has_exception_contract = xc ? 1 : 0;
    h->info.recommended_timestep_fs = rec_fs;
    h->bind();
    out = std::move(h);
    return PRISM_MD_OK;
}

// ---------------------------------------------------------------------------
// Writer
// ---------------------------------------------------------------------------
void put_double(std::ostringstream &os, double v) {
    char buf[40];
    std::snprintf(buf, sizeof buf, "%.17g", v);  // round-trips binary64
    os << buf;
    if (!std::strpbrk(buf, ".eEn")) os << ".0";
}

}  // namespace

extern "C" {

int prism_topo_default_options(PrismTopoLoadOptionsV1 *opt) {
    if (!opt) return PRISM_MD_E_INVALID;
    std::memset(opt, 0, sizeof(*opt));
    opt->harmonic_convention = PRISM_TOPO_CONVENTION_FROM_CONTRACT;
    opt->apply_hmr = 1;
    opt->hmr_factor = 3.0;
    return PRISM_MD_OK;
}

int prism_topo_parse_json(const char *text, size_t len, const PrismTopoLoadOptionsV1 *opt, PrismTopoHandle **out) {
    if (!text || !out) return fail(PRISM_MD_E_INVALID, "null argument");
    *out = nullptr;
    PrismTopoLoadOptionsV1 o;
    if (opt) o = *opt;
    else prism_topo_default_options(&o);
    Json doc;
    std::string err;
    Parser p(text, len);
    if (!p.parse(doc, err)) return fail(PRISM_MD_E_INVALID, "Failed to parse topology JSON: " + err);
    std::unique_ptr<PrismTopoHandle> h;
    try {
        if (int rc = load(doc, o, h)) return rc;
    } catch (const std::bad_alloc &) {
        return fail(PRISM_MD_E_NOMEM, "out of host memory");
    }
    *out = h.release();
    g_error = "no error";
    return PRISM_MD_OK;
}

int prism_topo_load_json(const char *path, const PrismTopoLoadOptionsV1 *opt, PrismTopoHandle **out) {
    if (!path || !out) return fail(PRISM_MD_E_INVALID, "null argument");
    *out = nullptr;
    std::ifstream f(path, std::ios::binary);
    if (!f) return fail(PRISM_MD_E_INVALID, std::string("Failed to read topology file: ") + path);
    std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    return prism_topo_parse_json(text.data(), text.size(), opt, out);
}

const PrismMdTopologyV1 *prism_topo_view(const PrismTopoHandle *t) { return t ? &t->view : nullptr; }

int prism_topo_info(const PrismTopoHandle *t, PrismTopoInfoV1 *info) {
    if (!t || !info) return PRISM_MD_E_INVALID;
    *info = t->info;
    return PRISM_MD_OK;
}

int prism_topo_destroy(PrismTopoHandle *t) {
    delete t;
    return PRISM_MD_OK;
}

const char *prism_topo_last_error(void) { return g_error.c_str(); }

int prism_topo_sha256_hex(const void *data, size_t len, char out_hex[65]) {
    if ((!data && len) || !out_hex) return PRISM_MD_E_INVALID;
    Sha256 h;
    h.update(data, len);
    const std::string s = h.hex();
    std::memcpy(out_hex, s.c_str(), 65);
    return PRISM_MD_OK;
}

int prism_topo_write_json(const PrismMdTopologyV1 *t, int32_t hmr_applied, const char *path) {
    if (!t || !path || t->n_atoms < 1 || !t->exclusion_offsets) return fail(PRISM_MD_E_INVALID, "null argument");
    const int n = t->n_atoms;
    std::ostringstream os;
    os << "{\"schema_kind\":\"prismos_openmm_topology\",\"source_pdb\":\"synthetic\",\"n_atoms\":" << n
       << ",\"n_residues\":1,\"n_chains\":1";
    auto farr = [&](const char *k, const float *a, size_t m) {
        os << ",\"" << k << "\":[";
        for (size_t i = 0; i < m; ++i) {
            if (i) os << ',';
            put_double(os, (double)a[i]);
        }
        os << ']';
    };
    auto sarr = [&](const char *k, const char *v) {
        os << ",\"" << k << "\":[";
        for (int i = 0; i < n; ++i) os << (i ? "," : "") << '"' << v << '"';
        os << ']';
    };
    farr("positions", t->positions, 3 * (size_t)n);
    sarr("elements", "X");
    sarr("atom_names", "X");
    sarr("residue_names", "UNK");
    os << ",\"residue_ids\":[";
    for (int i = 0; i < n; ++i) os << (i ? ",0" : "0");
    os << ']';
    sarr("chain_ids", "A");
    farr("charges", t->charges, (size_t)n);
    farr("masses", t->masses, (size_t)n);
    os << ",\"ca_indices\":[],\"bonds\":[";
    for (int b = 0; b < t->n_bonds; ++b) {
        os << (b ? "," : "") << "{\"i\":" << t->bond_atoms[2 * b] << ",\"j\":" << t->bond_atoms[2 * b + 1] << ",\"r0\":";
        put_double(os, t->bond_r0[b]);
        os << ",\"k\":";
        put_double(os, t->bond_k[b]);
        os << '}';
    }
    os << "],\"angles\":[";
    for (int a = 0; a < t->n_angles; ++a) {
        os << (a ? "," : "") << "{\"i\":" << t->angle_atoms[3 * a] << ",\"j\":" << t->angle_atoms[3 * a + 1]
           << ",\"k_idx\":" << t->angle_atoms[3 * a + 2] << ",\"theta0\":";
        put_double(os, t->angle_theta0[a]);
        os << ",\"force_k\":";
        put_double(os, t->angle_k[a]);
        os << '}';
    }
    os << ']';
    // harmonic contract with its canonical digest
    Json hc;
    hc.kind = Json::Object;
    auto add_s = [&](const char *k, const char *v) {
        Json x;
        x.kind = Json::String;
        x.str = v;
        hc.obj.emplace_back(k, x);
    };
    auto add_i = [&](const char *k, long long v) {
        Json x;
        x.kind = Json::Number;
        x.integral = true;
        x.ival = v;
        hc.obj.emplace_back(k, x);
    };
    auto add_b = [&](const char *k, bool v) {
        Json x;
        x.kind = Json::Bool;
        x.b = v;
        hc.obj.emplace_back(k, x);
    };
    add_s("schema_kind", "prismos_harmonic_parameter_contract");
    add_i("schema_version", 1);
    add_s("source_provider", "prism_md synthetic generator");
    add_s("bond_convention", "AMBER_NO_HALF_ENERGY");
    add_s("angle_convention", "AMBER_NO_HALF_ENERGY");
    add_s("bond_coefficient_unit", "kcal/mol/angstrom^2");
    add_s("angle_coefficient_unit", "kcal/mol/radian^2");
    add_s("bond_equilibrium_unit", "angstrom");
    add_s("angle_equilibrium_unit", "radian");
    add_b("bond_uses_periodic_boundary_conditions", false);
    add_b("angle_uses_periodic_boundary_conditions", false);
    add_i("bond_count", t->n_bonds);
    add_i("angle_count", t->n_angles);
    std::string canon;
    canonical(hc, canon);
    Sha256 sh;
    sh.update(canon);
    const std::string hc_digest = sh.hex();
    os << ",\"harmonic_parameter_contract\":" << canon.substr(0, canon.size() - 1) << ",\"contract_sha256\":\""
       << hc_digest << "\"}";
    os << ",\"dihedrals\":[";
    for (int d = 0; d < t->n_dihedrals; ++d) {
        os << (d ? "," : "") << "{\"i\":" << t->dihedral_atoms[4 * d] << ",\"j\":" << t->dihedral_atoms[4 * d + 1]
           << ",\"k_idx\":" << t->dihedral_atoms[4 * d + 2] << ",\"l\":" << t->dihedral_atoms[4 * d + 3]
           << ",\"periodicity\":" << t->dihedral_periodicity[d] << ",\"phase\":";
        put_double(os, t->dihedral_phase[d]);
        os << ",\"force_k\":";
        put_double(os, t->dihedral_k[d]);
        os << ",\"improper\":" << ((t->dihedral_improper && t->dihedral_improper[d]) ? "true" : "false") << '}';
    }
    os << "],\"lj_params\":[";
    for (int i = 0; i < n; ++i) {
        os << (i ? "," : "") << "{\"sigma\":";
        put_double(os, t->sigma[i]);
        os << ",\"epsilon\":";
        put_double(os, t->epsilon[i]);
        os << '}';
    }
    os << ']';
    // strict exclusions = exclusions minus exception partners
    std::set<std::pair<long long, long long>> expairs;
    std::vector<ExceptionRow> rows;
    for (int e = 0; e < t->n_exceptions; ++e) {
        long long i = t->exception_atoms[2 * e], j = t->exception_atoms[2 * e + 1];
        if (i > j) std::swap(i, j);
        expairs.insert({i, j});
        rows.push_back({i, j, (double)t->exception_charge_product[e], (double)t->exception_sigma[e],
                        (double)t->exception_epsilon[e]});
    }
    std::set<std::pair<long long, long long>> strict;
    os << ",\"exclusions\":[";
    for (int i = 0; i < n; ++i) {
        os << (i ? ",[" : "[");
        bool first = true;
        for (int e = t->exclusion_offsets[i]; e < t->exclusion_offsets[i + 1]; ++e) {
            const long long q = t->exclusion_atoms[e];
            const std::pair<long long, long long> key{std::min<long long>(i, q), std::max<long long>(i, q)};
            if (expairs.count(key)) continue;
            strict.insert(key);
            os << (first ? "" : ",") << q;
            first = false;
        }
        os << ']';
    }
    os << "],\"nonbonded_exceptions\":[";
    for (size_t e = 0; e < rows.size(); ++e) {
        os << (e ? "," : "") << "{\"i\":" << rows[e].i << ",\"j\":" << rows[e].j << ",\"charge_product_e2\":";
        put_double(os, rows[e].qq);
        os << ",\"sigma_A\":";
        put_double(os, rows[e].sigma);
        os << ",\"epsilon_kcal_mol\":";
        put_double(os, rows[e].eps);
        os << '}';
    }
    os << ']';
    static const char *prmtop = "0000000000000000000000000000000000000000000000000000000000000000";
    if (!rows.empty()) {
        os << ",\"generated_prmtop_sha256\":\"" << prmtop << "\"";
        os << ",\"nonbonded_exception_contract\":{\"schema_kind\":\"prismos_exact_nonbonded_exceptions_v2\","
              "\"schema_version\":2,\"semantics\":\"OPENMM_NONBONDEDFORCE_ABSOLUTE_EXCEPTION_PARAMETERS_V1\","
              "\"dynamic_charge_scaling_semantics\":"
              "\"EXCITED_STATE_MULTIPLIES_GROUND_STATE_CHARGE_PRODUCT_BY_PER_ATOM_SCALE_PRODUCT_V1\","
              "\"exact_scaled_pair_count\":"
           << rows.size() << ",\"strict_exclusion_pair_count\":" << strict.size()
           << ",\"total_openmm_exception_count\":" << rows.size() + strict.size()
           << ",\"pair_digest_algorithm\":\"SHA256_PRISMOS_NBEX_ACTIVE_V1_LE_U64_U64_F64_F64_F64\","
              "\"pair_digest_sha256\":\""
           << exception_digest(rows)
           << "\",\"strict_exclusion_digest_algorithm\":\"SHA256_PRISMOS_NBEX_STRICT_V1_LE_U64_U64\","
              "\"strict_exclusion_digest_sha256\":\""
           << strict_digest(strict) << "\",\"source_prmtop_sha256\":\"" << prmtop
           << "\",\"hardcoded_uniform_scaling_allowed\":false,\"complete\":true,\"fallback_used\":false}";
    }
    os << ",\"h_clusters\":[";
    for (int c = 0; c < t->n_clusters; ++c) {
        const int nh = t->cluster_n_hydrogens[c];
        os << (c ? "," : "") << "{\"type\":" << nh << ",\"central_atom\":" << t->cluster_central[c] << ",\"hydrogen_atoms\":[";
        for (int s = 0; s < 3; ++s) os << (s ? "," : "") << (s < nh ? t->cluster_hydrogens[3 * c + s] : -1);
        os << "],\"bond_lengths\":[";
        for (int s = 0; s < 3; ++s) {
            if (s) os << ',';
            put_double(os, s < nh ? (double)t->cluster_bond_lengths[3 * c + s] : 0.0);
        }
        os << "],\"n_hydrogens\":" << nh << ",\"inv_mass_central\":";
        put_double(os, t->cluster_inv_mass_central[c]);
        os << ",\"inv_mass_h\":";
        put_double(os, t->cluster_inv_mass_h[c]);
        os << '}';
    }
    os << ']';
    if (hmr_applied >= 0) os << ",\"hmr_applied\":" << (hmr_applied ? "true" : "false");
    os << "}\n";
    std::ofstream f(path, std::ios::binary);
    if (!f) return fail(PRISM_MD_E_INVALID, std::string("cannot write ") + path);
    const std::string s = os.str();
    f.write(s.data(), (std::streamsize)s.size());
    return f ? PRISM_MD_OK : fail(PRISM_MD_E_INVALID, std::string("write failed: ") + path);
}

}  // extern "C"
