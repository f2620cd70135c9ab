// smcrt_config.cpp — host-side input/output of libsmcrt.so that needs no GPU: the reference's
// TOML input schema (src/parse/parse.f90, parse_source.f90, parse_detectors.f90) read into the
// flat smcrt_config struct, and the detector writer (src/writer.f90:53-84).
//
// The parser accepts the TOML subset the reference's input files use: [tables], [[arrays of
// tables]], bare or quoted keys, basic and literal strings, integers, floats, booleans and
// (possibly multi-line) arrays, with # comments.  Anything else is a parse error naming the line.
#include <cctype>
#include <cerrno>
#include <cstdarg>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/smcrt.h"

int smcrt_internal_fail(int code, const char* msg);

namespace {

struct Value {
    enum Kind { STR, INT, FLT, BOOL, ARR } kind = STR;
    std::string s;
    long long i = 0;
    double f = 0.0;
    bool b = false;
    std::vector<Value> a;
    int line = 0;
};
using Table = std::map<std::string, Value>;

struct Doc {
    Table root;
    std::map<std::string, Table> tables;
    std::map<std::string, std::vector<Table>> arrays;
};

int failf(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int failf(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    return smcrt_internal_fail(code, buf);
}

class Parser {
  public:
    explicit Parser(const std::string& text) : t_(text) {}

    bool parse(Doc& doc, std::string& err) {
        Table* cur = &doc.root;
        while (true) {
            skip_blank_lines();
            if (at_end()) return true;
            if (peek() == '[') {
                bool arr = peek(1) == '[';
                pos_ += arr ? 2 : 1;
                skip_ws();
                std::string name;
                while (!at_end() && peek() != ']' && peek() != '\n') name += get();
                while (!name.empty() && std::isspace((unsigned char)name.back())) name.pop_back();
                if (name.empty()) return error(err, "empty table name");
                if (arr) {
                    if (peek() != ']' || peek(1) != ']') return error(err, "expected ]] after table name");
                    pos_ += 2;
                    doc.arrays[name].emplace_back();
                    cur = &doc.arrays[name].back();
                } else {
                    if (peek() != ']') return error(err, "expected ] after table name");
                    pos_ += 1;
                    if (doc.tables.count(name)) return error(err, "table [" + name + "] defined twice");
                    cur = &doc.tables[name];
                }
                if (!end_of_line(err)) return false;
                continue;
            }
            std::string key;
            if (!parse_key(key, err)) return false;
            skip_ws();
            if (peek() != '=') return error(err, "expected = after key '" + key + "'");
            ++pos_;
            skip_ws();
            Value v;
            if (!parse_value(v, err)) return false;
            if (cur->count(key)) return error(err, "duplicate key '" + key + "'");
            (*cur)[key] = v;
            if (!end_of_line(err)) return false;
        }
    }

  private:
    const std::string& t_;
    size_t pos_ = 0;
    int line_ = 1;

    bool at_end() const { return pos_ >= t_.size(); }
    char peek(size_t k = 0) const { return pos_ + k < t_.size() ? t_[pos_ + k] : '\0'; }
    char get() {
        char c = t_[pos_++];
        if (c == '\n') ++line_;
        return c;
    }
    bool error(std::string& err, const std::string& msg) {
        err = "line " + std::to_string(line_) + ": " + msg;
        return false;
    }
    void skip_ws() {
        while (!at_end() && (peek() == ' ' || peek() == '\t' || peek() == '\r')) ++pos_;
    }
    void skip_comment() {
        if (peek() == '#')
            while (!at_end() && peek() != '\n') ++pos_;
    }
    void skip_blank_lines() {
        while (!at_end()) {
            skip_ws();
            skip_comment();
            if (peek() == '\n') get();
            else break;
        }
    }
    // whitespace, comments and newlines inside arrays
    void skip_all() {
        while (!at_end()) {
            skip_ws();
            skip_comment();
            if (peek() == '\n') get();
            else break;
        }
    }
    bool end_of_line(std::string& err) {
        skip_ws();
        skip_comment();
        if (at_end()) return true;
        if (peek() != '\n') return error(err, std::string("unexpected '") + peek() + "' after value");
        get();
        return true;
    }
    bool parse_key(std::string& key, std::string& err) {
        if (peek() == '"' || peek() == '\'') {
            Value v;
            if (!parse_string(v, err)) return false;
            key = v.s;
            return true;
        }
        while (!at_end() && (std::isalnum((unsigned char)peek()) || peek() == '_' || peek() == '-')) key += get();
        if (key.empty()) return error(err, std::string("expected a key, found '") + peek() + "'");
        return true;
    }
    bool parse_string(Value& v, std::string& err) {
        const char q = get();
        if (peek() == q && peek(1) == q) return error(err, "multi-line strings are not supported");
        v.kind = Value::STR;
        v.line = line_;
        while (true) {
            if (at_end() || peek() == '\n') return error(err, "unterminated string");
            char c = get();
            if (c == q) return true;
            if (q == '"' && c == '\\') {
                char e = at_end() ? '\0' : get();
                switch (e) {
                case 'n': v.s += '\n'; break;
                case 't': v.s += '\t'; break;
                case 'r': v.s += '\r'; break;
                case 'b': v.s += '\b'; break;
                case 'f': v.s += '\f'; break;
                case '"': v.s += '"'; break;
                case '\\': v.s += '\\'; break;
                default: return error(err, "unsupported escape in string");
                }
            } else {
                v.s += c;
            }
        }
    }
    bool parse_value(Value& v, std::string& err) {
        v.line = line_;
        const char c = peek();
        if (c == '"' || c == '\'') return parse_string(v, err);
        if (c == '[') {
            get();
            v.kind = Value::ARR;
            while (true) {
                skip_all();
                if (peek() == ']') { get(); return true; }
                Value e;
                if (!parse_value(e, err)) return false;
                v.a.push_back(e);
                skip_all();
                if (peek() == ',') { get(); continue; }
                if (peek() == ']') { get(); return true; }
                return error(err, "expected , or ] in array");
            }
        }
        if (c == '{') return error(err, "inline tables are not supported");
        std::string tok;
        while (!at_end() && !std::isspace((unsigned char)peek()) && peek() != ',' && peek() != ']' && peek() != '#')
            tok += get();
        if (tok == "true" || tok == "false") {
            v.kind = Value::BOOL;
            v.b = tok == "true";
            return true;
        }
        std::string num;
        for (char ch : tok)
            if (ch != '_') num += ch;
        if (num.empty()) return error(err, "expected a value");
        const bool is_float = num.find_first_of(".eE") != std::string::npos || num.find("inf") != std::string::npos ||
                              num.find("nan") != std::string::npos;
        char* end = nullptr;
        errno = 0;
        if (is_float) {
            v.kind = Value::FLT;
            v.f = std::strtod(num.c_str(), &end);
        } else {
            v.kind = Value::INT;
            v.i = std::strtoll(num.c_str(), &end, 10);
        }
        if (!end || *end != '\0' || errno == ERANGE) return error(err, "malformed value '" + tok + "'");
        return true;
    }
};

// ---- typed access with the reference's defaults ----------------------------------------------
struct Ctx {
    const std::string& file;
    std::string section;
};

int bad(const Ctx& c, const Value& v, const std::string& key, const char* what) {
    return failf(SMCRT_EINVAL, "%s:%d: [%s] %s: %s", c.file.c_str(), v.line, c.section.c_str(), key.c_str(), what);
}

int get_real(const Ctx& c, const Table& t, const std::string& k, double dflt, double* out) {
    auto it = t.find(k);
    if (it == t.end()) { *out = dflt; return SMCRT_OK; }
    const Value& v = it->second;
    if (v.kind == Value::FLT) *out = v.f;
    else if (v.kind == Value::INT) *out = (double)v.i;
    else return bad(c, v, k, "expected a number");
    return SMCRT_OK;
}

template <class I>
int get_int(const Ctx& c, const Table& t, const std::string& k, long long dflt, I* out) {
    auto it = t.find(k);
    if (it == t.end()) { *out = (I)dflt; return SMCRT_OK; }
    const Value& v = it->second;
    if (v.kind != Value::INT) return bad(c, v, k, "expected an integer");
    *out = (I)v.i;
    return SMCRT_OK;
}

int get_bool(const Ctx& c, const Table& t, const std::string& k, bool dflt, int32_t* out) {
    auto it = t.find(k);
    if (it == t.end()) { *out = dflt; return SMCRT_OK; }
    const Value& v = it->second;
    if (v.kind != Value::BOOL) return bad(c, v, k, "expected true or false");
    *out = v.b ? 1 : 0;
    return SMCRT_OK;
}

int get_str(const Ctx& c, const Table& t, const std::string& k, const char* dflt, char* out, size_t cap) {
    auto it = t.find(k);
    std::string s = dflt;
    if (it != t.end()) {
        if (it->second.kind != Value::STR) return bad(c, it->second, k, "expected a string");
        s = it->second.s;
    }
    if (s.size() + 1 > cap) {
        return failf(SMCRT_ERANGE, "%s: [%s] %s: string longer than %zu characters", c.file.c_str(), c.section.c_str(),
                     k.c_str(), cap - 1);
    }
    std::memcpy(out, s.c_str(), s.size() + 1);
    return SMCRT_OK;
}

// `get_vector` (src/parse/parse_helpers.f90:14-49): a 3-element numeric array.
// Returns 1 (found), 0 (absent) or a negative status.
int get_vec3(const Ctx& c, const Table& t, const std::string& k, double out[3]) {
    auto it = t.find(k);
    if (it == t.end()) return 0;
    const Value& v = it->second;
    if (v.kind != Value::ARR || v.a.size() != 3) return bad(c, v, k, ("Expected vector of size 3 for " + k).c_str());
    for (int j = 0; j < 3; ++j) {
        if (v.a[j].kind == Value::FLT) out[j] = v.a[j].f;
        else if (v.a[j].kind == Value::INT) out[j] = (double)v.a[j].i;
        else return bad(c, v, k, ("Expected vector of size 3 for " + k).c_str());
    }
    return 1;
}

#define TRY(x)                 \
    do {                       \
        int rc_ = (x);         \
        if (rc_ < 0) return rc_; \
    } while (0)

int parse_source(const std::string& file, const Doc& doc, smcrt_config* o) {
    auto it = doc.tables.find("source");
    if (it == doc.tables.end()) return failf(SMCRT_EINVAL, "%s: Simulation needs Source table", file.c_str());
    const Table& t = it->second;
    Ctx c{file, "source"};
    TRY(get_str(c, t, "name", "point", o->source_name, sizeof o->source_name));
    TRY(get_int(c, t, "nphotons", 1000000, &o->nphotons));
    const std::string name = o->source_name;
    int got = get_vec3(c, t, "position", o->position);
    TRY(got);
    o->has_position = got;
    if (name != "uniform" && !got)
        return failf(SMCRT_EINVAL, "%s: [source] Expected vector of size 3 for position", file.c_str());
    // direction: a 3-vector, or a cardinal string (src/parse/parse_source.f90:64-108)
    auto d = t.find("direction");
    o->has_direction = 0;
    if (d != t.end() && d->second.kind == Value::ARR) {
        TRY(get_vec3(c, t, "direction", o->direction));
        o->has_direction = 1;
    } else if (d != t.end() && d->second.kind == Value::STR) {
        static const struct { const char* s; double v[3]; } card[] = {
            {"x", {1, 0, 0}}, {"-x", {-1, 0, 0}}, {"y", {0, 1, 0}}, {"-y", {0, -1, 0}}, {"z", {0, 0, 1}}, {"-z", {0, 0, -1}}};
        bool ok = false;
        for (const auto& e : card)
            if (d->second.s == e.s) {
                for (int j = 0; j < 3; ++j) o->direction[j] = e.v[j];
                ok = true;
            }
        if (!ok) return bad(c, d->second, "direction", "Direction needs a cardinal direction i.e x, y, or z");
        o->has_direction = 1;
    } else if (d != t.end()) {
        return bad(c, d->second, "direction", "expected a vector or a cardinal direction");
    } else if (name != "point") {
        return failf(SMCRT_EINVAL, "%s: [source] Need to specify direction for source type!", file.c_str());
    }
    static const double corners[3][3] = {{-1, -1, 1}, {2, 0, 0}, {0, 2, 0}};
    double* pts[3] = {o->point1, o->point2, o->point3};
    for (int p = 0; p < 3; ++p) {
        const std::string key = "point" + std::to_string(p + 1);
        for (int j = 0; j < 3; ++j) pts[p][j] = corners[p][j];
        auto pi = t.find(key);
        if (pi != t.end()) {
            if (pi->second.kind != Value::ARR || pi->second.a.size() < 3)
                return bad(c, pi->second, key, "Need a matrix row for points");
            for (int j = 0; j < 3; ++j) {
                const Value& e = pi->second.a[j];
                if (e.kind == Value::FLT) pts[p][j] = e.f;
                else if (e.kind == Value::INT) pts[p][j] = (double)e.i;
                else return bad(c, pi->second, key, "Need a matrix row for points");
            }
        } else if (name == "uniform") {
            return failf(SMCRT_EINVAL, "%s: [source] Uniform source requires %s variable", file.c_str(), key.c_str());
        }
    }
    TRY(get_real(c, t, "radius", 0.5, &o->radius));
    TRY(get_real(c, t, "beta", 5.0, &o->beta));
    TRY(get_real(c, t, "radius_hi", 0.6, &o->radius_hi));
    TRY(get_str(c, t, "annulus_type", "gaussian", o->annulus_type, sizeof o->annulus_type));
    return SMCRT_OK;
}

const Table* need(const std::string& file, const Doc& doc, const char* name, int* rc) {
    auto it = doc.tables.find(name);
    if (it == doc.tables.end()) {
        *rc = failf(SMCRT_EINVAL, "%s: Need %s table in input param file", file.c_str(), name);
        return nullptr;
    }
    *rc = SMCRT_OK;
    return &it->second;
}

int parse_detectors(const std::string& file, const Doc& doc, smcrt_config* o) {
    std::strcpy(o->history_file, "photPos.obj");
    auto it = doc.arrays.find("detectors");
    if (it == doc.arrays.end()) return SMCRT_OK;
    const std::vector<Table>& arr = it->second;
    if (arr.size() > SMCRT_MAX_DETECTORS)
        return failf(SMCRT_ERANGE, "%s: %zu detectors, at most %d supported", file.c_str(), arr.size(), SMCRT_MAX_DETECTORS);
    // the reference orders the detector array by type: circles, then annuli, then cameras
    // (src/parse/parse_detectors.f90:96-109)
    int n = 0;
    for (int want = SMCRT_DET_CIRCLE; want <= SMCRT_DET_CAMERA; ++want) {
        for (size_t i = 0; i < arr.size(); ++i) {
            const Table& t = arr[i];
            Ctx c{file, "detectors." + std::to_string(i + 1)};
            char type[32];
            auto ti = t.find("type");
            if (ti == t.end()) return failf(SMCRT_EINVAL, "%s: [[detectors]] %zu: missing type", file.c_str(), i + 1);
            TRY(get_str(c, t, "type", "", type, sizeof type));
            const std::string ty = type;
            const int kind = ty == "circle" ? SMCRT_DET_CIRCLE : ty == "annulus" ? SMCRT_DET_ANNULUS : ty == "camera" ? SMCRT_DET_CAMERA : 0;
            if (!kind) return bad(c, ti->second, "type", "Invalid detector type. Valid types are [circle, annulus, camera]");
            if (kind != want) continue;
            TRY(get_str(c, t, "historyFileName", "photPos.obj", o->history_file, sizeof o->history_file));
            smcrt_detector& d = o->detectors[n++];
            std::memset(&d, 0, sizeof d);
            d.kind = kind;
            TRY(get_int(c, t, "layer", 1, &d.layer));
            TRY(get_int(c, t, "nbins", 100, &d.nbins));
            TRY(get_real(c, t, "maxval", 100.0, &d.maxval));
            TRY(get_bool(c, t, "trackHistory", false, &d.track_history));
            if (kind == SMCRT_DET_CAMERA) {
                static const double p1[3] = {-1, -1, -1}, p2[3] = {2, 0, 0}, p3[3] = {0, 2, 0};
                for (int j = 0; j < 3; ++j) { d.pos[j] = p1[j]; d.p2[j] = p2[j]; d.p3[j] = p3[j]; }
                TRY(get_vec3(c, t, "p1", d.pos));
                TRY(get_vec3(c, t, "p2", d.p2));
                TRY(get_vec3(c, t, "p3", d.p3));
            } else {
                int got = get_vec3(c, t, "position", d.pos);
                TRY(got);
                if (!got) return failf(SMCRT_EINVAL, "%s: [[detectors]] %zu: Expected vector of size 3 for position", file.c_str(), i + 1);
                d.dir[0] = 0; d.dir[1] = 0; d.dir[2] = -1;
                TRY(get_vec3(c, t, "direction", d.dir));
                TRY(get_real(c, t, "radius1", 0.0, &d.radius1));
                if (kind == SMCRT_DET_ANNULUS) {
                    TRY(get_real(c, t, "radius2", 0.0, &d.radius2));
                    if (d.radius2 <= d.radius1)
                        return failf(SMCRT_EINVAL, "%s: [[detectors]] %zu: Radii are invalid: Expected radius2 (%g) > radius 1 (%g)",
                                     file.c_str(), i + 1, d.radius2, d.radius1);
                }
            }
        }
    }
    o->n_detectors = n;
    return SMCRT_OK;
}

}  // namespace

extern "C" int smcrt_config_load(const char* path, smcrt_config* o) {
    if (!path || !o) return failf(SMCRT_EINVAL, "null path or config");
    std::memset(o, 0, sizeof *o);
    FILE* f = std::fopen(path, "rb");
    if (!f) return failf(SMCRT_EIO, "cannot open %s", path);
    std::string text;
    char buf[65536];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, n);
    std::fclose(f);
    const std::string file = path;
    Doc doc;
    std::string err;
    if (!Parser(text).parse(doc, err)) return failf(SMCRT_EINVAL, "%s: %s", path, err.c_str());

    TRY(parse_source(file, doc, o));
    int rc;
    // [grid] (src/parse/parse.f90: parse_grid)
    const Table* t = need(file, doc, "grid", &rc);
    if (!t) return rc;
    {
        Ctx c{file, "grid"};
        TRY(get_int(c, *t, "nxg", 200, &o->nxg));
        TRY(get_int(c, *t, "nyg", 200, &o->nyg));
        TRY(get_int(c, *t, "nzg", 200, &o->nzg));
        TRY(get_real(c, *t, "xmax", 1.0, &o->xmax));
        TRY(get_real(c, *t, "ymax", 1.0, &o->ymax));
        TRY(get_real(c, *t, "zmax", 1.0, &o->zmax));
        TRY(get_str(c, *t, "units", "cm", o->units, sizeof o->units));
    }
    // [geometry] (parse_geometry)
    t = need(file, doc, "geometry", &rc);
    if (!t) return rc;
    {
        Ctx c{file, "geometry"};
        TRY(get_str(c, *t, "geom_name", "sphere", o->geom_name, sizeof o->geom_name));
        TRY(get_real(c, *t, "tau", 10.0, &o->tau));
        TRY(get_int(c, *t, "num_spheres", 10, &o->num_spheres));
        TRY(get_real(c, *t, "musb", 0.0, &o->musb));
        TRY(get_real(c, *t, "muab", 0.01, &o->muab));
        TRY(get_real(c, *t, "musc", 0.0, &o->musc));
        TRY(get_real(c, *t, "muac", 0.01, &o->muac));
        TRY(get_real(c, *t, "hgg", 0.7, &o->hgg));
    }
    TRY(parse_detectors(file, doc, o));
    // [output] (parse_output)
    t = need(file, doc, "output", &rc);
    if (!t) return rc;
    {
        Ctx c{file, "output"};
        TRY(get_str(c, *t, "fluence", "fluence.nrrd", o->fluence, sizeof o->fluence));
        TRY(get_str(c, *t, "absorb", "absorb.nrrd", o->absorb, sizeof o->absorb));
        TRY(get_str(c, *t, "render", "geom_render.nrrd", o->render, sizeof o->render));
        TRY(get_bool(c, *t, "render_geom", false, &o->render_geom));
        o->render_size[0] = o->render_size[1] = o->render_size[2] = 200;
        auto rs = t->find("render_size");
        if (rs != t->end()) {
            if (rs->second.kind != Value::ARR || rs->second.a.size() < 3)
                return bad(c, rs->second, "render_size", "Need a vector of size 3 for render_size.");
            for (int j = 0; j < 3; ++j) {
                if (rs->second.a[j].kind != Value::INT) return bad(c, rs->second, "render_size", "expected integers");
                o->render_size[j] = (int32_t)rs->second.a[j].i;
            }
        }
        TRY(get_bool(c, *t, "overwrite", false, &o->overwrite));
    }
    // [simulation] (parse_simulation)
    t = need(file, doc, "simulation", &rc);
    if (!t) return rc;
    {
        Ctx c{file, "simulation"};
        TRY(get_int(c, *t, "iseed", 123456789, &o->iseed));
        TRY(get_bool(c, *t, "tev", false, &o->tev));
        TRY(get_bool(c, *t, "absorb", false, &o->absorb_flag));
        TRY(get_bool(c, *t, "load_checkpoint", false, &o->load_checkpoint));
        TRY(get_str(c, *t, "checkpoint_file", "check.ckpt", o->checkpoint_file, sizeof o->checkpoint_file));
        TRY(get_int(c, *t, "checkpoint_every_n", 1000000, &o->checkpoint_every_n));
    }
    return SMCRT_OK;
}

extern "C" int smcrt_write_detector(const char* path, const smcrt_detector* d, const uint64_t* counts) {
    if (!path || !d || (!counts && d->kind == SMCRT_DET_CIRCLE)) return failf(SMCRT_EINVAL, "bad arguments");
    FILE* f = std::fopen(path, "w");
    if (!f) return failf(SMCRT_EIO, "cannot open %s for writing", path);
    if (d->kind == SMCRT_DET_CIRCLE) {
        const int nb = d->nbins + 1;
        const double bw = d->nbins == 0 ? 1.0 : d->maxval / (double)(d->nbins - 1);
        for (int j = 1; j <= nb; ++j) std::fprintf(f, " %.16E %.16E\n", (double)j * bw, (double)counts[j - 1]);
    }
    if (std::fclose(f) != 0) return failf(SMCRT_EIO, "short write to %s", path);
    return SMCRT_OK;
}
