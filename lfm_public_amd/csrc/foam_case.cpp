// foam_case.cpp -- OpenFOAM ASCII case input and time-directory output
// (include/lfm_case.h).
//
// The reference reads its inputs through OpenFOAM: the polyMesh with
// fvMesh (polyMeshReaderOF/polyMeshReaderOF.cpp:5-60), the processor
// addressing files (:21-44), the fields p, T, U, alpha
// (polyMeshReaderOF.cpp:690-760, cfd_v0.cpp:919-935), the dictionaries with
// IOdictionary lookups along a '/'-separated sub-dictionary path
// (dictReaderOF/dictReaderOF.cpp) and the run controls with Foam::Time
// (runTimeManagerOF/runTimeManagerOF.cpp).  This file parses the ASCII
// formats directly: FoamFile headers, // and /* */ comments, sub-dictionaries,
// counted lists "N ( ... )", uniform lists "N{v}", "uniform" /
// "nonuniform List<...>" fields.  Files whose header says "format binary"
// hold their primitive lists as raw little-endian bytes, "N(<bytes>)" and
// "N{<bytes>}", with label and scalar sizes from the header's arch entry
// ("LSB;label=32;scalar=64" by default); everything else in them is text.
#include <sys/stat.h>

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <dirent.h>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "lfm_case.h"
#include "lfm_hip.h"

namespace lfm {
int fail(int code, const std::string &msg);
}

namespace {

using lfm::fail;

struct ParseError {
    std::string msg;
};

[[noreturn]] void perr(const std::string &file, const std::string &msg) { throw ParseError{file + ": " + msg}; }

bool read_text(const std::string &path, std::string &out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::ostringstream ss;
    ss << f.rdbuf();
    out = ss.str();
    return true;
}

bool is_dir(const std::string &p) {
    struct stat st;
    return stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

// ---------------------------------------------------------------------------
// lexer over an in-memory file
// ---------------------------------------------------------------------------
struct Lexer {
    const std::string &s;
    std::string file;
    size_t i = 0;
    // binary format: primitive lists are raw bytes of these sizes
    bool binary = false;
    int label_bytes = 4, scalar_bytes = 8;
    Lexer(const std::string &text, const std::string &name) : s(text), file(name) {}

    void skip() {
        for (;;) {
            while (i < s.size() && std::isspace((unsigned char)s[i])) i++;
            if (i + 1 < s.size() && s[i] == '/' && s[i + 1] == '/') {
                while (i < s.size() && s[i] != '\n') i++;
                continue;
            }
            if (i + 1 < s.size() && s[i] == '/' && s[i + 1] == '*') {
                size_t e = s.find("*/", i + 2);
                i = e == std::string::npos ? s.size() : e + 2;
                continue;
            }
            break;
        }
    }
    bool eof() {
        skip();
        return i >= s.size();
    }
    char peek() {
        skip();
        return i < s.size() ? s[i] : '\0';
    }
    static bool punct(char c) { return c == '{' || c == '}' || c == '(' || c == ')' || c == ';' || c == '[' || c == ']'; }
    // next token: a punctuation character, a quoted string (with quotes) or a word
    std::string next() {
        skip();
        if (i >= s.size()) perr(file, "unexpected end of file");
        const char c = s[i];
        if (punct(c)) {
            i++;
            return std::string(1, c);
        }
        if (c == '"') {
            size_t e = s.find('"', i + 1);
            if (e == std::string::npos) perr(file, "unterminated string");
            std::string t = s.substr(i, e + 1 - i);
            i = e + 1;
            return t;
        }
        size_t b = i;
        while (i < s.size() && !std::isspace((unsigned char)s[i]) && !punct(s[i])) {
            if (s[i] == '/' && i + 1 < s.size() && (s[i + 1] == '/' || s[i + 1] == '*')) break;
            i++;
        }
        return s.substr(b, i - b);
    }
    void expect(char c) {
        std::string t = next();
        if (t.size() != 1 || t[0] != c) perr(file, std::string("expected '") + c + "', found '" + t + "'");
    }
    // fast numbers for the large lists
    double num() {
        skip();
        const char *b = s.c_str() + i;
        char *e = nullptr;
        double v = std::strtod(b, &e);
        if (e == b) perr(file, "expected a number near offset " + std::to_string(i));
        i += (size_t)(e - b);
        return v;
    }
    long lab() {
        skip();
        const char *b = s.c_str() + i;
        char *e = nullptr;
        long v = std::strtol(b, &e, 10);
        if (e == b) perr(file, "expected an integer near offset " + std::to_string(i));
        i += (size_t)(e - b);
        return v;
    }
    // a list size: non-negative and within the 32-bit labels of this
    // library; with item_bytes > 0 (a list written out item by item), also
    // at most what the rest of the file can hold
    long count(size_t item_bytes) {
        const long n = lab();
        if (n < 0) perr(file, "negative list size");
        if (n > INT32_MAX) perr(file, "list size beyond the 32-bit label range");
        const size_t left = i < s.size() ? s.size() - i : 0;
        if (item_bytes > 0 && (size_t)n > left / item_bytes)
            perr(file, "list of " + std::to_string(n) + " items is longer than the file");
        return n;
    }
    // binary list payload: the opening bracket (after white space only --
    // the bytes may contain anything), `bytes` raw bytes, the closing one
    const char *raw(size_t bytes, char open, char close) {
        while (i < s.size() && std::isspace((unsigned char)s[i])) i++;
        if (i >= s.size() || s[i] != open) perr(file, std::string("expected '") + open + "' before binary data");
        const char *p = s.data() + i + 1;
        if (bytes >= s.size() - (i + 1) || s[i + 1 + bytes] != close)
            perr(file, "binary list of " + std::to_string(bytes) + " bytes is truncated");
        i += bytes + 2;
        return p;
    }
    // which bracket follows (white space only skipped)
    char bracket() {
        size_t j = i;
        while (j < s.size() && std::isspace((unsigned char)s[j])) j++;
        return j < s.size() ? s[j] : '\0';
    }
    double scalar_at(const char *p, size_t k) const {
        if (scalar_bytes == 8) {
            double v;
            std::memcpy(&v, p + 8 * k, 8);
            return v;
        }
        float v;
        std::memcpy(&v, p + 4 * k, 4);
        return v;
    }
    long label_at(const char *p, size_t k) const {
        if (label_bytes == 4) {
            int32_t v;
            std::memcpy(&v, p + 4 * k, 4);
            return v;
        }
        int64_t v;
        std::memcpy(&v, p + 8 * k, 8);
        return (long)v;
    }
    // components of one element of a List<type>
    static int ncomp_of(const std::string &type) {
        if (type == "scalar" || type == "label") return 1;
        if (type == "vector") return 3;
        if (type == "symmTensor") return 6;
        if (type == "tensor") return 9;
        if (type == "sphericalTensor") return 1;
        return -1;
    }
};

// ---------------------------------------------------------------------------
// dictionaries
// ---------------------------------------------------------------------------
struct Dict;
struct Entry {
    std::string key;
    std::vector<std::string> tokens;  // value tokens when not a dictionary
    std::unique_ptr<Dict> dict;
};
struct Dict {
    std::vector<Entry> entries;
    const Entry *find(const std::string &k) const {
        // the last definition wins, as in OpenFOAM
        for (size_t n = entries.size(); n-- > 0;)
            if (entries[n].key == k) return &entries[n];
        return nullptr;
    }
    const Dict *sub(const std::string &k) const {
        const Entry *e = find(k);
        return e && e->dict ? e->dict.get() : nullptr;
    }
};

// entries until '}' (nested) or end of file (top level)
void parse_dict(Lexer &L, Dict &d, bool nested) {
    for (;;) {
        if (L.eof()) {
            if (nested) perr(L.file, "missing '}'");
            return;
        }
        if (L.peek() == '}') {
            if (!nested) perr(L.file, "unbalanced '}'");
            L.next();
            return;
        }
        std::string key = L.next();
        if (key == ";") continue;
        if (!key.empty() && key[0] == '#') {  // directives (#include, #inputMode ...): skip the line
            while (L.i < L.s.size() && L.s[L.i] != '\n') L.i++;
            continue;
        }
        Entry e;
        e.key = key;
        if (L.peek() == '{') {
            L.next();
            e.dict = std::make_unique<Dict>();
            parse_dict(L, *e.dict, true);
        } else {
            int depth = 0;
            for (;;) {
                std::string t = L.next();
                if (L.binary && t.rfind("List<", 0) == 0 && t.back() == '>') {
                    // "List<type> N(<bytes>)": step over the payload
                    const int nc = Lexer::ncomp_of(t.substr(5, t.size() - 6));
                    if (nc < 0) perr(L.file, "unsupported binary " + t);
                    e.tokens.push_back(t);
                    const long n = L.lab();
                    const size_t el = (size_t)nc * (t == "List<label>" ? L.label_bytes : L.scalar_bytes);
                    const char br = L.bracket();
                    L.raw(br == '{' ? el : (size_t)n * el, br == '{' ? '{' : '(', br == '{' ? '}' : ')');
                    e.tokens.push_back("<binary>");
                    continue;
                }
                if (t == "(" || t == "[" || t == "{") depth++;
                if (t == ")" || t == "]" || t == "}") depth--;
                if (depth < 0) perr(L.file, "unbalanced bracket in entry " + key);
                if (t == ";" && depth == 0) break;
                e.tokens.push_back(t);
            }
        }
        d.entries.push_back(std::move(e));
    }
}

void skip_header(Lexer &L, Dict *header) {
    size_t save = L.i;
    if (!L.eof() && L.next() == "FoamFile" && L.peek() == '{') {
        L.next();
        Dict h;
        parse_dict(L, h, true);
        if (header) *header = std::move(h);
        return;
    }
    L.i = save;
}

std::string header_word(const Dict &h, const char *k) {
    const Entry *e = h.find(k);
    if (!e || e->tokens.empty()) return "";
    std::string t = e->tokens[0];
    if (t.size() >= 2 && t.front() == '"') t = t.substr(1, t.size() - 2);
    return t;
}

// format and arch of a FoamFile header into the lexer
void set_format(Lexer &L, const Dict &h) {
    L.binary = header_word(h, "format") == "binary";
    if (!L.binary) return;
    const std::string arch = header_word(h, "arch");
    if (arch.find("MSB") != std::string::npos) perr(L.file, "big-endian binary data is not supported");
    auto size_of = [&](const char *key, int dflt) {
        const size_t p = arch.find(key);
        if (p == std::string::npos) return dflt;
        const int bits = std::atoi(arch.c_str() + p + std::strlen(key));
        if (bits != 32 && bits != 64) perr(L.file, std::string("unsupported ") + key + " size in arch");
        return bits / 8;
    };
    L.label_bytes = size_of("label=", 4);
    L.scalar_bytes = size_of("scalar=", 8);
}

struct DictFile {
    std::string path;
    Dict root;
};

std::vector<std::string> split_path(const std::string &p) {
    std::vector<std::string> out;
    std::string cur;
    for (char ch : p) {
        if (ch == '/') {
            if (!cur.empty()) out.push_back(cur);
            cur.clear();
        } else {
            cur += ch;
        }
    }
    if (!cur.empty()) out.push_back(cur);
    return out;
}

const Entry *lookup(const DictFile &f, const std::string &sub, const std::string &key) {
    const Dict *d = &f.root;
    for (const std::string &s : split_path(sub)) {
        d = d->sub(s);
        if (!d) return nullptr;
    }
    return d->find(key);
}

// readDictScalar / readDictBool / readDictWord with (default, allowDefault)
double dict_scalar(const DictFile &f, const std::string &sub, const std::string &key, double dflt, bool allow) {
    const Entry *e = lookup(f, sub, key);
    if (!e || e->tokens.empty()) {
        if (allow) return dflt;
        perr(f.path, "missing scalar " + (sub.empty() ? key : sub + "/" + key));
    }
    const char *b = e->tokens[0].c_str();
    char *end = nullptr;
    double v = std::strtod(b, &end);
    if (end == b || *end) perr(f.path, "not a number: " + key + " " + e->tokens[0]);
    return v;
}

bool dict_bool(const DictFile &f, const std::string &sub, const std::string &key, bool dflt, bool allow) {
    const Entry *e = lookup(f, sub, key);
    if (!e || e->tokens.empty()) {
        if (allow) return dflt;
        perr(f.path, "missing switch " + key);
    }
    const std::string &t = e->tokens[0];
    if (t == "true" || t == "yes" || t == "on" || t == "y" || t == "t" || t == "1") return true;
    if (t == "false" || t == "no" || t == "off" || t == "n" || t == "f" || t == "0" || t == "none") return false;
    perr(f.path, "not a switch: " + key + " " + t);
}

std::string dict_word(const DictFile &f, const std::string &sub, const std::string &key, const std::string &dflt,
                      bool allow) {
    const Entry *e = lookup(f, sub, key);
    if (!e || e->tokens.empty()) {
        if (allow) return dflt;
        perr(f.path, "missing word " + key);
    }
    std::string t = e->tokens[0];
    if (t.size() >= 2 && t.front() == '"') t = t.substr(1, t.size() - 2);
    return t;
}

DictFile load_dict(const std::string &path) {
    DictFile f;
    f.path = path;
    std::string text;
    if (!read_text(path, text)) perr(path, "cannot read");
    Lexer L(text, path);
    skip_header(L, nullptr);
    parse_dict(L, f.root, false);
    return f;
}

// ---------------------------------------------------------------------------
// polyMesh lists
// ---------------------------------------------------------------------------
struct ListFile {
    std::string text;
    std::unique_ptr<Lexer> L;
    Dict header;
};

void open_list(const std::string &path, ListFile &lf) {
    if (!read_text(path, lf.text)) perr(path, "cannot read");
    lf.L = std::make_unique<Lexer>(lf.text, path);
    skip_header(*lf.L, &lf.header);
    set_format(*lf.L, lf.header);
}

int32_t to_label32(Lexer &L, long x) {
    if (x < INT32_MIN || x > INT32_MAX) perr(L.file, "label out of the 32-bit range");
    return (int32_t)x;
}

// "N ( a b ... )" or "N{v}" of labels (raw bytes in a binary file)
std::vector<int32_t> read_labels(Lexer &L) {
    long n = L.count(0);
    std::vector<int32_t> v;
    const bool uniform = L.bracket() == '{';
    // item by item: at least a label's bytes (binary) or one character each
    if (!uniform && (size_t)n > (L.s.size() - std::min(L.i, L.s.size())) / (L.binary ? L.label_bytes : 1))
        perr(L.file, "list of " + std::to_string(n) + " labels is longer than the file");
    if (L.binary) {
        const bool uni = uniform;
        const char *p = uni ? L.raw((size_t)L.label_bytes, '{', '}') : L.raw((size_t)n * L.label_bytes, '(', ')');
        v.resize((size_t)n);
        for (long k = 0; k < n; k++) v[k] = to_label32(L, L.label_at(p, uni ? 0 : (size_t)k));
        return v;
    }
    if (L.peek() == '{') {
        L.next();
        long x = L.lab();
        L.expect('}');
        v.assign((size_t)n, (int32_t)x);
        return v;
    }
    L.expect('(');
    v.resize((size_t)n);
    for (long k = 0; k < n; k++) v[k] = (int32_t)L.lab();
    L.expect(')');
    return v;
}

std::vector<int32_t> read_label_file(const std::string &path) {
    ListFile lf;
    open_list(path, lf);
    return read_labels(*lf.L);
}

// ---------------------------------------------------------------------------
// fields
// ---------------------------------------------------------------------------
struct Field {
    std::vector<double> v;          // n * ncomp
    std::string boundary_text;      // the boundaryField block as written
    std::string dimensions;
    bool binary = false;            // boundary_text may hold raw lists
    int label_bytes = 4, scalar_bytes = 8;
};

// The boundaryField block with its binary lists ("List<T> N(<bytes>)")
// spelled out in ASCII, for the ASCII time-directory writer.
std::string boundary_ascii(const Field &F) {
    if (!F.binary) return F.boundary_text;
    const std::string &t = F.boundary_text;
    Lexer L(t, "boundaryField");
    L.binary = true;
    L.label_bytes = F.label_bytes;
    L.scalar_bytes = F.scalar_bytes;
    std::string out;
    size_t copied = 0;
    for (;;) {
        const size_t at = t.find("List<", L.i);
        if (at == std::string::npos) break;
        const size_t close = t.find('>', at);
        if (close == std::string::npos) break;
        const std::string type = t.substr(at + 5, close - at - 5);
        const int nc = Lexer::ncomp_of(type);
        if (nc < 0) perr("boundaryField", "unsupported binary List<" + type + ">");
        L.i = close + 1;
        const long n = L.lab();
        const bool lab = type == "label";
        const size_t el = (size_t)nc * (lab ? L.label_bytes : L.scalar_bytes);
        const bool uni = L.bracket() == '{';
        const char *p = uni ? L.raw(el, '{', '}') : L.raw((size_t)n * el, '(', ')');
        out += t.substr(copied, close + 1 - copied);
        char buf[64];
        out += " " + std::to_string(n) + (uni ? "{" : "\n(\n");
        const long m = uni ? 1 : n;
        for (long c = 0; c < m; c++) {
            if (nc > 1) out += "(";
            for (int k = 0; k < nc; k++) {
                if (lab)
                    std::snprintf(buf, sizeof buf, "%ld", L.label_at(p, (size_t)c * nc + k));
                else
                    std::snprintf(buf, sizeof buf, "%.17g", L.scalar_at(p, (size_t)c * nc + k));
                out += (k ? " " : "") + std::string(buf);
            }
            out += nc > 1 ? ")" : "";
            out += uni ? "" : "\n";
        }
        out += uni ? "}" : ")";
        copied = L.i;
    }
    return out + t.substr(copied);
}

Field read_field(const std::string &path, long n, int ncomp) {
    Field F;
    std::string text;
    if (!read_text(path, text)) perr(path, "cannot read");
    Lexer L(text, path);
    Dict header;
    skip_header(L, &header);
    set_format(L, header);
    F.binary = L.binary;
    F.label_bytes = L.label_bytes;
    F.scalar_bytes = L.scalar_bytes;
    bool got = false;
    while (!L.eof()) {
        size_t start = L.i;
        std::string key = L.next();
        if (key == "internalField") {
            std::string kind = L.next();
            F.v.resize((size_t)n * ncomp);
            if (kind == "uniform") {
                double u[3] = {0, 0, 0};
                if (ncomp == 1) {
                    u[0] = L.num();
                } else {
                    L.expect('(');
                    for (int k = 0; k < ncomp; k++) u[k] = L.num();
                    L.expect(')');
                }
                for (long c = 0; c < n; c++)
                    for (int k = 0; k < ncomp; k++) F.v[(size_t)c * ncomp + k] = u[k];
            } else if (kind == "nonuniform") {
                L.next();  // List<scalar> / List<vector>
                long m = L.lab();
                if (m != n) perr(path, "internalField has " + std::to_string(m) + " values for " + std::to_string(n) + " cells");
                if (L.binary) {
                    const bool uni = L.bracket() == '{';
                    const size_t el = (size_t)ncomp * L.scalar_bytes;
                    const char *p = uni ? L.raw(el, '{', '}') : L.raw((size_t)n * el, '(', ')');
                    for (long c = 0; c < n; c++)
                        for (int k = 0; k < ncomp; k++)
                            F.v[(size_t)c * ncomp + k] = L.scalar_at(p, (uni ? 0 : (size_t)c * ncomp) + k);
                } else if (L.peek() == '{') {
                    L.next();
                    double u[3] = {0, 0, 0};
                    if (ncomp == 1) {
                        u[0] = L.num();
                    } else {
                        L.expect('(');
                        for (int k = 0; k < ncomp; k++) u[k] = L.num();
                        L.expect(')');
                    }
                    L.expect('}');
                    for (long c = 0; c < n; c++)
                        for (int k = 0; k < ncomp; k++) F.v[(size_t)c * ncomp + k] = u[k];
                } else {
                    L.expect('(');
                    for (long c = 0; c < n; c++) {
                        if (ncomp == 1) {
                            F.v[c] = L.num();
                        } else {
                            L.expect('(');
                            for (int k = 0; k < ncomp; k++) F.v[(size_t)c * ncomp + k] = L.num();
                            L.expect(')');
                        }
                    }
                    L.expect(')');
                }
            } else {
                perr(path, "internalField must be uniform or nonuniform");
            }
            L.expect(';');
            got = true;
            continue;
        }
        if (key == "boundaryField" && L.peek() == '{') {
            size_t b = L.i;
            L.next();
            Dict tmp;
            parse_dict(L, tmp, true);
            F.boundary_text = L.s.substr(b, L.i - b);
            continue;
        }
        // any other entry: skip it (dimensions kept for the writer)
        L.i = start;
        Dict one;
        {
            std::string k2 = L.next();
            Entry e;
            e.key = k2;
            if (L.peek() == '{') {
                L.next();
                e.dict = std::make_unique<Dict>();
                parse_dict(L, *e.dict, true);
            } else {
                int depth = 0;
                std::string val;
                for (;;) {
                    std::string t = L.next();
                    if (t == "(" || t == "[") depth++;
                    if (t == ")" || t == "]") depth--;
                    if (t == ";" && depth == 0) break;
                    val += (val.empty() || t == "]" || val.back() == '[' ? "" : " ") + t;
                }
                if (k2 == "dimensions") F.dimensions = val;
            }
        }
    }
    if (!got) perr(path, "no internalField");
    return F;
}

// OpenFOAM's "general" time name (timePrecision 6 by default)
std::string time_name(double t) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.12g", t);
    return buf;
}

}  // namespace

// ---------------------------------------------------------------------------
// the case
// ---------------------------------------------------------------------------
struct lfm_case {
    std::string root, dir;  // case root; mesh/field directory (root or processorN)
    int rank = -1;
    // mesh
    std::vector<double> points;
    std::vector<int32_t> face_off, face_pts, owner, neighbour;
    int32_t n_cells = 0;
    std::vector<std::string> patch_name, patch_type;
    std::vector<int32_t> patch_start, patch_size, patch_role, patch_nbr;
    std::vector<int32_t> cell_submesh, face_key;
    // fields
    std::vector<double> p, T, U, alpha;
    bool have_alpha = false;
    Field fp, fT, fU, falpha;
    // controls
    lfm_params prm{};
    double start = 0, end = 0, dt = 0, write_interval = 1;
    int write_control = 0;
    double R = 0, gamma = 0;
};

static void read_mesh(lfm_case &c) {
    const std::string pm = c.dir + "/constant/polyMesh/";
    {
        ListFile lf;
        open_list(pm + "points", lf);
        Lexer &L = *lf.L;
        long n = L.count(L.binary ? 3 * (size_t)L.scalar_bytes : 7);  // "(x y z)"
        c.points.resize((size_t)n * 3);
        if (L.binary) {
            const char *p = L.raw((size_t)n * 3 * L.scalar_bytes, '(', ')');
            for (size_t k = 0; k < (size_t)n * 3; k++) c.points[k] = L.scalar_at(p, k);
        } else {
            L.expect('(');
            for (long i = 0; i < n; i++) {
                L.expect('(');
                for (int k = 0; k < 3; k++) c.points[(size_t)i * 3 + k] = L.num();
                L.expect(')');
            }
            L.expect(')');
        }
    }
    {
        ListFile lf;
        open_list(pm + "faces", lf);
        Lexer &L = *lf.L;
        // binary faces are always written compact (offsets, then labels)
        if (header_word(lf.header, "class") == "faceCompactList" || L.binary) {
            std::vector<int32_t> off = read_labels(L);
            std::vector<int32_t> pts = read_labels(L);
            if (off.empty() || off[0] != 0 || (size_t)off.back() != pts.size())
                perr(pm + "faces", "face offsets do not span the point labels");
            for (size_t k = 1; k < off.size(); k++)
                if (off[k] < off[k - 1]) perr(pm + "faces", "face offsets decrease");
            c.face_off = off;
            c.face_pts = pts;
        } else {
            long n = L.count(5);  // "k(...)"
            L.expect('(');
            c.face_off.assign(1, 0);
            c.face_off.reserve((size_t)n + 1);
            for (long i = 0; i < n; i++) {
                long k = L.count(2);
                L.expect('(');
                for (long j = 0; j < k; j++) c.face_pts.push_back((int32_t)L.lab());
                L.expect(')');
                c.face_off.push_back((int32_t)c.face_pts.size());
            }
            L.expect(')');
        }
    }
    c.owner = read_label_file(pm + "owner");
    c.neighbour = read_label_file(pm + "neighbour");
    const long nf = (long)c.face_off.size() - 1;
    if ((long)c.owner.size() != nf) perr(pm + "owner", "size differs from faces");
    {
        // the cell count is in the owner header's note ("nCells: N ..."), else max index + 1
        int32_t m = -1;
        for (int32_t o : c.owner) m = std::max(m, o);
        for (int32_t o : c.neighbour) m = std::max(m, o);
        // every cell owns or neighbours at least one face
        if ((long)m >= nf) perr(pm + "owner", "cell label " + std::to_string(m) + " beyond the face count");
        c.n_cells = m + 1;
    }
    {
        ListFile lf;
        open_list(pm + "boundary", lf);
        Lexer &L = *lf.L;
        long n = L.count(3);  // "p{}"
        L.expect('(');
        for (long i = 0; i < n; i++) {
            std::string name = L.next();
            L.expect('{');
            Dict d;
            parse_dict(L, d, true);
            DictFile tmp;
            tmp.path = pm + "boundary";
            tmp.root = std::move(d);
            const std::string type = dict_word(tmp, "", "type", "", false);
            c.patch_name.push_back(name);
            c.patch_type.push_back(type);
            c.patch_start.push_back((int32_t)dict_scalar(tmp, "", "startFace", 0, false));
            c.patch_size.push_back((int32_t)dict_scalar(tmp, "", "nFaces", 0, false));
            int32_t nbr = -1;
            if (type == "processor") nbr = (int32_t)dict_scalar(tmp, "", "neighbProcNo", 0, false);
            c.patch_role.push_back(lfm_patch_role(type.c_str(), name.c_str()));
            // cyclic partner resolved below by name
            c.patch_nbr.push_back(nbr);
            if (type == "cyclic") {
                std::string pn = dict_word(tmp, "", "neighbourPatch", "", true);
                c.patch_name.back() = name;
                c.patch_type.back() = type + ":" + pn;
            }
        }
        L.expect(')');
        for (size_t i = 0; i < c.patch_type.size(); i++) {
            const std::string &t = c.patch_type[i];
            if (t.rfind("cyclic:", 0) == 0) {
                const std::string pn = t.substr(7);
                for (size_t j = 0; j < c.patch_name.size(); j++)
                    if (c.patch_name[j] == pn) c.patch_nbr[i] = (int32_t)j;
                c.patch_type[i] = "cyclic";
                if (c.patch_nbr[i] < 0) perr(pm + "boundary", "cyclic patch " + c.patch_name[i] + " without neighbourPatch");
            }
        }
    }
    struct stat st;
    if (stat((pm + "cellSubmesh").c_str(), &st) == 0) {
        c.cell_submesh = read_label_file(pm + "cellSubmesh");
        if ((int32_t)c.cell_submesh.size() != c.n_cells) perr(pm + "cellSubmesh", "size differs from the cell count");
    }
    const int32_t nb = (int32_t)(nf - (long)c.neighbour.size());
    c.face_key.assign((size_t)std::max(nb, 0), -1);
    if (c.rank >= 0) {
        std::vector<int32_t> fpa = read_label_file(pm + "faceProcAddressing");
        if ((long)fpa.size() != nf) perr(pm + "faceProcAddressing", "size differs from faces");
        // faceProcAddressing: +-(global face + 1), sign = orientation
        for (int32_t f = (int32_t)c.neighbour.size(); f < (int32_t)nf; f++)
            c.face_key[f - c.neighbour.size()] = std::abs(fpa[f]) - 1;
    }
}

// the smallest (latest = false) or largest numeric time directory of a case
// directory; "constant" and other names are not times
static double time_dir_extreme(const std::string &dir, bool latest) {
    DIR *d = opendir(dir.c_str());
    if (!d) perr(dir, "cannot list");
    bool any = false;
    double best = 0.0;
    while (dirent *e = readdir(d)) {
        const char *n = e->d_name;
        char *end = nullptr;
        const double v = std::strtod(n, &end);
        if (end == n || *end || !std::isfinite(v) || !is_dir(dir + "/" + n)) continue;
        if (!any || (latest ? v > best : v < best)) best = v;
        any = true;
    }
    closedir(d);
    if (!any) perr(dir, "no time directory to start from");
    return best;
}

static void read_controls(lfm_case &c) {
    const std::string r = c.root + "/";
    DictFile control = load_dict(r + "system/controlDict");
    DictFile schemes = load_dict(r + "system/fvSchemes");
    DictFile sponge = load_dict(r + "constant/spongeDict");
    DictFile thermo = load_dict(r + "constant/thermophysicalProperties");
    DictFile turb = load_dict(r + "constant/turbulenceProperties");
    lfm_params &p = c.prm;
    std::memset(&p, 0, sizeof p);
    // CInputReader::initialize, src/inputReader.cpp:13-64
    p.have_res = dict_bool(control, "lfm/post", "haveResiduals", false, true);
    p.max_co = dict_scalar(control, "", "maxCo", 1, true);
    const bool dbl = dict_bool(control, "lfm", "doublePrecision", true, true);
    p.precision = dbl ? LFM_PREC_FP64 : LFM_PREC_FP32P;
    p.solver = (int32_t)dict_scalar(schemes, "lfm", "solver", 0, true);
    p.dim = (int32_t)dict_scalar(schemes, "lfm", "dimension", 2, true);
    p.rk_order = (int32_t)dict_scalar(schemes, "lfm", "rkOrder", 5, true);
    p.minmod = dict_bool(schemes, "lfm", "minmodExists", false, true);
    p.p_inf = dict_scalar(sponge, "", "pinf", 0, false);
    p.T_inf = dict_scalar(sponge, "", "Tinf", 0, false);
    p.U_inf[0] = dict_scalar(sponge, "", "Uinf_x", 0, true);
    p.U_inf[1] = dict_scalar(sponge, "", "Uinf_y", 0, true);
    p.U_inf[2] = dict_scalar(sponge, "", "Uinf_z", 0, true);
    p.Ls = dict_scalar(sponge, "", "Ls", 0, false);
    p.mach = dict_scalar(sponge, "", "M", 0, false);
    p.k_sponge = dict_scalar(sponge, "", "k", 4.34294481903252e-01, true);
    p.cp = dict_scalar(thermo, "mixture/thermodynamics", "Cp", 0, false);
    p.mol_weight = dict_scalar(thermo, "mixture/specie", "molWeight", 0, false);
    p.mu = dict_scalar(thermo, "mixture/transport", "mu", 0, false);
    p.pr = dict_scalar(thermo, "mixture/transport", "Pr", 0, false);
    p.laminar = dict_word(turb, "", "simulationType", "", false) == "laminar";
    // Foam::Time controls (runTimeManagerOF.cpp): startFrom picks the start
    // time as Foam::Time does -- the startTime entry (the default), or the
    // first / latest time directory of the case (a restart)
    const std::string from = dict_word(control, "", "startFrom", "startTime", true);
    if (from == "startTime")
        c.start = dict_scalar(control, "", "startTime", 0, true);
    else if (from == "firstTime" || from == "latestTime")
        c.start = time_dir_extreme(c.dir, from == "latestTime");
    else
        perr(r + "system/controlDict", "startFrom must be startTime, firstTime or latestTime, not " + from);
    c.end = dict_scalar(control, "", "endTime", 0, false);
    c.dt = dict_scalar(control, "", "deltaT", 0, false);
    p.dt = c.dt;
    p.adjust_dt = dict_bool(control, "", "adjustTimeStep", false, true);
    c.write_interval = dict_scalar(control, "", "writeInterval", 1, true);
    const std::string wc = dict_word(control, "", "writeControl", "timeStep", true);
    c.write_control = wc == "timeStep" ? 0 : 1;
    c.R = 8.31447 / p.mol_weight * 1000;
    c.gamma = p.cp / (p.cp - c.R);
}

static std::string start_dir(const lfm_case &c) {
    // the directory whose numeric name equals startTime
    DIR *d = opendir(c.dir.c_str());
    if (!d) perr(c.dir, "cannot list");
    std::string best;
    while (dirent *e = readdir(d)) {
        const char *n = e->d_name;
        char *end = nullptr;
        double v = std::strtod(n, &end);
        if (end == n || *end) continue;
        if (std::fabs(v - c.start) <= 1e-12 * std::max(1.0, std::fabs(c.start)) && is_dir(c.dir + "/" + n)) best = n;
    }
    closedir(d);
    if (best.empty()) perr(c.dir, "no time directory for startTime " + time_name(c.start));
    return c.dir + "/" + best;
}

static void read_fields(lfm_case &c) {
    const std::string t = start_dir(c) + "/";
    const long n = c.n_cells;
    c.fp = read_field(t + "p", n, 1);
    c.fT = read_field(t + "T", n, 1);
    c.fU = read_field(t + "U", n, 3);
    c.p = c.fp.v;
    c.T = c.fT.v;
    c.U = c.fU.v;
    struct stat st;
    c.have_alpha = stat((t + "alpha").c_str(), &st) == 0;
    if (c.have_alpha) {
        c.falpha = read_field(t + "alpha", n, 1);
        c.alpha = c.falpha.v;
    }
}

static void write_field(const std::string &path, const char *cls, const char *name, const std::string &dims,
                        const std::vector<double> &v, int ncomp, const std::string &boundary) {
    std::FILE *f = std::fopen(path.c_str(), "w");
    if (!f) perr(path, "cannot write");
    const size_t n = v.size() / ncomp;
    std::fprintf(f, "FoamFile\n{\n    version     2.0;\n    format      ascii;\n    class       %s;\n"
                    "    object      %s;\n}\n\n",
                 cls, name);
    if (!dims.empty()) std::fprintf(f, "dimensions      %s;\n\n", dims.c_str());
    std::fprintf(f, "internalField   nonuniform List<%s> \n%zu\n(\n", ncomp == 1 ? "scalar" : "vector", n);
    for (size_t i = 0; i < n; i++) {
        if (ncomp == 1)
            std::fprintf(f, "%.17g\n", v[i]);
        else
            std::fprintf(f, "(%.17g %.17g %.17g)\n", v[3 * i], v[3 * i + 1], v[3 * i + 2]);
    }
    std::fprintf(f, ")\n;\n\nboundaryField\n%s\n", boundary.empty() ? "{\n}" : boundary.c_str());
    std::fclose(f);
}

// boundaryField of a derived field: calculated on every patch (empty and
// coupled patches keep their type)
static std::string calculated_boundary(const lfm_case &c) {
    std::string bnd = "{\n";
    for (size_t i = 0; i < c.patch_name.size(); i++) {
        const std::string &t = c.patch_type[i];
        const bool keep = t == "empty" || t == "cyclic" || t == "processor";
        bnd += "    " + c.patch_name[i] + "\n    {\n        type            " + (keep ? t : "calculated") +
               ";\n" + (keep ? "" : "        value           uniform 0;\n") + "    }\n";
    }
    return bnd + "}";
}

extern "C" {

int lfm_case_open(const char *case_dir, int32_t rank, lfm_case **out) {
    if (!case_dir || !out) return fail(LFM_EINVAL, "null argument");
    *out = nullptr;
    auto c = std::make_unique<lfm_case>();
    c->root = case_dir;
    c->rank = rank;
    c->dir = rank < 0 ? c->root : c->root + "/processor" + std::to_string(rank);
    if (!is_dir(c->dir)) return fail(LFM_EINVAL, c->dir + ": no such case directory");
    try {
        read_controls(*c);
        read_mesh(*c);
        read_fields(*c);
    } catch (const ParseError &e) {
        return fail(LFM_EINVAL, e.msg);
    } catch (const std::bad_alloc &) {
        return fail(LFM_ENOMEM, "out of host memory reading the case");
    } catch (const std::exception &e) {
        return fail(LFM_EINVAL, c->dir + ": " + e.what());
    }
    *out = c.release();
    return LFM_OK;
}

int lfm_case_close(lfm_case *c) {
    delete c;
    return LFM_OK;
}

int lfm_case_mesh(lfm_case *c, lfm_polymesh *m, lfm_partition *part) {
    if (!c || !m) return fail(LFM_EINVAL, "null argument");
    std::memset(m, 0, sizeof *m);
    m->n_points = (int32_t)(c->points.size() / 3);
    m->points = c->points.data();
    m->n_faces = (int32_t)c->owner.size();
    m->face_off = c->face_off.data();
    m->face_pts = c->face_pts.data();
    m->owner = c->owner.data();
    m->n_internal = (int32_t)c->neighbour.size();
    m->neighbour = c->neighbour.data();
    m->n_cells = c->n_cells;
    m->n_patches = (int32_t)c->patch_start.size();
    m->patch_start = c->patch_start.data();
    m->patch_size = c->patch_size.data();
    m->patch_role = c->patch_role.data();
    m->patch_nbr = c->patch_nbr.data();
    m->cell_submesh = c->cell_submesh.empty() ? nullptr : c->cell_submesh.data();
    if (part) {
        part->rank = c->rank;
        part->face_key = c->rank >= 0 ? c->face_key.data() : nullptr;
    }
    return LFM_OK;
}

int lfm_case_params(lfm_case *c, lfm_params *prm) {
    if (!c || !prm) return fail(LFM_EINVAL, "null argument");
    *prm = c->prm;
    return LFM_OK;
}

int lfm_case_fields(lfm_case *c, const double **p, const double **T, const double **U3, const double **alpha) {
    if (!c || !p || !T || !U3 || !alpha) return fail(LFM_EINVAL, "null argument");
    *p = c->p.data();
    *T = c->T.data();
    *U3 = c->U.data();
    *alpha = c->have_alpha ? c->alpha.data() : nullptr;
    return LFM_OK;
}

int lfm_case_time(lfm_case *c, double *start, double *end, double *delta_t, double *write_interval,
                  int32_t *write_control) {
    if (!c) return fail(LFM_EINVAL, "null argument");
    if (start) *start = c->start;
    if (end) *end = c->end;
    if (delta_t) *delta_t = c->dt;
    if (write_interval) *write_interval = c->write_interval;
    if (write_control) *write_control = c->write_control;
    return LFM_OK;
}

int lfm_case_dict_scalar(lfm_case *c, const char *dict, const char *subdict, const char *key, double dflt,
                         int32_t allow_default, double *out) {
    if (!c || !dict || !key || !out) return fail(LFM_EINVAL, "null argument");
    try {
        DictFile f = load_dict(c->root + "/" + dict);
        *out = dict_scalar(f, subdict ? subdict : "", key, dflt, allow_default != 0);
    } catch (const ParseError &e) {
        return fail(LFM_EINVAL, e.msg);
    } catch (const std::exception &e) {
        return fail(LFM_EINVAL, e.what());
    }
    return LFM_OK;
}

int lfm_case_post(lfm_case *c, lfm_post *post) {
    if (!c || !post) return fail(LFM_EINVAL, "null argument");
    std::memset(post, 0, sizeof *post);
    try {
        DictFile control = load_dict(c->root + "/system/controlDict");
        // CInputReader::initialize, src/inputReader.cpp:19-23
        post->have_average = dict_bool(control, "lfm/post", "haveAverage", false, true);
        const double t_avg = dict_scalar(control, "lfm/post", "tStartAverage", 0, true);
        post->have_forces = dict_bool(control, "lfm/post", "haveForces", false, true);
        post->forces_every = (int32_t)dict_scalar(control, "lfm/post", "saveForcesStep", 1, true);
        // Mesh::solve, mesh_solver.cpp:249-254: int indexStartAverage =
        // (tStartAverage - time) / dt, truncated toward zero
        post->average_start = (int32_t)((t_avg - c->start) / c->dt);
    } catch (const ParseError &e) {
        return fail(LFM_EINVAL, e.msg);
    } catch (const std::exception &e) {
        return fail(LFM_EINVAL, e.what());
    }
    post->time = c->start;
    return LFM_OK;
}

int lfm_case_write_scalar(lfm_case *c, double time, const char *name, const char *dimensions, const double *v) {
    if (!c || !name || !v) return fail(LFM_EINVAL, "null argument");
    const std::string d = c->dir + "/" + time_name(time);
    if (!is_dir(d) && mkdir(d.c_str(), 0755) != 0) return fail(LFM_EINVAL, d + ": " + std::strerror(errno));
    try {
        std::vector<double> vals(v, v + c->n_cells);
        write_field(d + "/" + name, "volScalarField", name, dimensions ? dimensions : "", vals, 1,
                    calculated_boundary(*c));
    } catch (const ParseError &e) {
        return fail(LFM_EINVAL, e.msg);
    } catch (const std::exception &e) {
        return fail(LFM_EINVAL, e.what());
    }
    return LFM_OK;
}

int lfm_case_write_time(lfm_case *c, double time, const double *q, int32_t dim) {
    if (!c || !q || (dim != 2 && dim != 3)) return fail(LFM_EINVAL, "bad argument");
    const size_t n = (size_t)c->n_cells;
    std::vector<double> rho(n), E(n), p(n), T(n), U(3 * n, 0.0);
    const double gm1 = c->gamma - 1.0;
    // updateSolutionPrimitives, cfd_v0.cpp:3000-3031; T = p / (rho R) is
    // written as well (the reference leaves T at its initial field)
    for (size_t i = 0; i < n; i++) {
        const double r = q[i];
        double u2 = 0.0;
        for (int k = 0; k < dim; k++) {
            U[3 * i + k] = q[(size_t)(k + 1) * n + i] / r;
            u2 += U[3 * i + k] * U[3 * i + k];
        }
        const double e = q[(size_t)(dim + 1) * n + i] / r;
        rho[i] = r;
        E[i] = e;
        p[i] = r * gm1 * (e - 0.5 * u2);
        T[i] = p[i] / (r * c->R);
    }
    const std::string d = c->dir + "/" + time_name(time);
    if (!is_dir(d) && mkdir(d.c_str(), 0755) != 0) return fail(LFM_EINVAL, d + ": " + std::strerror(errno));
    try {
        write_field(d + "/p", "volScalarField", "p", c->fp.dimensions, p, 1, boundary_ascii(c->fp));
        write_field(d + "/T", "volScalarField", "T", c->fT.dimensions, T, 1, boundary_ascii(c->fT));
        write_field(d + "/U", "volVectorField", "U", c->fU.dimensions, U, 3, boundary_ascii(c->fU));
        const std::string bnd = calculated_boundary(*c);
        write_field(d + "/rho", "volScalarField", "rho", "[1 -3 0 0 0 0 0]", rho, 1, bnd);
        write_field(d + "/E", "volScalarField", "E", "[0 2 -2 0 0 0 0]", E, 1, bnd);
        // the sponge distance is read with AUTO_WRITE and so lands in every
        // written time directory (polyMeshReaderOF.cpp:700-709), where a
        // restart from that time reads it again
        if (c->have_alpha)
            write_field(d + "/alpha", "volScalarField", "alpha", c->falpha.dimensions, c->alpha, 1,
                        boundary_ascii(c->falpha));
    } catch (const ParseError &e) {
        return fail(LFM_EINVAL, e.msg);
    } catch (const std::exception &e) {
        return fail(LFM_EINVAL, e.what());
    }
    return LFM_OK;
}

}  // extern "C"
