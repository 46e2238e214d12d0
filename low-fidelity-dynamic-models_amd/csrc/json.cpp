#include "json.h"

#include <cctype>
#include <charconv>
#include <fstream>
#include <sstream>

namespace lfdm {

const JValue* JValue::find(const std::string& key) const {
    if (kind != Object) return nullptr;
    for (const auto& kv : obj)
        if (kv.first == key) return &kv.second;
    return nullptr;
}

namespace {

// Nesting deeper than this is refused (the reference's parameter files nest three levels).
constexpr int kMaxDepth = 64;

struct Parser {
    const std::string& s;
    size_t i = 0;
    std::string err;
    int depth = 0;

    void ws() {
        while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
    }
    bool fail(const char* what) {
        if (err.empty()) err = std::string(what) + " at byte " + std::to_string(i);
        return false;
    }
    bool lit(const char* w) {
        size_t n = 0;
        while (w[n]) ++n;
        if (s.compare(i, n, w) != 0) return fail("bad literal");
        i += n;
        return true;
    }
    bool string(std::string& out) {
        if (s[i] != '"') return fail("expected string");
        ++i;
        while (i < s.size() && s[i] != '"') {
            if (s[i] == '\\') {
                ++i;
                if (i >= s.size()) return fail("bad escape");
                char c = s[i];
                out.push_back(c == 'n' ? '\n' : c == 't' ? '\t' : c == 'r' ? '\r' : c);
            } else {
                out.push_back(s[i]);
            }
            ++i;
        }
        if (i >= s.size()) return fail("unterminated string");
        ++i;
        return true;
    }
    bool value(JValue& v) {
        if (++depth > kMaxDepth) return fail("nesting too deep");
        const bool ok = value_at(v);
        --depth;
        return ok;
    }
    bool value_at(JValue& v) {
        ws();
        if (i >= s.size()) return fail("unexpected end");
        char c = s[i];
        if (c == '{') {
            v.kind = JValue::Object;
            ++i;
            ws();
            if (i < s.size() && s[i] == '}') { ++i; return true; }
            while (true) {
                ws();
                std::string key;
                if (i >= s.size() || !string(key)) return fail("expected key");
                ws();
                if (i >= s.size() || s[i] != ':') return fail("expected ':'");
                ++i;
                v.obj.emplace_back(key, JValue());
                if (!value(v.obj.back().second)) return false;
                ws();
                if (i < s.size() && s[i] == ',') { ++i; continue; }
                if (i < s.size() && s[i] == '}') { ++i; return true; }
                return fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.kind = JValue::Array;
            ++i;
            ws();
            if (i < s.size() && s[i] == ']') { ++i; return true; }
            while (true) {
                v.arr.emplace_back();
                if (!value(v.arr.back())) return false;
                ws();
                if (i < s.size() && s[i] == ',') { ++i; continue; }
                if (i < s.size() && s[i] == ']') { ++i; return true; }
                return fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.kind = JValue::String;
            return string(v.str);
        }
        if (c == 't') { v.kind = JValue::Bool; v.b = true; return lit("true"); }
        if (c == 'f') { v.kind = JValue::Bool; v.b = false; return lit("false"); }
        if (c == 'n') { v.kind = JValue::Null; return lit("null"); }
        if (c == '-' || (c >= '0' && c <= '9')) {
            // std::from_chars: correctly rounded and independent of the process's C locale (strtod
            // would read "0.5" as 0 under a comma-decimal LC_NUMERIC)
            const char* p = s.data() + i;
            v.kind = JValue::Number;
            const auto r = std::from_chars(p, s.data() + s.size(), v.num);
            if (r.ec != std::errc() || r.ptr == p) return fail("bad number");
            i += (size_t)(r.ptr - p);
            return true;
        }
        return fail("unexpected character");
    }
};

}  // namespace

bool json_parse(const std::string& text, JValue& out, std::string& err) {
    Parser p{text};
    out = JValue();
    if (!p.value(out)) {
        err = p.err;
        return false;
    }
    p.ws();
    if (p.i != text.size()) {
        err = "trailing characters at byte " + std::to_string(p.i);
        return false;
    }
    return true;
}

bool json_load_file(const std::string& path, JValue& out, std::string& err) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    if (!json_parse(ss.str(), out, err)) {
        err = path + ": JSON parse error: " + err;
        return false;
    }
    return true;
}

}  // namespace lfdm
