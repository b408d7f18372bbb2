// Recursive-descent JSON reader (RFC 8259): objects, arrays, strings with
// escapes (\uXXXX encoded as UTF-8, surrogate pairs joined), numbers via
// strtod, literals.  Depth is bounded so hostile input cannot exhaust the
// stack.
#include "json_lite.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace lenia {
namespace json {

namespace {

constexpr int kMaxDepth = 256;

struct Reader {
    const char *p, *end, *begin;
    std::string err;

    bool fail(const char *what) {
        if (err.empty()) err = std::string(what) + " at byte " + std::to_string(p - begin);
        return false;
    }
    void ws() {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    }
    bool lit(const char *s, Value &v, Value::Kind k, bool b) {
        size_t n = std::strlen(s);
        if ((size_t)(end - p) < n || std::memcmp(p, s, n) != 0) return fail("invalid literal");
        p += n;
        v.kind = k;
        v.b = b;
        return true;
    }
    static void utf8(std::string &o, uint32_t c) {
        if (c < 0x80) {
            o += (char)c;
        } else if (c < 0x800) {
            o += (char)(0xC0 | (c >> 6));
            o += (char)(0x80 | (c & 0x3F));
        } else if (c < 0x10000) {
            o += (char)(0xE0 | (c >> 12));
            o += (char)(0x80 | ((c >> 6) & 0x3F));
            o += (char)(0x80 | (c & 0x3F));
        } else {
            o += (char)(0xF0 | (c >> 18));
            o += (char)(0x80 | ((c >> 12) & 0x3F));
            o += (char)(0x80 | ((c >> 6) & 0x3F));
            o += (char)(0x80 | (c & 0x3F));
        }
    }
    bool hex4(uint32_t &c) {
        if (end - p < 4) return fail("truncated \\u escape");
        c = 0;
        for (int i = 0; i < 4; ++i) {
            char h = *p++;
            c <<= 4;
            if (h >= '0' && h <= '9') c |= (uint32_t)(h - '0');
            else if (h >= 'a' && h <= 'f') c |= (uint32_t)(h - 'a' + 10);
            else if (h >= 'A' && h <= 'F') c |= (uint32_t)(h - 'A' + 10);
            else return fail("bad hex digit in \\u escape");
        }
        return true;
    }
    bool string(std::string &o) {
        ++p;  // opening quote
        while (true) {
            if (p >= end) return fail("unterminated string");
            char c = *p++;
            if (c == '"') return true;
            if ((unsigned char)c < 0x20) return fail("control character in string");
            if (c != '\\') {
                o += c;
                continue;
            }
            if (p >= end) return fail("unterminated escape");
            char e = *p++;
            switch (e) {
                case '"': o += '"'; break;
                case '\\': o += '\\'; break;
                case '/': o += '/'; break;
                case 'b': o += '\b'; break;
                case 'f': o += '\f'; break;
                case 'n': o += '\n'; break;
                case 'r': o += '\r'; break;
                case 't': o += '\t'; break;
                case 'u': {
                    uint32_t c1;
                    if (!hex4(c1)) return false;
                    if (c1 >= 0xD800 && c1 < 0xDC00) {
                        uint32_t c2;
                        if (end - p < 2 || p[0] != '\\' || p[1] != 'u') return fail("lone surrogate");
                        p += 2;
                        if (!hex4(c2)) return false;
                        if (c2 < 0xDC00 || c2 >= 0xE000) return fail("bad surrogate pair");
                        c1 = 0x10000 + ((c1 - 0xD800) << 10) + (c2 - 0xDC00);
                    }
                    utf8(o, c1);
                    break;
                }
                default: return fail("bad escape");
            }
        }
    }
    bool number(Value &v) {
        const char *s = p;
        if (p < end && *p == '-') ++p;
        if (p >= end) return fail("bad number");
        if (*p == '0') {
            ++p;
        } else if (*p >= '1' && *p <= '9') {
            while (p < end && *p >= '0' && *p <= '9') ++p;
        } else {
            return fail("bad number");
        }
        if (p < end && *p == '.') {
            ++p;
            if (p >= end || *p < '0' || *p > '9') return fail("bad fraction");
            while (p < end && *p >= '0' && *p <= '9') ++p;
        }
        if (p < end && (*p == 'e' || *p == 'E')) {
            ++p;
            if (p < end && (*p == '+' || *p == '-')) ++p;
            if (p >= end || *p < '0' || *p > '9') return fail("bad exponent");
            while (p < end && *p >= '0' && *p <= '9') ++p;
        }
        std::string tok(s, p);
        v.kind = Value::Number;
        v.num = std::strtod(tok.c_str(), nullptr);
        return true;
    }
    bool value(Value &v, int depth) {
        if (depth > kMaxDepth) return fail("nesting too deep");
        ws();
        if (p >= end) return fail("unexpected end of input");
        switch (*p) {
            case '{': {
                ++p;
                v.kind = Value::Object;
                ws();
                if (p < end && *p == '}') {
                    ++p;
                    return true;
                }
                while (true) {
                    ws();
                    if (p >= end || *p != '"') return fail("expected object key");
                    std::string key;
                    if (!string(key)) return false;
                    ws();
                    if (p >= end || *p != ':') return fail("expected ':'");
                    ++p;
                    Value m;
                    if (!value(m, depth + 1)) return false;
                    v.obj.emplace_back(std::move(key), std::move(m));
                    ws();
                    if (p < end && *p == ',') {
                        ++p;
                        continue;
                    }
                    if (p < end && *p == '}') {
                        ++p;
                        return true;
                    }
                    return fail("expected ',' or '}'");
                }
            }
            case '[': {
                ++p;
                v.kind = Value::Array;
                ws();
                if (p < end && *p == ']') {
                    ++p;
                    return true;
                }
                while (true) {
                    Value e;
                    if (!value(e, depth + 1)) return false;
                    v.arr.push_back(std::move(e));
                    ws();
                    if (p < end && *p == ',') {
                        ++p;
                        continue;
                    }
                    if (p < end && *p == ']') {
                        ++p;
                        return true;
                    }
                    return fail("expected ',' or ']'");
                }
            }
            case '"': v.kind = Value::String; return string(v.str);
            case 't': return lit("true", v, Value::Bool, true);
            case 'f': return lit("false", v, Value::Bool, false);
            case 'n': return lit("null", v, Value::Null, false);
            default: return number(v);
        }
    }
};

}  // namespace

const Value *Value::get(const std::string &key) const {
    if (kind != Object) return nullptr;
    // the last duplicate wins, as in the reference's parser
    const Value *found = nullptr;
    for (const auto &m : obj)
        if (m.first == key) found = &m.second;
    return found;
}

bool parse(const char *text, size_t len, Value &out, std::string &err) {
    Reader r{text, text + len, text, {}};
    out = Value();
    if (!r.value(out, 0)) {
        err = r.err;
        return false;
    }
    r.ws();
    if (r.p != r.end) {
        r.fail("trailing characters");
        err = r.err;
        return false;
    }
    return true;
}

static void dump_string(const std::string &s, std::string &o) {
    o += '"';
    for (unsigned char c : s) {
        switch (c) {
            case '"': o += "\\\""; break;
            case '\\': o += "\\\\"; break;
            case '\n': o += "\\n"; break;
            case '\r': o += "\\r"; break;
            case '\t': o += "\\t"; break;
            default:
                if (c < 0x20) {
                    char b[8];
                    std::snprintf(b, sizeof(b), "\\u%04x", c);
                    o += b;
                } else {
                    o += (char)c;
                }
        }
    }
    o += '"';
}

static void dump_to(const Value &v, std::string &o) {
    switch (v.kind) {
        case Value::Null: o += "null"; break;
        case Value::Bool: o += v.b ? "true" : "false"; break;
        case Value::Number: {
            char b[32];
            if (v.num == (double)(long long)v.num && v.num > -1e15 && v.num < 1e15)
                std::snprintf(b, sizeof(b), "%lld", (long long)v.num);
            else
                std::snprintf(b, sizeof(b), "%.17g", v.num);
            o += b;
            break;
        }
        case Value::String: dump_string(v.str, o); break;
        case Value::Array:
            o += '[';
            for (size_t i = 0; i < v.arr.size(); ++i) {
                if (i) o += ',';
                dump_to(v.arr[i], o);
            }
            o += ']';
            break;
        case Value::Object:
            o += '{';
            for (size_t i = 0; i < v.obj.size(); ++i) {
                if (i) o += ',';
                dump_string(v.obj[i].first, o);
                o += ':';
                dump_to(v.obj[i].second, o);
            }
            o += '}';
            break;
    }
}

std::string dump(const Value &v) {
    std::string o;
    dump_to(v, o);
    return o;
}

bool get_number(const Value &o, const std::string &key, double dflt, double &out, std::string &err) {
    const Value *v = o.get(key);
    if (!v) {
        out = dflt;
        return true;
    }
    if (v->kind == Value::Number) {
        out = v->num;
        return true;
    }
    if (v->kind == Value::Bool) {  // nlohmann converts booleans to numbers
        out = v->b ? 1.0 : 0.0;
        return true;
    }
    err = "\"" + key + "\" must be a number";
    return false;
}

bool get_string(const Value &o, const std::string &key, const std::string &dflt, std::string &out,
                std::string &err) {
    const Value *v = o.get(key);
    if (!v) {
        out = dflt;
        return true;
    }
    if (v->kind == Value::String) {
        out = v->str;
        return true;
    }
    err = "\"" + key + "\" must be a string";
    return false;
}

bool get_bool(const Value &o, const std::string &key, bool dflt, bool &out, std::string &err) {
    const Value *v = o.get(key);
    if (!v) {
        out = dflt;
        return true;
    }
    if (v->kind == Value::Bool) {
        out = v->b;
        return true;
    }
    err = "\"" + key + "\" must be true or false";
    return false;
}

}  // namespace json
}  // namespace lenia
