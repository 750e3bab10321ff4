// Recursive-descent JSON reader (RFC 8259 subset sufficient for sidecars:
// objects, arrays, strings with escapes, numbers, true/false/null).
#include "json.h"

#include <cstdlib>
#include <cstring>

namespace chd::json {

const Value *Value::get(const char *key) const
{
    if (type != Object) return nullptr;
    for (const auto &kv : o)
        if (kv.first == key) return &kv.second;
    return nullptr;
}

namespace {

struct Parser {
    const char *p;
    const char *end;
    std::string err;
    int depth = 0;

    void ws()
    {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    }
    bool fail(const char *m)
    {
        if (err.empty()) err = m;
        return false;
    }
    bool lit(const char *w)
    {
        const size_t n = std::strlen(w);
        if ((size_t)(end - p) < n || std::strncmp(p, w, n) != 0) return fail("invalid literal");
        p += n;
        return true;
    }
    static void put_utf8(std::string &s, unsigned cp)
    {
        if (cp < 0x80) s += (char)cp;
        else if (cp < 0x800) { s += (char)(0xC0 | (cp >> 6)); s += (char)(0x80 | (cp & 0x3F)); }
        else if (cp < 0x10000) {
            s += (char)(0xE0 | (cp >> 12)); s += (char)(0x80 | ((cp >> 6) & 0x3F)); s += (char)(0x80 | (cp & 0x3F));
        } else {
            s += (char)(0xF0 | (cp >> 18)); s += (char)(0x80 | ((cp >> 12) & 0x3F));
            s += (char)(0x80 | ((cp >> 6) & 0x3F)); s += (char)(0x80 | (cp & 0x3F));
        }
    }
    bool hex4(unsigned &v)
    {
        if (end - p < 4) return fail("short \\u escape");
        v = 0;
        for (int i = 0; i < 4; ++i) {
            const char c = *p++;
            v <<= 4;
            if (c >= '0' && c <= '9') v |= c - '0';
            else if (c >= 'a' && c <= 'f') v |= c - 'a' + 10;
            else if (c >= 'A' && c <= 'F') v |= c - 'A' + 10;
            else return fail("bad \\u escape");
        }
        return true;
    }
    bool str(std::string &s)
    {
        if (p >= end || *p != '"') return fail("expected string");
        ++p;
        while (p < end && *p != '"') {
            if (*p == '\\') {
                ++p;
                if (p >= end) return fail("unterminated escape");
                const char c = *p++;
                switch (c) {
                case '"': s += '"'; break;
                case '\\': s += '\\'; break;
                case '/': s += '/'; break;
                case 'b': s += '\b'; break;
                case 'f': s += '\f'; break;
                case 'n': s += '\n'; break;
                case 'r': s += '\r'; break;
                case 't': s += '\t'; break;
                case 'u': {
                    unsigned cp;
                    if (!hex4(cp)) return false;
                    if (cp >= 0xD800 && cp < 0xDC00 && end - p >= 6 && p[0] == '\\' && p[1] == 'u') {
                        p += 2;
                        unsigned lo;
                        if (!hex4(lo)) return false;
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    put_utf8(s, cp);
                    break;
                }
                default: return fail("bad escape");
                }
            } else {
                s += *p++;
            }
        }
        if (p >= end) return fail("unterminated string");
        ++p;
        return true;
    }
    bool value(Value &v)
    {
        if (++depth > 256) return fail("nesting too deep");
        ws();
        if (p >= end) return fail("unexpected end");
        bool ok = true;
        switch (*p) {
        case '{': {
            ++p;
            v.type = Value::Object;
            ws();
            if (p < end && *p == '}') { ++p; break; }
            for (;;) {
                ws();
                std::string k;
                if (!str(k)) return false;
                ws();
                if (p >= end || *p != ':') return fail("expected ':'");
                ++p;
                v.o.emplace_back(std::move(k), Value());
                if (!value(v.o.back().second)) return false;
                ws();
                if (p < end && *p == ',') { ++p; continue; }
                if (p < end && *p == '}') { ++p; break; }
                return fail("expected ',' or '}'");
            }
            break;
        }
        case '[': {
            ++p;
            v.type = Value::Array;
            ws();
            if (p < end && *p == ']') { ++p; break; }
            for (;;) {
                v.a.emplace_back();
                if (!value(v.a.back())) return false;
                ws();
                if (p < end && *p == ',') { ++p; continue; }
                if (p < end && *p == ']') { ++p; break; }
                return fail("expected ',' or ']'");
            }
            break;
        }
        case '"':
            v.type = Value::String;
            ok = str(v.s);
            break;
        case 't': v.type = Value::Bool; v.b = true; ok = lit("true"); break;
        case 'f': v.type = Value::Bool; v.b = false; ok = lit("false"); break;
        case 'n': v.type = Value::Null; ok = lit("null"); break;
        default: {
            const char *q = p;
            if (q < end && (*q == '-' || *q == '+')) ++q;
            while (q < end && ((*q >= '0' && *q <= '9') || *q == '.' || *q == 'e' || *q == 'E' || *q == '-' || *q == '+')) ++q;
            if (q == p) return fail("unexpected character");
            std::string num(p, q);
            char *e2 = nullptr;
            v.type = Value::Number;
            v.n = std::strtod(num.c_str(), &e2);
            if (e2 == nullptr || *e2 != '\0') return fail("bad number");
            p = q;
        }
        }
        --depth;
        return ok;
    }
};

}  // namespace

bool parse(const std::string &text, Value &out, std::string &err)
{
    Parser ps{text.data(), text.data() + text.size(), std::string()};
    out = Value();
    if (!ps.value(out)) {
        err = ps.err;
        return false;
    }
    ps.ws();
    if (ps.p != ps.end) {
        err = "trailing characters after the document";
        return false;
    }
    return true;
}

}  // namespace chd::json
