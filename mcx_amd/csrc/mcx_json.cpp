/*
 * mcx_json.cpp -- recursive-descent JSON reader and pretty printer.
 */
#include "mcx_json.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mcx_internal.h"

namespace mcx {

namespace {

struct Parser {
    const char* s;
    const char* e;
    explicit Parser(const std::string& t) : s(t.data()), e(t.data() + t.size()) {}

    [[noreturn]] void fail(const char* what) {
        throw Error(-1, std::string("JSON parse error: ") + what);
    }
    void ws() {
        while (s < e && (*s == ' ' || *s == '\t' || *s == '\n' || *s == '\r')) s++;
    }
    bool lit(const char* w) {
        size_t n = strlen(w);
        if ((size_t)(e - s) >= n && memcmp(s, w, n) == 0) {
            s += n;
            return true;
        }
        return false;
    }
    static void utf8(std::string& out, unsigned cp) {
        if (cp < 0x80) {
            out += (char)cp;
        } else if (cp < 0x800) {
            out += (char)(0xC0 | (cp >> 6));
            out += (char)(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            out += (char)(0xE0 | (cp >> 12));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        } else {
            out += (char)(0xF0 | (cp >> 18));
            out += (char)(0x80 | ((cp >> 12) & 0x3F));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        }
    }
    std::string string() {
        if (s >= e || *s != '"') fail("expected string");
        s++;
        std::string out;
        while (s < e && *s != '"') {
            char c = *s++;
            if (c != '\\') {
                out += c;
                continue;
            }
            if (s >= e) fail("bad escape");
            char x = *s++;
            switch (x) {
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                case '/': out += '/'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'n': out += '\n'; break;
                case 'r': out += '\r'; break;
                case 't': out += '\t'; break;
                case 'u': {
                    if (e - s < 4) fail("bad \\u escape");
                    unsigned cp = (unsigned)strtoul(std::string(s, 4).c_str(), nullptr, 16);
                    s += 4;
                    if (cp >= 0xD800 && cp < 0xDC00 && e - s >= 6 && s[0] == '\\' && s[1] == 'u') {
                        unsigned lo = (unsigned)strtoul(std::string(s + 2, 4).c_str(), nullptr, 16);
                        s += 6;
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    utf8(out, cp);
                    break;
                }
                default: fail("bad escape");
            }
        }
        if (s >= e) fail("unterminated string");
        s++;
        return out;
    }
    Json value() {
        ws();
        if (s >= e) fail("unexpected end");
        Json j;
        char c = *s;
        if (c == '{') {
            s++;
            j.type = Json::Object;
            ws();
            if (s < e && *s == '}') {
                s++;
                return j;
            }
            while (true) {
                ws();
                std::string k = string();
                ws();
                if (s >= e || *s != ':') fail("expected ':'");
                s++;
                j.obj.emplace_back(k, value());
                ws();
                if (s < e && *s == ',') {
                    s++;
                    continue;
                }
                if (s < e && *s == '}') {
                    s++;
                    break;
                }
                fail("expected ',' or '}'");
            }
        } else if (c == '[') {
            s++;
            j.type = Json::Array;
            ws();
            if (s < e && *s == ']') {
                s++;
                return j;
            }
            while (true) {
                j.arr.push_back(value());
                ws();
                if (s < e && *s == ',') {
                    s++;
                    continue;
                }
                if (s < e && *s == ']') {
                    s++;
                    break;
                }
                fail("expected ',' or ']'");
            }
        } else if (c == '"') {
            j.type = Json::String;
            j.str = string();
        } else if (lit("true")) {
            j.type = Json::Bool;
            j.b = true;
        } else if (lit("false")) {
            j.type = Json::Bool;
            j.b = false;
        } else if (lit("null")) {
            j.type = Json::Null;
        } else {
            char* end = nullptr;
            std::string tmp(s, (size_t)std::min<ptrdiff_t>(e - s, 64));
            double v = strtod(tmp.c_str(), &end);
            if (end == tmp.c_str()) fail("unexpected token");
            s += (end - tmp.c_str());
            j.type = Json::Number;
            j.num = v;
        }
        return j;
    }
};

void esc(std::string& o, const std::string& s) {
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
                    snprintf(b, sizeof(b), "\\u%04x", c);
                    o += b;
                } else {
                    o += (char)c;
                }
        }
    }
    o += '"';
}

void num(std::string& o, double v) {
    char b[40];
    if (v == floor(v) && fabs(v) < 1e15)
        snprintf(b, sizeof(b), "%.0f", v);
    else
        snprintf(b, sizeof(b), "%.9g", v);
    o += b;
}

void dump_rec(const Json& j, std::string& o, int indent, int level) {
    auto nl = [&](int l) {
        if (indent > 0) {
            o += '\n';
            o.append((size_t)(l * indent), '\t');
        }
    };
    switch (j.type) {
        case Json::Null: o += "null"; break;
        case Json::Bool: o += j.b ? "true" : "false"; break;
        case Json::Number: num(o, j.num); break;
        case Json::String: esc(o, j.str); break;
        case Json::Array: {
            bool flat = true;
            for (const auto& x : j.arr)
                if (x.type == Json::Array || x.type == Json::Object) flat = false;
            o += '[';
            for (size_t i = 0; i < j.arr.size(); i++) {
                if (i) o += flat ? ", " : ",";
                if (!flat) nl(level + 1);
                dump_rec(j.arr[i], o, indent, level + 1);
            }
            if (!flat && !j.arr.empty()) nl(level);
            o += ']';
            break;
        }
        case Json::Object: {
            o += '{';
            for (size_t i = 0; i < j.obj.size(); i++) {
                if (i) o += ',';
                nl(level + 1);
                esc(o, j.obj[i].first);
                o += indent > 0 ? ":\t" : ":";
                dump_rec(j.obj[i].second, o, indent, level + 1);
            }
            if (!j.obj.empty()) nl(level);
            o += '}';
            break;
        }
    }
}

}  // namespace

Json Json::parse(const std::string& text) {
    Parser p(text);
    Json j = p.value();
    p.ws();
    if (p.s != p.e) p.fail("trailing characters");
    return j;
}

std::string Json::dump(int indent) const {
    std::string o;
    dump_rec(*this, o, indent, 0);
    return o;
}

const Json* Json::get(const char* key) const {
    if (type != Object) return nullptr;
    for (const auto& kv : obj)
        if (kv.first == key) return &kv.second;
    return nullptr;
}

}  // namespace mcx
