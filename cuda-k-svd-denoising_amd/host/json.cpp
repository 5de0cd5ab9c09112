#include "json.h"

#include <cctype>
#include <cstdlib>
#include <iterator>
#include <sstream>

namespace json {

namespace detail {
template <> bool is<double>(const Value &v) { return v.kind == Value::NUMBER; }
template <> bool is<bool>(const Value &v) { return v.kind == Value::BOOL; }
template <> bool is<std::string>(const Value &v) { return v.kind == Value::STRING; }
template <> bool is<Array>(const Value &v) { return v.kind == Value::ARRAY; }
template <> bool is<Object>(const Value &v) { return v.kind == Value::OBJECT; }
template <> double as<double>(const Value &v) { return v.num; }
template <> bool as<bool>(const Value &v) { return v.b; }
template <> std::string as<std::string>(const Value &v) { return v.str; }
template <> Array as<Array>(const Value &v) { return *v.arr; }
template <> Object as<Object>(const Value &v) { return *v.obj; }
}  // namespace detail

namespace {

struct Parser {
    const std::string &s;
    size_t i = 0;
    std::string err;

    explicit Parser(const std::string &text) : s(text) {}

    void ws()
    {
        while (i < s.size() && std::isspace((unsigned char)s[i])) i++;
    }
    bool fail(const std::string &m)
    {
        if (err.empty()) err = m + " at offset " + std::to_string(i);
        return false;
    }
    bool lit(const char *w)
    {
        size_t n = 0;
        while (w[n]) n++;
        if (s.compare(i, n, w) != 0) return false;
        i += n;
        return true;
    }
    bool string(std::string &out)
    {
        if (i >= s.size() || s[i] != '"') return fail("expected string");
        i++;
        out.clear();
        while (i < s.size() && s[i] != '"') {
            char c = s[i++];
            if (c == '\\') {
                if (i >= s.size()) return fail("bad escape");
                char e = s[i++];
                switch (e) {
                case 'n': out += '\n'; break;
                case 't': out += '\t'; break;
                case 'r': out += '\r'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'u': {
                    if (i + 4 > s.size()) return fail("bad \\u escape");
                    unsigned cp = (unsigned)strtoul(s.substr(i, 4).c_str(), nullptr, 16);
                    i += 4;
                    if (cp < 0x80) out += (char)cp;
                    else if (cp < 0x800) { out += (char)(0xC0 | (cp >> 6)); out += (char)(0x80 | (cp & 0x3F)); }
                    else {
                        out += (char)(0xE0 | (cp >> 12));
                        out += (char)(0x80 | ((cp >> 6) & 0x3F));
                        out += (char)(0x80 | (cp & 0x3F));
                    }
                    break;
                }
                default: out += e; break;
                }
            } else {
                out += c;
            }
        }
        if (i >= s.size()) return fail("unterminated string");
        i++;
        return true;
    }
    bool value(Value &v)
    {
        ws();
        if (i >= s.size()) return fail("unexpected end");
        char c = s[i];
        if (c == '{') {
            v.kind = Value::OBJECT;
            v.obj = std::make_shared<Object>();
            return object(*v.obj);
        }
        if (c == '[') {
            v.kind = Value::ARRAY;
            v.arr = std::make_shared<Array>();
            i++;
            ws();
            if (i < s.size() && s[i] == ']') { i++; return true; }
            for (;;) {
                Value item;
                if (!value(item)) return false;
                v.arr->items.push_back(item);
                ws();
                if (i < s.size() && s[i] == ',') { i++; continue; }
                if (i < s.size() && s[i] == ']') { i++; return true; }
                return fail("expected , or ]");
            }
        }
        if (c == '"') {
            v.kind = Value::STRING;
            return string(v.str);
        }
        if (lit("true")) { v.kind = Value::BOOL; v.b = true; return true; }
        if (lit("false")) { v.kind = Value::BOOL; v.b = false; return true; }
        if (lit("null")) { v.kind = Value::NUL; return true; }
        if (c == '-' || c == '+' || std::isdigit((unsigned char)c) || c == '.') {
            const char *b = s.c_str() + i;
            char *e = nullptr;
            v.num = strtod(b, &e);
            if (e == b) return fail("bad number");
            v.kind = Value::NUMBER;
            i += (size_t)(e - b);
            return true;
        }
        return fail(std::string("unexpected character '") + c + "'");
    }
    bool object(Object &o)
    {
        ws();
        if (i >= s.size() || s[i] != '{') return fail("expected {");
        i++;
        ws();
        if (i < s.size() && s[i] == '}') { i++; return true; }
        for (;;) {
            ws();
            std::string key;
            if (!string(key)) return false;
            ws();
            if (i >= s.size() || s[i] != ':') return fail("expected :");
            i++;
            Value v;
            if (!value(v)) return false;
            o.kv[key] = v;
            ws();
            if (i < s.size() && s[i] == ',') { i++; continue; }
            if (i < s.size() && s[i] == '}') { i++; return true; }
            return fail("expected , or }");
        }
    }
};

}  // namespace

bool Object::parse(const std::string &text)
{
    kv.clear();
    Parser p(text);
    if (!p.object(*this)) {
        last_error = p.err;
        return false;
    }
    p.ws();
    if (p.i != text.size()) {
        last_error = "trailing characters at offset " + std::to_string(p.i);
        return false;
    }
    return true;
}

bool Object::parse(std::istream &in)
{
    std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    return parse(text);
}

}  // namespace json
