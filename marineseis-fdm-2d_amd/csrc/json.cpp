#include "json.hpp"

#include <cerrno>
#include <climits>
#include <clocale>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace msfdm {

namespace {

class Parser {
public:
    explicit Parser(const std::string &t) : s_(t) {}

    std::unique_ptr<JsonValue> document() {
        skip_ws();
        auto v = value(0);
        skip_ws();
        if (p_ != s_.size()) fail("trailing characters after the JSON value");
        return v;
    }

private:
    const std::string &s_;
    size_t p_ = 0;

    [[noreturn]] void fail(const std::string &why) const {
        size_t line = 1, col = 1;
        for (size_t i = 0; i < p_ && i < s_.size(); ++i) {
            if (s_[i] == '\n') { ++line; col = 1; } else { ++col; }
        }
        throw JsonError("JSON parse error at line " + std::to_string(line) + ", column " +
                        std::to_string(col) + ": " + why);
    }
    void skip_ws() {
        while (p_ < s_.size() && (s_[p_] == ' ' || s_[p_] == '\t' || s_[p_] == '\n' || s_[p_] == '\r')) ++p_;
    }
    bool eat(char c) {
        if (p_ < s_.size() && s_[p_] == c) { ++p_; return true; }
        return false;
    }
    void expect_word(const char *w) {
        for (const char *c = w; *c; ++c) {
            if (p_ >= s_.size() || s_[p_] != *c) fail(std::string("invalid literal, expected '") + w + "'");
            ++p_;
        }
    }

    std::unique_ptr<JsonValue> value(int depth) {
        if (depth > 256) fail("nesting too deep");
        if (p_ >= s_.size()) fail("unexpected end of input");
        auto v = std::make_unique<JsonValue>();
        const char c = s_[p_];
        if (c == '{') {
            v->kind = JsonValue::Object;
            ++p_;
            skip_ws();
            if (eat('}')) return v;
            for (;;) {
                skip_ws();
                if (p_ >= s_.size() || s_[p_] != '"') fail("expected a string key");
                std::string key = string_lit();
                skip_ws();
                if (!eat(':')) fail("expected ':'");
                skip_ws();
                v->members.emplace_back(std::move(key), value(depth + 1));
                skip_ws();
                if (eat(',')) continue;
                if (eat('}')) break;
                fail("expected ',' or '}'");
            }
        } else if (c == '[') {
            v->kind = JsonValue::Array;
            ++p_;
            skip_ws();
            if (eat(']')) return v;
            for (;;) {
                skip_ws();
                v->items.push_back(value(depth + 1));
                skip_ws();
                if (eat(',')) continue;
                if (eat(']')) break;
                fail("expected ',' or ']'");
            }
        } else if (c == '"') {
            v->kind = JsonValue::String;
            v->str = string_lit();
        } else if (c == 't') {
            expect_word("true"); v->kind = JsonValue::Bool; v->b = true;
        } else if (c == 'f') {
            expect_word("false"); v->kind = JsonValue::Bool; v->b = false;
        } else if (c == 'n') {
            expect_word("null"); v->kind = JsonValue::Null;
        } else if (c == '-' || (c >= '0' && c <= '9')) {
            v->kind = JsonValue::Number;
            v->num = number();
        } else {
            fail(std::string("unexpected character '") + c + "'");
        }
        return v;
    }

    double number() {
        const size_t b = p_;
        if (s_[p_] == '-') ++p_;
        if (p_ >= s_.size() || !isdigit((unsigned char)s_[p_])) fail("invalid number");
        if (s_[p_] == '0') ++p_;
        else while (p_ < s_.size() && isdigit((unsigned char)s_[p_])) ++p_;
        if (p_ < s_.size() && s_[p_] == '.') {
            ++p_;
            if (p_ >= s_.size() || !isdigit((unsigned char)s_[p_])) fail("invalid fraction");
            while (p_ < s_.size() && isdigit((unsigned char)s_[p_])) ++p_;
        }
        if (p_ < s_.size() && (s_[p_] == 'e' || s_[p_] == 'E')) {
            ++p_;
            if (p_ < s_.size() && (s_[p_] == '+' || s_[p_] == '-')) ++p_;
            if (p_ >= s_.size() || !isdigit((unsigned char)s_[p_])) fail("invalid exponent");
            while (p_ < s_.size() && isdigit((unsigned char)s_[p_])) ++p_;
        }
        // The grammar above only admits '.' as decimal separator, so parse with
        // the "C" numeric convention regardless of the process locale.
        const std::string tok = s_.substr(b, p_ - b);
        std::istringstream is(tok);
        is.imbue(std::locale::classic());
        double d = 0.0;
        is >> d;
        if (is.fail()) {
            // overflow: istream leaves d at +-max; keep IEEE semantics of strtod.
            d = (tok[0] == '-') ? -HUGE_VAL : HUGE_VAL;
        }
        return d;
    }

    static void put_utf8(std::string &o, unsigned cp) {
        if (cp < 0x80) o += char(cp);
        else if (cp < 0x800) { o += char(0xC0 | (cp >> 6)); o += char(0x80 | (cp & 0x3F)); }
        else if (cp < 0x10000) { o += char(0xE0 | (cp >> 12)); o += char(0x80 | ((cp >> 6) & 0x3F)); o += char(0x80 | (cp & 0x3F)); }
        else { o += char(0xF0 | (cp >> 18)); o += char(0x80 | ((cp >> 12) & 0x3F)); o += char(0x80 | ((cp >> 6) & 0x3F)); o += char(0x80 | (cp & 0x3F)); }
    }
    unsigned hex4() {
        if (p_ + 4 > s_.size()) fail("truncated \\u escape");
        unsigned v = 0;
        for (int i = 0; i < 4; ++i) {
            const char c = s_[p_++];
            v <<= 4;
            if (c >= '0' && c <= '9') v |= unsigned(c - '0');
            else if (c >= 'a' && c <= 'f') v |= unsigned(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= unsigned(c - 'A' + 10);
            else fail("invalid \\u escape");
        }
        return v;
    }
    std::string string_lit() {
        ++p_;  // opening quote
        std::string o;
        for (;;) {
            if (p_ >= s_.size()) fail("unterminated string");
            const char c = s_[p_++];
            if (c == '"') break;
            if ((unsigned char)c < 0x20) fail("control character in string");
            if (c != '\\') { o += c; continue; }
            if (p_ >= s_.size()) fail("unterminated escape");
            const char e = s_[p_++];
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
                    unsigned cp = hex4();
                    if (cp >= 0xD800 && cp < 0xDC00 && p_ + 1 < s_.size() && s_[p_] == '\\' && s_[p_ + 1] == 'u') {
                        p_ += 2;
                        const unsigned lo = hex4();
                        if (lo >= 0xDC00 && lo < 0xE000) cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                        else fail("invalid surrogate pair");
                    }
                    put_utf8(o, cp);
                    break;
                }
                default: fail("invalid escape");
            }
        }
        return o;
    }
};

}  // namespace

int json_valueint(double v) {
    if (v >= (double)INT_MAX) return INT_MAX;
    if (v <= (double)INT_MIN) return INT_MIN;
    if (std::isnan(v)) return 0;
    return (int)v;
}

JsonDoc JsonDoc::parse(const std::string &text) {
    if (text.empty()) throw JsonError("JSON text is empty");
    JsonDoc d;
    Parser p(text);
    d.root_ = p.document();
    return d;
}

JsonDoc JsonDoc::load_file(const std::string &path) {
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) throw JsonError("Cannot open JSON file: " + path);
    std::stringstream buf;
    buf << f.rdbuf();
    const std::string text = buf.str();
    if (text.empty()) throw JsonError("JSON file is empty: " + path);
    try {
        return parse(text);
    } catch (const JsonError &e) {
        throw JsonError("Failed to parse JSON file: " + path + ": " + e.what());
    }
}

// Path syntax of include/json_config.hpp:241-356.
const JsonValue *JsonDoc::find(const std::string &path) const {
    const JsonValue *cur = root_.get();
    if (!cur) return nullptr;
    size_t pos = 0;
    while (pos < path.size()) {
        if (path[pos] == '.') { ++pos; continue; }
        if (path[pos] != '[') {
            const size_t b = pos;
            while (pos < path.size() && path[pos] != '.' && path[pos] != '[') ++pos;
            const std::string key = path.substr(b, pos - b);
            if (cur->kind != JsonValue::Object) return nullptr;
            const JsonValue *next = nullptr;
            for (const auto &m : cur->members)
                if (m.first == key) { next = m.second.get(); break; }
            if (!next) return nullptr;
            cur = next;
        }
        while (pos < path.size() && path[pos] == '[') {
            const size_t close = path.find(']', pos);
            if (close == std::string::npos) return nullptr;
            const std::string idx = path.substr(pos + 1, close - pos - 1);
            if (idx.empty()) return nullptr;
            char *end = nullptr;
            const long i = std::strtol(idx.c_str(), &end, 10);
            if (end == idx.c_str() || *end != '\0' || i < 0) return nullptr;
            if (cur->kind != JsonValue::Array || (size_t)i >= cur->items.size()) return nullptr;
            cur = cur->items[(size_t)i].get();
            pos = close + 1;
        }
        if (pos < path.size() && path[pos] == '.') ++pos;
    }
    return cur;
}

static const JsonValue &need(const JsonDoc &d, const std::string &path) {
    const JsonValue *v = d.find(path);
    if (!v) throw JsonError("JSON path not found: " + path);
    return *v;
}

int JsonDoc::get_int(const std::string &path) const {
    const JsonValue &v = need(*this, path);
    if (v.kind != JsonValue::Number) throw JsonError("JSON type mismatch at '" + path + "': expected int");
    return json_valueint(v.num);
}

int JsonDoc::get_int(const std::string &path, size_t index) const {
    const JsonValue &a = need(*this, path);
    if (a.kind != JsonValue::Array) throw JsonError("JSON node is not an array: " + path);
    if (index >= a.items.size())
        throw JsonError("Array index out of range: " + path + "[" + std::to_string(index) + "]");
    const JsonValue &v = *a.items[index];
    if (v.kind != JsonValue::Number)
        throw JsonError("JSON type mismatch at '" + path + "[" + std::to_string(index) + "]': expected int");
    return json_valueint(v.num);
}

float JsonDoc::get_float(const std::string &path) const {
    const JsonValue &v = need(*this, path);
    if (v.kind != JsonValue::Number) throw JsonError("JSON type mismatch at '" + path + "': expected float");
    return (float)v.num;
}

std::string JsonDoc::get_string(const std::string &path) const {
    const JsonValue &v = need(*this, path);
    if (v.kind != JsonValue::String) throw JsonError("JSON type mismatch at '" + path + "': expected string");
    return v.str;
}

size_t JsonDoc::array_size(const std::string &path) const {
    const JsonValue &v = need(*this, path);
    if (v.kind != JsonValue::Array) throw JsonError("JSON node is not an array: " + path);
    return v.items.size();
}

}  // namespace msfdm
