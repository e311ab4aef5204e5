// TOML reader for scene files (see sol/toml.h for the subset).
#include "sol/toml.h"

#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <limits>

#include "sol/scene_loader.h"

namespace sol::toml {

const Value* Value::get(std::string_view key) const {
    if (!is_table()) return nullptr;
    for (auto& [k, v] : table)
        if (k == key) return v.get();
    return nullptr;
}

std::optional<double> Value::number() const {
    if (kind == Kind::Float) return real;
    if (kind == Kind::Integer) return static_cast<double>(integer);
    return std::nullopt;
}

std::optional<std::string> Value::string() const {
    if (kind == Kind::String) return str;
    return std::nullopt;
}

namespace {

bool is_bare(char c) {
    return (c >= 'A' && c <= 'Z') || (c >= 'a' && c <= 'z') || (c >= '0' && c <= '9') || c == '_' || c == '-';
}
bool is_dec(char c) { return c >= '0' && c <= '9'; }

void append_utf8(std::string& out, uint32_t cp) {
    if (cp < 0x80) {
        out += static_cast<char>(cp);
    } else if (cp < 0x800) {
        out += static_cast<char>(0xC0 | (cp >> 6));
        out += static_cast<char>(0x80 | (cp & 0x3F));
    } else if (cp < 0x10000) {
        out += static_cast<char>(0xE0 | (cp >> 12));
        out += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
        out += static_cast<char>(0x80 | (cp & 0x3F));
    } else {
        out += static_cast<char>(0xF0 | (cp >> 18));
        out += static_cast<char>(0x80 | ((cp >> 12) & 0x3F));
        out += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
        out += static_cast<char>(0x80 | (cp & 0x3F));
    }
}

class Parser {
public:
    Parser(std::string_view text, const std::string& file) : s_(text), file_(file) {}

    std::unique_ptr<Value> run() {
        auto root = std::make_unique<Value>(Value::Kind::Table);
        Value* current = root.get();
        while (true) {
            skip_blank();
            if (eof()) break;
            if (peek() == '[') {
                current = header(root.get());
            } else {
                key_value(current);
            }
            end_of_line();
        }
        return root;
    }

private:
    std::string_view s_;
    const std::string& file_;
    size_t pos_ = 0, line_ = 1, col_ = 1;

    bool eof() const { return pos_ >= s_.size(); }
    char peek(size_t k = 0) const { return pos_ + k < s_.size() ? s_[pos_ + k] : '\0'; }
    char get() {
        char c = s_[pos_++];
        if (c == '\n') {
            ++line_;
            col_ = 1;
        } else {
            ++col_;
        }
        return c;
    }
    [[noreturn]] void error(const std::string& msg) const { throw SourceError(file_, line_, col_, msg); }
    bool newline_here() const { return peek() == '\n' || (peek() == '\r' && peek(1) == '\n'); }
    void take_newline() {
        if (peek() == '\r') get();
        get();
    }
    void skip_ws() {
        while (peek() == ' ' || peek() == '\t') get();
    }
    void skip_comment() {
        if (peek() == '#')
            while (!eof() && !newline_here()) get();
    }
    // White space, comments and newlines (between statements and inside arrays).
    void skip_blank() {
        while (true) {
            skip_ws();
            skip_comment();
            if (!eof() && newline_here()) {
                take_newline();
                continue;
            }
            return;
        }
    }
    void end_of_line() {
        skip_ws();
        skip_comment();
        if (eof()) return;
        if (!newline_here()) error("expected the end of the line");
        take_newline();
    }
    void expect(char c, const char* what) {
        if (peek() != c) error(std::string("expected ") + what);
        get();
    }

    std::string simple_key() {
        if (peek() == '"') {
            get();
            return basic_string(false);
        }
        if (peek() == '\'') {
            get();
            return literal_string(false);
        }
        std::string k;
        while (is_bare(peek())) k += get();
        if (k.empty()) error("expected a key");
        return k;
    }

    std::vector<std::string> dotted_key() {
        std::vector<std::string> keys{simple_key()};
        while (true) {
            skip_ws();
            if (peek() != '.') return keys;
            get();
            skip_ws();
            keys.push_back(simple_key());
        }
    }

    // The table a dotted key or header path goes through; created if missing.
    Value* descend(Value* t, const std::string& k) {
        for (auto& [key, v] : t->table) {
            if (key != k) continue;
            if (v->is_table()) {
                if (v->sealed) error("inline table '" + k + "' cannot be extended");
                return v.get();
            }
            if (v->is_array() && v->table_array) return v->array.back().get();
            error("key '" + k + "' is not a table");
        }
        t->table.emplace_back(k, std::make_unique<Value>(Value::Kind::Table));
        return t->table.back().second.get();
    }

    Value* header(Value* root) {
        get();
        const bool is_array = peek() == '[';
        if (is_array) get();
        skip_ws();
        std::vector<std::string> keys = dotted_key();
        skip_ws();
        expect(']', "']'");
        if (is_array) expect(']', "']]'");
        Value* t = root;
        for (size_t i = 0; i + 1 < keys.size(); ++i) t = descend(t, keys[i]);
        const std::string& last = keys.back();
        for (auto& [key, v] : t->table) {
            if (key != last) continue;
            if (is_array) {
                if (!v->is_array() || !v->table_array) error("key '" + last + "' is not an array of tables");
                v->array.push_back(std::make_unique<Value>(Value::Kind::Table));
                return v->array.back().get();
            }
            if (!v->is_table() || v->sealed) error("key '" + last + "' is not a table");
            if (v->header_defined) error("table '" + last + "' is defined twice");
            v->header_defined = true;
            return v.get();
        }
        if (is_array) {
            auto arr = std::make_unique<Value>(Value::Kind::Array);
            arr->table_array = true;
            arr->array.push_back(std::make_unique<Value>(Value::Kind::Table));
            Value* first = arr->array.back().get();
            t->table.emplace_back(last, std::move(arr));
            return first;
        }
        auto tab = std::make_unique<Value>(Value::Kind::Table);
        tab->header_defined = true;
        t->table.emplace_back(last, std::move(tab));
        return t->table.back().second.get();
    }

    void key_value(Value* t) {
        std::vector<std::string> keys = dotted_key();
        skip_ws();
        expect('=', "'='");
        skip_ws();
        for (size_t i = 0; i + 1 < keys.size(); ++i) t = descend(t, keys[i]);
        if (t->get(keys.back())) error("duplicate key '" + keys.back() + "'");
        auto v = value();
        t->table.emplace_back(keys.back(), std::move(v));
    }

    std::unique_ptr<Value> value() {
        const size_t line = line_, col = col_;
        std::unique_ptr<Value> v;
        char c = peek();
        if (c == '"' || c == '\'') {
            v = std::make_unique<Value>(Value::Kind::String);
            bool multi = peek(1) == c && peek(2) == c;
            get();
            if (multi) {
                get();
                get();
            }
            v->str = c == '"' ? basic_string(multi) : literal_string(multi);
        } else if (s_.substr(pos_, 4) == "true") {
            for (int i = 0; i < 4; ++i) get();
            v = std::make_unique<Value>(Value::Kind::Boolean);
            v->boolean = true;
        } else if (s_.substr(pos_, 5) == "false") {
            for (int i = 0; i < 5; ++i) get();
            v = std::make_unique<Value>(Value::Kind::Boolean);
        } else if (c == '[') {
            v = array();
        } else if (c == '{') {
            v = inline_table();
        } else if (is_dec(c) || c == '+' || c == '-' || c == 'i' || c == 'n') {
            v = number();
        } else {
            error("expected a value");
        }
        v->line = line;
        v->col = col;
        return v;
    }

    uint32_t hex_digits(int n) {
        uint32_t cp = 0;
        for (int i = 0; i < n; ++i) {
            char c = peek();
            int d = is_dec(c) ? c - '0' : (c >= 'a' && c <= 'f') ? c - 'a' + 10 : (c >= 'A' && c <= 'F') ? c - 'A' + 10 : -1;
            if (d < 0) error("invalid unicode escape");
            get();
            cp = cp * 16 + static_cast<uint32_t>(d);
        }
        if (cp > 0x10FFFF || (cp >= 0xD800 && cp <= 0xDFFF)) error("invalid unicode scalar value");
        return cp;
    }

    // After the opening quote(s).
    std::string basic_string(bool multi) {
        std::string out;
        if (multi && newline_here()) take_newline();
        while (true) {
            if (eof()) error("unterminated string");
            char c = peek();
            if (c == '"') {
                if (!multi) {
                    get();
                    return out;
                }
                if (peek(1) == '"' && peek(2) == '"' && peek(3) != '"') {
                    get();
                    get();
                    get();
                    return out;
                }
                out += get();
                continue;
            }
            if (!multi && newline_here()) error("newline in a single-line string");
            if (c == '\\') {
                get();
                char e = peek();
                if (multi && (e == ' ' || e == '\t' || newline_here())) {
                    // line-ending backslash: drop white space and newlines up to the next text
                    skip_ws();
                    if (!newline_here()) error("invalid escape");
                    while (peek() == ' ' || peek() == '\t' || newline_here()) {
                        if (newline_here()) take_newline();
                        else get();
                    }
                    continue;
                }
                get();
                switch (e) {
                    case 'b': out += '\b'; break;
                    case 't': out += '\t'; break;
                    case 'n': out += '\n'; break;
                    case 'f': out += '\f'; break;
                    case 'r': out += '\r'; break;
                    case '"': out += '"'; break;
                    case '\\': out += '\\'; break;
                    case 'u': append_utf8(out, hex_digits(4)); break;
                    case 'U': append_utf8(out, hex_digits(8)); break;
                    default: error(std::string("invalid escape '\\") + e + "'");
                }
                continue;
            }
            out += get();
        }
    }

    std::string literal_string(bool multi) {
        std::string out;
        if (multi && newline_here()) take_newline();
        while (true) {
            if (eof()) error("unterminated string");
            char c = peek();
            if (c == '\'') {
                if (!multi) {
                    get();
                    return out;
                }
                if (peek(1) == '\'' && peek(2) == '\'' && peek(3) != '\'') {
                    get();
                    get();
                    get();
                    return out;
                }
            } else if (!multi && newline_here()) {
                error("newline in a single-line string");
            }
            out += get();
        }
    }

    std::unique_ptr<Value> array() {
        get();
        auto v = std::make_unique<Value>(Value::Kind::Array);
        v->sealed = true;
        while (true) {
            skip_blank();
            if (peek() == ']') {
                get();
                return v;
            }
            v->array.push_back(value());
            skip_blank();
            if (peek() == ',') {
                get();
                continue;
            }
            if (peek() == ']') {
                get();
                return v;
            }
            error("expected ',' or ']' in an array");
        }
    }

    std::unique_ptr<Value> inline_table() {
        get();
        auto v = std::make_unique<Value>(Value::Kind::Table);
        skip_ws();
        if (peek() == '}') {
            get();
            v->sealed = true;
            return v;
        }
        while (true) {
            skip_ws();
            key_value(v.get());
            skip_ws();
            if (peek() == ',') {
                get();
                continue;
            }
            if (peek() == '}') {
                get();
                v->sealed = true;
                return v;
            }
            error("expected ',' or '}' in an inline table");
        }
    }

    // Digits with single '_' between them; returns them without the separators.
    std::string digits(const std::string& tok, size_t from, size_t to, bool (*ok)(char)) {
        std::string out;
        for (size_t i = from; i < to; ++i) {
            char c = tok[i];
            if (c == '_') {
                if (i == from || i + 1 >= to || !ok(tok[i - 1]) || !ok(tok[i + 1])) error("misplaced '_' in a number");
                continue;
            }
            if (!ok(c)) error("invalid number '" + tok + "'");
            out += c;
        }
        if (out.empty()) error("invalid number '" + tok + "'");
        return out;
    }

    std::unique_ptr<Value> number() {
        std::string tok;
        while (!eof()) {
            char c = peek();
            if (is_bare(c) || c == '+' || c == '.' || c == ':') tok += get();
            else break;
        }
        if (tok.size() >= 5 && is_dec(tok[0]) && is_dec(tok[1]) && is_dec(tok[2]) && is_dec(tok[3]) && tok[4] == '-')
            error("date-time values are not supported");
        if (tok.size() >= 3 && is_dec(tok[0]) && is_dec(tok[1]) && tok[2] == ':')
            error("date-time values are not supported");
        size_t i = 0;
        bool neg = false;
        if (tok[0] == '+' || tok[0] == '-') {
            neg = tok[0] == '-';
            i = 1;
        }
        std::string body = tok.substr(i);
        if (body == "inf" || body == "nan") {
            auto v = std::make_unique<Value>(Value::Kind::Float);
            v->real = body == "inf" ? std::numeric_limits<double>::infinity() : std::numeric_limits<double>::quiet_NaN();
            if (neg) v->real = -v->real;
            return v;
        }
        if (body.size() > 2 && body[0] == '0' && (body[1] == 'x' || body[1] == 'o' || body[1] == 'b')) {
            if (i) error("prefixed integers take no sign");
            int base = body[1] == 'x' ? 16 : body[1] == 'o' ? 8 : 2;
            auto ok = base == 16 ? +[](char c) { return is_dec(c) || (c >= 'a' && c <= 'f') || (c >= 'A' && c <= 'F'); }
                    : base == 8 ? +[](char c) { return c >= '0' && c <= '7'; }
                                : +[](char c) { return c == '0' || c == '1'; };
            std::string d = digits(body, 2, body.size(), ok);
            errno = 0;
            unsigned long long u = std::strtoull(d.c_str(), nullptr, base);
            if (errno == ERANGE || u > static_cast<unsigned long long>(std::numeric_limits<int64_t>::max()))
                error("integer out of range");
            auto v = std::make_unique<Value>(Value::Kind::Integer);
            v->integer = static_cast<int64_t>(u);
            return v;
        }
        // decimal: integer part, optional fraction, optional exponent
        size_t dot = body.find('.'), exp = body.find_first_of("eE");
        size_t int_end = std::min(dot, exp) == std::string::npos ? body.size() : std::min(dot, exp);
        std::string ip = digits(body, 0, int_end, is_dec);
        if (ip.size() > 1 && ip[0] == '0') error("leading zeros in '" + tok + "'");
        std::string clean = (neg ? "-" : "") + ip;
        bool is_float = false;
        if (dot != std::string::npos && dot == int_end) {
            size_t frac_end = exp == std::string::npos ? body.size() : exp;
            clean += "." + digits(body, dot + 1, frac_end, is_dec);
            is_float = true;
        }
        if (exp != std::string::npos) {
            size_t e = exp + 1;
            std::string sign;
            if (e < body.size() && (body[e] == '+' || body[e] == '-')) sign = body[e++];
            clean += "e" + sign + digits(body, e, body.size(), is_dec);
            is_float = true;
        } else if (!is_float && int_end != body.size()) {
            error("invalid number '" + tok + "'");
        }
        if (is_float) {
            auto v = std::make_unique<Value>(Value::Kind::Float);
            v->real = std::strtod(clean.c_str(), nullptr);
            return v;
        }
        errno = 0;
        long long n = std::strtoll(clean.c_str(), nullptr, 10);
        if (errno == ERANGE) error("integer out of range");
        auto v = std::make_unique<Value>(Value::Kind::Integer);
        v->integer = n;
        return v;
    }
};

} // namespace

std::unique_ptr<Value> parse(std::string_view text, const std::string& file_name) {
    return Parser(text, file_name).run();
}

} // namespace sol::toml
