// Minimal JSON reader for plan documents (objects, arrays, strings, integers,
// booleans, null).  Numbers keep their integer value (plans carry int64
// literals); fractional numbers are rejected.
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace sdb {
namespace json {

struct Value {
    enum Kind { Null, Bool, Int, String, Array, Object } kind = Null;
    bool b = false;
    int64_t i = 0;
    std::string s;
    std::vector<Value> a;
    // object members in document order (plan objects have a handful of
    // keys: a linear search beats a tree); a repeated key keeps its last value
    std::vector<std::pair<std::string, Value>> o;

    const Value *find(const char *k) const
    {
        if (kind != Object) return nullptr;
        for (auto it = o.rbegin(); it != o.rend(); ++it)
            if (it->first == k) return &it->second;
        return nullptr;
    }
    bool has(const char *k) const { return find(k) != nullptr; }
    bool has(const std::string &k) const { return find(k.c_str()) != nullptr; }
    const Value &at(const char *k) const
    {
        const Value *v = find(k);
        if (!v) throw std::invalid_argument(std::string("plan: missing field '") + k + "'");
        return *v;
    }
    const Value &at(const std::string &k) const { return at(k.c_str()); }
    int64_t as_int() const
    {
        if (kind != Int) throw std::invalid_argument("plan: expected an integer");
        return i;
    }
    const std::string &as_str() const
    {
        if (kind != String) throw std::invalid_argument("plan: expected a string");
        return s;
    }
    const std::vector<Value> &as_arr() const
    {
        if (kind != Array) throw std::invalid_argument("plan: expected an array");
        return a;
    }
};

class Parser {
  public:
    explicit Parser(const std::string &t) : t_(t) {}
    Value parse()
    {
        Value v = value();
        ws();
        if (p_ != t_.size()) err("trailing characters");
        return v;
    }

  private:
    const std::string &t_;
    size_t p_ = 0;

    [[noreturn]] void err(const char *m)
    {
        throw std::invalid_argument(std::string("plan JSON: ") + m + " at offset " + std::to_string(p_));
    }
    void ws()
    {
        while (p_ < t_.size() && (t_[p_] == ' ' || t_[p_] == '\n' || t_[p_] == '\t' || t_[p_] == '\r')) p_++;
    }
    bool lit(const char *w)
    {
        size_t n = std::char_traits<char>::length(w);
        if (t_.compare(p_, n, w) == 0) {
            p_ += n;
            return true;
        }
        return false;
    }
    Value value()
    {
        ws();
        if (p_ >= t_.size()) err("unexpected end");
        char c = t_[p_];
        Value v;
        if (c == '{') {
            v.kind = Value::Object;
            p_++;
            ws();
            if (p_ < t_.size() && t_[p_] == '}') {
                p_++;
                return v;
            }
            for (;;) {
                ws();
                if (p_ >= t_.size() || t_[p_] != '"') err("expected key");
                std::string k = str();
                ws();
                if (p_ >= t_.size() || t_[p_] != ':') err("expected ':'");
                p_++;
                v.o.emplace_back(std::move(k), value());
                ws();
                if (p_ < t_.size() && t_[p_] == ',') {
                    p_++;
                    continue;
                }
                if (p_ < t_.size() && t_[p_] == '}') {
                    p_++;
                    return v;
                }
                err("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.kind = Value::Array;
            p_++;
            ws();
            if (p_ < t_.size() && t_[p_] == ']') {
                p_++;
                return v;
            }
            for (;;) {
                v.a.push_back(value());
                ws();
                if (p_ < t_.size() && t_[p_] == ',') {
                    p_++;
                    continue;
                }
                if (p_ < t_.size() && t_[p_] == ']') {
                    p_++;
                    return v;
                }
                err("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.kind = Value::String;
            v.s = str();
            return v;
        }
        if (lit("true")) {
            v.kind = Value::Bool;
            v.b = true;
            return v;
        }
        if (lit("false")) {
            v.kind = Value::Bool;
            return v;
        }
        if (lit("null")) return v;
        if (c == '-' || (c >= '0' && c <= '9')) {
            size_t s = p_;
            if (t_[p_] == '-') p_++;
            while (p_ < t_.size() && t_[p_] >= '0' && t_[p_] <= '9') p_++;
            if (p_ < t_.size() && (t_[p_] == '.' || t_[p_] == 'e' || t_[p_] == 'E')) err("non-integer number");
            v.kind = Value::Int;
            // digits accumulated in uint64 with an overflow check (the
            // range of int64, as std::stoll accepts)
            const bool neg = t_[s] == '-';
            const size_t d0 = s + (neg ? 1 : 0);
            if (d0 == p_) err("expected digits");
            uint64_t mag = 0;
            const uint64_t limit = neg ? uint64_t(1) << 63 : (uint64_t(1) << 63) - 1;
            for (size_t q = d0; q < p_; q++) {
                const uint64_t d = (uint64_t)(t_[q] - '0');
                if (mag > (limit - d) / 10) err("integer out of range");
                mag = mag * 10 + d;
            }
            v.i = neg ? (int64_t)(uint64_t(0) - mag) : (int64_t)mag;
            return v;
        }
        err("unexpected character");
    }
    std::string str()
    {
        p_++; // opening quote
        // fast path: no escape before the closing quote
        {
            size_t q = p_;
            while (q < t_.size() && t_[q] != '"' && t_[q] != '\\') q++;
            if (q < t_.size() && t_[q] == '"') {
                std::string out(t_, p_, q - p_);
                p_ = q + 1;
                return out;
            }
        }
        std::string out;
        while (p_ < t_.size() && t_[p_] != '"') {
            char c = t_[p_++];
            if (c == '\\') {
                if (p_ >= t_.size()) err("bad escape");
                char e = t_[p_++];
                switch (e) {
                case 'n': out += '\n'; break;
                case 't': out += '\t'; break;
                case 'r': out += '\r'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'u':
                    if (p_ + 4 > t_.size()) err("bad unicode escape");
                    out += '?';
                    p_ += 4;
                    break;
                default: out += e; break;
                }
            } else {
                out += c;
            }
        }
        if (p_ >= t_.size()) err("unterminated string");
        p_++;
        return out;
    }
};

inline Value parse(const std::string &text) { return Parser(text).parse(); }

} // namespace json
} // namespace sdb
