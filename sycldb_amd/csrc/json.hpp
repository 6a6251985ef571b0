// Minimal JSON reader for plan documents (objects, arrays, strings, integers,
// booleans, null).  Numbers keep their integer value (plans carry int64
// literals); fractional numbers are rejected.
//
// A Document parses the whole text into one contiguous array of nodes: the
// members of an object (or the elements of an array) sit next to each other,
// and strings are views into the text, so parsing allocates a few vectors
// rather than a string and two vectors per value.  A string with escapes is
// unescaped into storage the Document owns.  Nodes and views stay valid while
// the Document and the text live.
#pragma once

#include <cstdint>
#include <deque>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

namespace sdb {
namespace json {

struct Value {
    enum Kind : uint8_t { Null, Bool, Int, String, Array, Object } kind = Null;
    bool b = false;
    int64_t i = 0;
    std::string_view s;   // String: the unescaped text
    std::string_view key; // the member's key when the value sits in an object
    const Value *kids = nullptr; // Array elements / Object members, contiguous
    uint32_t first = 0, n = 0;   // (first: the kids' index while parsing)

    struct Range {
        const Value *b, *e;
        const Value *begin() const { return b; }
        const Value *end() const { return e; }
        size_t size() const { return (size_t)(e - b); }
    };

    // object members in document order (plan objects have a handful of
    // keys: a linear search); a repeated key keeps its last value
    const Value *find(std::string_view k) const
    {
        if (kind != Object) return nullptr;
        for (uint32_t j = n; j-- > 0;)
            if (kids[j].key == k) return &kids[j];
        return nullptr;
    }
    bool has(std::string_view k) const { return find(k) != nullptr; }
    const Value &at(std::string_view k) const
    {
        const Value *v = find(k);
        if (!v) throw std::invalid_argument("plan: missing field '" + std::string(k) + "'");
        return *v;
    }
    int64_t as_int() const
    {
        if (kind != Int) throw std::invalid_argument("plan: expected an integer");
        return i;
    }
    std::string_view as_str() const
    {
        if (kind != String) throw std::invalid_argument("plan: expected a string");
        return s;
    }
    Range as_arr() const
    {
        if (kind != Array) throw std::invalid_argument("plan: expected an array");
        return {kids, kids + n};
    }
};

class Document {
  public:
    explicit Document(std::string_view text) : t_(text)
    {
        nodes_.reserve(t_.size() / 8 + 4);
        pending_.reserve(64);
        value(0);
        ws();
        if (p_ != t_.size()) err("trailing characters");
        nodes_.push_back(pending_.back());
        for (Value &v : nodes_)
            if (v.kind == Value::Array || v.kind == Value::Object) v.kids = nodes_.data() + v.first;
    }
    Document(const Document &) = delete;
    Document &operator=(const Document &) = delete;
    const Value &root() const { return nodes_.back(); }

  private:
    static constexpr int MAX_DEPTH = 256;
    std::string_view t_;
    size_t p_ = 0;
    std::vector<Value> nodes_;   // finished containers' children, then the root
    std::vector<Value> pending_; // values whose container is still open
    std::deque<std::string> unescaped_;

    [[noreturn]] void err(const char *m) const
    {
        throw std::invalid_argument(std::string("plan JSON: ") + m + " at offset " + std::to_string(p_));
    }
    void ws()
    {
        while (p_ < t_.size() && (t_[p_] == ' ' || t_[p_] == '\n' || t_[p_] == '\t' || t_[p_] == '\r')) p_++;
    }
    bool lit(std::string_view w)
    {
        if (t_.compare(p_, w.size(), w) == 0) {
            p_ += w.size();
            return true;
        }
        return false;
    }
    // the children pending since `start` move to the node array, contiguous
    void close(Value &v, size_t start)
    {
        v.first = (uint32_t)nodes_.size();
        v.n = (uint32_t)(pending_.size() - start);
        nodes_.insert(nodes_.end(), pending_.begin() + (std::ptrdiff_t)start, pending_.end());
        pending_.resize(start);
    }
    // parses one value and leaves it on top of the pending list
    void value(int depth)
    {
        ws();
        if (p_ >= t_.size()) err("unexpected end");
        if (depth > MAX_DEPTH) err("nesting too deep");
        const char c = t_[p_];
        Value v;
        if (c == '{') {
            v.kind = Value::Object;
            p_++;
            ws();
            const size_t start = pending_.size();
            if (p_ < t_.size() && t_[p_] == '}') {
                p_++;
            } else {
                for (;;) {
                    ws();
                    if (p_ >= t_.size() || t_[p_] != '"') err("expected key");
                    const std::string_view k = str();
                    ws();
                    if (p_ >= t_.size() || t_[p_] != ':') err("expected ':'");
                    p_++;
                    value(depth + 1);
                    pending_.back().key = k;
                    ws();
                    if (p_ < t_.size() && t_[p_] == ',') {
                        p_++;
                        continue;
                    }
                    if (p_ < t_.size() && t_[p_] == '}') {
                        p_++;
                        break;
                    }
                    err("expected ',' or '}'");
                }
            }
            close(v, start);
        } else if (c == '[') {
            v.kind = Value::Array;
            p_++;
            ws();
            const size_t start = pending_.size();
            if (p_ < t_.size() && t_[p_] == ']') {
                p_++;
            } else {
                for (;;) {
                    value(depth + 1);
                    ws();
                    if (p_ < t_.size() && t_[p_] == ',') {
                        p_++;
                        continue;
                    }
                    if (p_ < t_.size() && t_[p_] == ']') {
                        p_++;
                        break;
                    }
                    err("expected ',' or ']'");
                }
            }
            close(v, start);
        } else if (c == '"') {
            v.kind = Value::String;
            v.s = str();
        } else if (lit("true")) {
            v.kind = Value::Bool;
            v.b = true;
        } else if (lit("false")) {
            v.kind = Value::Bool;
        } else if (lit("null")) {
        } else if (c == '-' || (c >= '0' && c <= '9')) {
            const size_t s = p_;
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
        } else {
            err("unexpected character");
        }
        pending_.push_back(v);
    }
    std::string_view str()
    {
        p_++; // opening quote
        // fast path: no escape before the closing quote -- a view of the text
        {
            size_t q = p_;
            while (q < t_.size() && t_[q] != '"' && t_[q] != '\\') q++;
            if (q < t_.size() && t_[q] == '"') {
                const std::string_view out = t_.substr(p_, q - p_);
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
        unescaped_.push_back(std::move(out));
        return unescaped_.back();
    }
};

} // namespace json
} // namespace sdb
