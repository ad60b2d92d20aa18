// json.hpp -- the small JSON value type the host side needs to read the
// reference's simulation files (include/SimulationConfig.hpp:73-137) and write
// its result dictionary (NewtonSolver.cpp:775-816): null, bool, number,
// string, array, object (insertion-ordered).  Numbers are doubles, written
// with the fewest significant digits (15 to 17) that read back exactly.
#pragma once
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace cchost {

class Json {
public:
    enum class Kind { Null, Bool, Number, String, Array, Object };

    Json() = default;
    Json(std::nullptr_t) {}
    Json(bool b) : kind_(Kind::Bool), b_(b) {}
    Json(double d) : kind_(Kind::Number), d_(d) {}
    Json(int v) : kind_(Kind::Number), int_(true), d_(v) {}
    Json(long v) : kind_(Kind::Number), int_(true), d_((double)v) {}
    Json(long long v) : kind_(Kind::Number), int_(true), d_((double)v) {}
    Json(unsigned long v) : kind_(Kind::Number), int_(true), d_((double)v) {}
    Json(const char *s) : kind_(Kind::String), s_(s) {}
    Json(std::string s) : kind_(Kind::String), s_(std::move(s)) {}
    Json(const std::vector<double> &v) : kind_(Kind::Array)
    {
        for (double x : v) a_.emplace_back(x);
    }

    static Json array() { Json j; j.kind_ = Kind::Array; return j; }
    static Json object() { Json j; j.kind_ = Kind::Object; return j; }

    Kind kind() const { return kind_; }
    bool is_null() const { return kind_ == Kind::Null; }
    bool is_object() const { return kind_ == Kind::Object; }
    bool is_array() const { return kind_ == Kind::Array; }
    bool is_number() const { return kind_ == Kind::Number; }

    bool contains(const std::string &k) const
    {
        if (kind_ != Kind::Object) return false;
        for (const auto &kv : o_)
            if (kv.first == k) return true;
        return false;
    }
    // object access; a missing key reads as null (as nlohmann's const operator[] on a null)
    const Json &operator[](const std::string &k) const
    {
        static const Json null_value;
        if (kind_ != Kind::Object) return null_value;
        for (const auto &kv : o_)
            if (kv.first == k) return kv.second;
        return null_value;
    }
    Json &operator[](const std::string &k)
    {
        if (kind_ == Kind::Null) kind_ = Kind::Object;
        if (kind_ != Kind::Object) throw std::runtime_error("json: not an object");
        for (auto &kv : o_)
            if (kv.first == k) return kv.second;
        o_.emplace_back(k, Json());
        return o_.back().second;
    }
    const std::vector<std::pair<std::string, Json>> &items() const { return o_; }
    const std::vector<Json> &elements() const { return a_; }
    void push_back(Json v)
    {
        if (kind_ == Kind::Null) kind_ = Kind::Array;
        if (kind_ != Kind::Array) throw std::runtime_error("json: not an array");
        a_.push_back(std::move(v));
    }
    size_t size() const { return kind_ == Kind::Array ? a_.size() : (kind_ == Kind::Object ? o_.size() : 0); }

    double number() const
    {
        if (kind_ == Kind::Number) return d_;
        if (kind_ == Kind::Bool) return b_ ? 1.0 : 0.0;
        throw std::runtime_error("json: not a number");
    }
    bool boolean() const
    {
        if (kind_ == Kind::Bool) return b_;
        if (kind_ == Kind::Number) return d_ != 0.0;
        throw std::runtime_error("json: not a boolean");
    }
    const std::string &string() const
    {
        if (kind_ != Kind::String) throw std::runtime_error("json: not a string");
        return s_;
    }
    std::vector<double> numbers() const
    {
        if (kind_ != Kind::Array) throw std::runtime_error("json: not an array");
        std::vector<double> v;
        v.reserve(a_.size());
        for (const auto &e : a_) v.push_back(e.number());
        return v;
    }

    static Json parse(const std::string &text)
    {
        size_t i = 0;
        Json j = parse_value(text, i);
        skip_ws(text, i);
        if (i != text.size()) throw std::runtime_error("json: trailing characters at " + std::to_string(i));
        return j;
    }

    std::string dump(int indent = 4) const
    {
        std::string out;
        write(out, indent, 0);
        return out;
    }

private:
    Kind kind_ = Kind::Null;
    bool int_ = false;        // an integer literal (written without a fraction)
    bool b_ = false;
    double d_ = 0.0;
    std::string s_;
    std::vector<Json> a_;
    std::vector<std::pair<std::string, Json>> o_;

    static void skip_ws(const std::string &t, size_t &i)
    {
        while (i < t.size() && (t[i] == ' ' || t[i] == '\n' || t[i] == '\r' || t[i] == '\t')) ++i;
    }
    static void expect(const std::string &t, size_t &i, const char *word)
    {
        for (const char *p = word; *p; ++p, ++i)
            if (i >= t.size() || t[i] != *p) throw std::runtime_error(std::string("json: expected ") + word);
    }
    static void put_utf8(std::string &s, unsigned cp)
    {
        if (cp < 0x80) {
            s += (char)cp;
        } else if (cp < 0x800) {
            s += (char)(0xC0 | (cp >> 6));
            s += (char)(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            s += (char)(0xE0 | (cp >> 12));
            s += (char)(0x80 | ((cp >> 6) & 0x3F));
            s += (char)(0x80 | (cp & 0x3F));
        } else {
            s += (char)(0xF0 | (cp >> 18));
            s += (char)(0x80 | ((cp >> 12) & 0x3F));
            s += (char)(0x80 | ((cp >> 6) & 0x3F));
            s += (char)(0x80 | (cp & 0x3F));
        }
    }
    static unsigned hex4(const std::string &t, size_t &i)
    {
        if (i + 4 > t.size()) throw std::runtime_error("json: bad \\u escape");
        const unsigned v = (unsigned)std::stoul(t.substr(i, 4), nullptr, 16);
        i += 4;
        return v;
    }
    static std::string parse_string(const std::string &t, size_t &i)
    {
        if (t[i] != '"') throw std::runtime_error("json: expected string");
        ++i;
        std::string s;
        while (i < t.size() && t[i] != '"') {
            char ch = t[i++];
            if (ch != '\\') {
                s += ch;
                continue;
            }
            if (i >= t.size()) break;
            const char e = t[i++];
            switch (e) {
            case '"': s += '"'; break;
            case '\\': s += '\\'; break;
            case '/': s += '/'; break;
            case 'b': s += '\b'; break;
            case 'f': s += '\f'; break;
            case 'n': s += '\n'; break;
            case 'r': s += '\r'; break;
            case 't': s += '\t'; break;
            case 'u': {
                unsigned cp = hex4(t, i);
                if (cp >= 0xD800 && cp < 0xDC00 && i + 6 <= t.size() && t[i] == '\\' && t[i + 1] == 'u') {
                    i += 2;
                    const unsigned lo = hex4(t, i);
                    cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                }
                put_utf8(s, cp);
                break;
            }
            default: throw std::runtime_error("json: bad escape");
            }
        }
        if (i >= t.size()) throw std::runtime_error("json: unterminated string");
        ++i;
        return s;
    }
    static Json parse_value(const std::string &t, size_t &i)
    {
        skip_ws(t, i);
        if (i >= t.size()) throw std::runtime_error("json: unexpected end");
        const char ch = t[i];
        if (ch == '{') {
            Json j = object();
            ++i;
            skip_ws(t, i);
            if (i < t.size() && t[i] == '}') {
                ++i;
                return j;
            }
            for (;;) {
                skip_ws(t, i);
                std::string k = parse_string(t, i);
                skip_ws(t, i);
                if (i >= t.size() || t[i] != ':') throw std::runtime_error("json: expected ':'");
                ++i;
                j.o_.emplace_back(std::move(k), parse_value(t, i));
                skip_ws(t, i);
                if (i < t.size() && t[i] == ',') {
                    ++i;
                    continue;
                }
                if (i < t.size() && t[i] == '}') {
                    ++i;
                    return j;
                }
                throw std::runtime_error("json: expected ',' or '}'");
            }
        }
        if (ch == '[') {
            Json j = array();
            ++i;
            skip_ws(t, i);
            if (i < t.size() && t[i] == ']') {
                ++i;
                return j;
            }
            for (;;) {
                j.a_.push_back(parse_value(t, i));
                skip_ws(t, i);
                if (i < t.size() && t[i] == ',') {
                    ++i;
                    continue;
                }
                if (i < t.size() && t[i] == ']') {
                    ++i;
                    return j;
                }
                throw std::runtime_error("json: expected ',' or ']'");
            }
        }
        if (ch == '"') return Json(parse_string(t, i));
        if (ch == 't') {
            expect(t, i, "true");
            return Json(true);
        }
        if (ch == 'f') {
            expect(t, i, "false");
            return Json(false);
        }
        if (ch == 'n') {
            expect(t, i, "null");
            return Json();
        }
        const char *begin = t.c_str() + i;
        char *end = nullptr;
        const double d = std::strtod(begin, &end);
        if (end == begin) throw std::runtime_error("json: bad value at " + std::to_string(i));
        Json j(d);
        j.int_ = true;
        for (const char *p = begin; p < end; ++p)
            if (*p == '.' || *p == 'e' || *p == 'E') j.int_ = false;
        i += (size_t)(end - begin);
        return j;
    }
    static void write_string(std::string &out, const std::string &s)
    {
        out += '"';
        for (unsigned char ch : s) {
            switch (ch) {
            case '"': out += "\\\""; break;
            case '\\': out += "\\\\"; break;
            case '\n': out += "\\n"; break;
            case '\r': out += "\\r"; break;
            case '\t': out += "\\t"; break;
            default:
                if (ch < 0x20) {
                    char buf[8];
                    std::snprintf(buf, sizeof buf, "\\u%04x", ch);
                    out += buf;
                } else {
                    out += (char)ch;
                }
            }
        }
        out += '"';
    }
    void write(std::string &out, int indent, int depth) const
    {
        const std::string pad = indent > 0 ? "\n" + std::string((size_t)indent * (depth + 1), ' ') : "";
        const std::string endpad = indent > 0 ? "\n" + std::string((size_t)indent * depth, ' ') : "";
        switch (kind_) {
        case Kind::Null: out += "null"; break;
        case Kind::Bool: out += b_ ? "true" : "false"; break;
        case Kind::Number: {
            if (!std::isfinite(d_)) {
                out += "null";
            } else if (d_ == std::floor(d_) && std::fabs(d_) < 1e15) {
                char buf[32];
                std::snprintf(buf, sizeof buf, int_ ? "%.0f" : "%.1f", d_);
                out += buf;
            } else {
                // shortest of 15, 16, 17 significant digits that reads back exactly
                char buf[32];
                for (int prec = 15; prec <= 17; ++prec) {
                    std::snprintf(buf, sizeof buf, "%.*g", prec, d_);
                    if (std::strtod(buf, nullptr) == d_) break;
                }
                out += buf;
            }
            break;
        }
        case Kind::String: write_string(out, s_); break;
        case Kind::Array:
            out += '[';
            for (size_t k = 0; k < a_.size(); ++k) {
                out += k ? "," : "";
                out += pad;
                a_[k].write(out, indent, depth + 1);
            }
            if (!a_.empty()) out += endpad;
            out += ']';
            break;
        case Kind::Object:
            out += '{';
            for (size_t k = 0; k < o_.size(); ++k) {
                out += k ? "," : "";
                out += pad;
                write_string(out, o_[k].first);
                out += indent > 0 ? ": " : ":";
                o_[k].second.write(out, indent, depth + 1);
            }
            if (!o_.empty()) out += endpad;
            out += '}';
            break;
        }
    }
};

}  // namespace cchost
