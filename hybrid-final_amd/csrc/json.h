// Minimal JSON reader for case decks.  Lookups follow the reference's parson
// usage (src/start.c:19-1431, src/parson.c:1319-1321): a missing key, a
// missing parent object or a value of the wrong type reads as 0 / empty.
#pragma once
#include <cctype>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace pic {

struct JValue {
    enum Kind { NUL, NUM, STR, BOOL, ARR, OBJ } kind = NUL;
    double num = 0.0;
    bool b = false;
    std::string str;
    std::vector<JValue> arr;
    std::vector<std::pair<std::string, JValue>> obj;

    const JValue *get(const std::string &k) const {
        if (kind != OBJ) return nullptr;
        for (auto &kv : obj)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
    // parson json_object_get_number: 0 unless the key holds a number
    double number(const std::string &k) const {
        const JValue *v = get(k);
        return (v && v->kind == NUM) ? v->num : 0.0;
    }
    const JValue &object(const std::string &k) const {
        static const JValue empty;
        const JValue *v = get(k);
        return (v && v->kind == OBJ) ? *v : empty;
    }
    const std::vector<JValue> &array(const std::string &k) const {
        static const std::vector<JValue> empty;
        const JValue *v = get(k);
        return (v && v->kind == ARR) ? v->arr : empty;
    }
};

class JParser {
  public:
    explicit JParser(const std::string &s) : s_(s), i_(0) {}
    JValue parse() {
        JValue v = value();
        ws();
        if (i_ != s_.size()) fail("trailing characters");
        return v;
    }

  private:
    const std::string &s_;
    size_t i_;
    [[noreturn]] void fail(const char *m) {
        throw std::runtime_error(std::string("deck JSON: ") + m + " at offset " + std::to_string(i_));
    }
    void ws() {
        while (i_ < s_.size() && std::isspace((unsigned char)s_[i_])) ++i_;
    }
    JValue value() {
        ws();
        if (i_ >= s_.size()) fail("unexpected end");
        char c = s_[i_];
        JValue v;
        if (c == '{') {
            v.kind = JValue::OBJ;
            ++i_;
            ws();
            if (i_ < s_.size() && s_[i_] == '}') { ++i_; return v; }
            for (;;) {
                ws();
                if (i_ >= s_.size() || s_[i_] != '"') fail("expected key");
                std::string k = string();
                ws();
                if (i_ >= s_.size() || s_[i_] != ':') fail("expected ':'");
                ++i_;
                JValue e = value();
                v.obj.emplace_back(std::move(k), std::move(e));
                ws();
                if (i_ < s_.size() && s_[i_] == ',') { ++i_; continue; }
                if (i_ < s_.size() && s_[i_] == '}') { ++i_; break; }
                fail("expected ',' or '}'");
            }
        } else if (c == '[') {
            v.kind = JValue::ARR;
            ++i_;
            ws();
            if (i_ < s_.size() && s_[i_] == ']') { ++i_; return v; }
            for (;;) {
                v.arr.push_back(value());
                ws();
                if (i_ < s_.size() && s_[i_] == ',') { ++i_; continue; }
                if (i_ < s_.size() && s_[i_] == ']') { ++i_; break; }
                fail("expected ',' or ']'");
            }
        } else if (c == '"') {
            v.kind = JValue::STR;
            v.str = string();
        } else if (s_.compare(i_, 4, "true") == 0) {
            v.kind = JValue::BOOL; v.b = true; i_ += 4;
        } else if (s_.compare(i_, 5, "false") == 0) {
            v.kind = JValue::BOOL; v.b = false; i_ += 5;
        } else if (s_.compare(i_, 4, "null") == 0) {
            v.kind = JValue::NUL; i_ += 4;
        } else {
            const char *start = s_.c_str() + i_;
            char *end = nullptr;
            double d = std::strtod(start, &end);
            if (end == start) fail("bad value");
            v.kind = JValue::NUM;
            v.num = d;
            i_ += (size_t)(end - start);
        }
        return v;
    }
    std::string string() {
        std::string out;
        ++i_;  // opening quote
        while (i_ < s_.size() && s_[i_] != '"') {
            char c = s_[i_++];
            if (c == '\\' && i_ < s_.size()) {
                char e = s_[i_++];
                switch (e) {
                case 'n': out += '\n'; break;
                case 't': out += '\t'; break;
                case 'r': out += '\r'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'u': i_ += 4; out += '?'; break;
                default: out += e;
                }
            } else {
                out += c;
            }
        }
        if (i_ >= s_.size()) fail("unterminated string");
        ++i_;
        return out;
    }
};

}  // namespace pic
