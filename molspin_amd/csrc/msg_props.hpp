// Property lists of MSD objects ("key=value; key=value; ..."), split by the
// rules of MSDParser::ObjectParser (MSDParser/ObjectParser.cpp:37-86): ';' ends a
// value unless inside ( ); keys are matched case-insensitively, and one pair of
// enclosing quotes is dropped from a value.  Shared by the plan resolver and the
// MSD front-end.
#pragma once
#include <cctype>
#include <map>
#include <string>
#include <utility>
#include <vector>

namespace msg {

inline std::string trim(const std::string &s) {
    size_t b = 0, e = s.size();
    while (b < e && std::isspace((unsigned char)s[b])) ++b;
    while (e > b && std::isspace((unsigned char)s[e - 1])) --e;
    return s.substr(b, e - b);
}

inline std::string lower(std::string s) {
    for (char &c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

struct Props {
    std::map<std::string, std::string> f;
    std::vector<std::pair<std::string, std::string>> raw;  // keys as written, in order

    explicit Props(const char *text) {
        std::string key, value;
        bool is_key = true;
        int paren = 0;
        auto flush = [&]() {
            std::string k = trim(key), v = trim(value);
            if (v.size() >= 2 && v.front() == '"' && v.back() == '"') v = v.substr(1, v.size() - 2);
            if (!k.empty() && !v.empty()) {
                f[lower(k)] = v;
                raw.emplace_back(k, v);
            }
            key.clear();
            value.clear();
            is_key = true;
            paren = 0;
        };
        for (const char *p = text; *p; ++p) {
            const char c = *p;
            if (is_key) {
                if (c == '=') is_key = false;
                else if (c == ';') key.clear();
                else key += c;
            } else {
                if (c == '(') ++paren;
                else if (c == ')') --paren;
                if (c == ';' && paren < 1) flush();
                else value += c;
            }
        }
        if (!is_key) flush();  // a last value without ';'
    }
    bool has(const char *k) const { return f.count(k) != 0; }
    // first key present, value lower-cased (HSExecutionPlan.cpp:52-64)
    bool str(std::initializer_list<const char *> keys, std::string &out) const {
        for (const char *k : keys) {
            auto it = f.find(k);
            if (it != f.end()) {
                out = lower(it->second);
                return true;
            }
        }
        return false;
    }
    // value as written (names keep their case)
    bool text(std::initializer_list<const char *> keys, std::string &out) const {
        for (const char *k : keys) {
            auto it = f.find(k);
            if (it != f.end()) {
                out = it->second;
                return true;
            }
        }
        return false;
    }
    // ObjectParser::Get(double): std::stod; an unparsable value reads as absent
    bool num(const char *k, double &out) const {
        auto it = f.find(k);
        if (it == f.end()) return false;
        try {
            out = std::stod(it->second);
        } catch (...) {
            return false;
        }
        return true;
    }
    bool integer(const char *k, int &out) const {
        auto it = f.find(k);
        if (it == f.end()) return false;
        try {
            out = std::stoi(it->second);
        } catch (...) {
            return false;
        }
        return true;
    }
    // ObjectParser::Get(bool): true/yes/1, false/no/0 (MSDParser/ObjectParser.cpp:187-202)
    bool boolean(const char *k, bool &out) const {
        auto it = f.find(k);
        if (it == f.end()) return false;
        const std::string &v = it->second;
        if (v == "true" || v == "yes" || v == "1") out = true;
        else if (v == "false" || v == "no" || v == "0") out = false;
        else return false;
        return true;
    }
};

inline bool one_of(const std::string &v, std::initializer_list<const char *> c) {
    for (const char *x : c)
        if (v == x) return true;
    return false;
}

}  // namespace msg
