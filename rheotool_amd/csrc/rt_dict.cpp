// rt_dict.cpp -- OpenFOAM dictionary text reader (see rt_dict.h).
#include "rt_dict.h"

#include <cctype>
#include <cstdlib>

#include "rt_internal.h"

namespace rt {

namespace {

std::vector<std::string> tokenize(const std::string& s) {
    std::vector<std::string> t;
    size_t i = 0, n = s.size();
    while (i < n) {
        char c = s[i];
        if (std::isspace((unsigned char)c)) {
            ++i;
        } else if (c == '/' && i + 1 < n && s[i + 1] == '/') {
            while (i < n && s[i] != '\n') ++i;
        } else if (c == '/' && i + 1 < n && s[i + 1] == '*') {
            size_t e = s.find("*/", i + 2);
            i = e == std::string::npos ? n : e + 2;
        } else if (c == '{' || c == '}' || c == ';' || c == '[' || c == ']' || c == '(' ||
                   c == ')') {
            t.emplace_back(1, c);
            ++i;
        } else if (c == '"') {
            size_t e = s.find('"', i + 1);
            if (e == std::string::npos) throw Error{RT_E_BAD_ARG, "unterminated string in dictionary"};
            t.push_back(s.substr(i + 1, e - i - 1));
            i = e + 1;
        } else {
            size_t j = i;
            while (j < n && !std::isspace((unsigned char)s[j]) && s[j] != '{' && s[j] != '}' &&
                   s[j] != ';' && s[j] != '[' && s[j] != ']' && s[j] != '(' && s[j] != ')')
                ++j;
            t.push_back(s.substr(i, j - i));
            i = j;
        }
    }
    return t;
}

void parseInto(const std::vector<std::string>& t, size_t& i, Dict& d, bool top) {
    while (i < t.size()) {
        if (t[i] == "}") {
            if (top) throw Error{RT_E_BAD_ARG, "unbalanced '}' in dictionary"};
            ++i;
            return;
        }
        if (t[i] == ";") {
            ++i;
            continue;
        }
        std::string key = t[i++];
        if (i < t.size() && t[i] == "{") {
            ++i;
            parseInto(t, i, d.subDicts[key], false);
            continue;
        }
        std::vector<std::string> vals;
        int depth = 0;
        while (i < t.size() && !(depth == 0 && t[i] == ";")) {
            if (t[i] == "[" || t[i] == "(") ++depth;
            if (t[i] == "]" || t[i] == ")") --depth;
            if (depth == 0 && t[i] == "}") break;
            vals.push_back(t[i++]);
        }
        if (i < t.size() && t[i] == ";") ++i;
        d.entries[key] = vals;
    }
    if (!top) throw Error{RT_E_BAD_ARG, "unbalanced '{' in dictionary"};
}

bool isNumber(const std::string& s, double& v) {
    if (s.empty()) return false;
    char* end = nullptr;
    v = std::strtod(s.c_str(), &end);
    return end && *end == '\0';
}

}  // namespace

bool Dict::scalar(const std::string& k, double& out) const {
    auto it = entries.find(k);
    if (it == entries.end()) return false;
    int depth = 0;
    bool found = false;
    for (const auto& tok : it->second) {
        if (tok == "[" || tok == "(") ++depth;
        else if (tok == "]" || tok == ")") --depth;
        else if (depth == 0) {
            double v;
            if (isNumber(tok, v)) {
                out = v;
                found = true;
            }
        }
    }
    return found;
}

bool Dict::word(const std::string& k, std::string& out) const {
    auto it = entries.find(k);
    if (it == entries.end() || it->second.empty()) return false;
    out = it->second.front();
    return true;
}

bool Dict::vector(const std::string& k, std::vector<double>& out) const {
    auto it = entries.find(k);
    if (it == entries.end()) return false;
    out.clear();
    bool open = false;
    for (const auto& tok : it->second) {
        if (tok == "(") {
            if (open) return false;  // nested lists are not vectors
            open = true;
        } else if (tok == ")") {
            return open;
        } else if (open) {
            double v;
            if (!isNumber(tok, v)) return false;
            out.push_back(v);
        }
    }
    return false;
}

bool Dict::dictList(const std::string& k, std::vector<std::pair<std::string, Dict>>& out) const {
    auto it = entries.find(k);
    if (it == entries.end()) return false;
    const std::vector<std::string>& t = it->second;
    out.clear();
    size_t i = 0;
    if (i >= t.size() || t[i] != "(") return false;
    ++i;
    while (i < t.size() && t[i] != ")") {
        std::string name = t[i++];
        if (i >= t.size() || t[i] != "{") return false;
        ++i;
        Dict sub;
        parseInto(t, i, sub, false);  // consumes the closing '}'
        out.emplace_back(name, std::move(sub));
    }
    return i < t.size();
}

std::vector<std::string> Dict::keys() const {
    std::vector<std::string> k;
    for (auto& e : entries) k.push_back(e.first);
    for (auto& e : subDicts) k.push_back(e.first);
    return k;
}

Dict parseDict(const std::string& text) {
    auto t = tokenize(text);
    Dict d;
    size_t i = 0;
    parseInto(t, i, d, true);
    return d;
}

const Dict& parametersOf(const Dict& d) {
    if (d.entries.count("type")) return d;
    auto it = d.subDicts.find("parameters");
    if (it != d.subDicts.end()) return it->second;
    return d;
}

}  // namespace rt
