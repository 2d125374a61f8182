// rt_dict.h -- a small reader for OpenFOAM dictionary text.
//
// Accepts the entries the stress step reads: `key value;`,
// `key name [0 0 1 0 0 0 0] value;`, `key [dims] value;`, nested
// `name { ... }` sub-dictionaries, and C/C++ comments.
#pragma once

#include <map>
#include <string>
#include <vector>

namespace rt {

struct Dict {
    std::map<std::string, std::vector<std::string>> entries;  // tokens up to ';'
    std::map<std::string, Dict> subDicts;
    bool has(const std::string& k) const { return entries.count(k) || subDicts.count(k); }
    // scalar: the last numeric token of the entry (skips "name [dims]")
    bool scalar(const std::string& k, double& out) const;
    // word: the first token of the entry
    bool word(const std::string& k, std::string& out) const;
    // vector: the numbers of the entry's first parenthesised list, `key (x y z);`
    bool vector(const std::string& k, std::vector<double>& out) const;
    std::vector<std::string> keys() const;
    // a list of named sub-dictionaries, `key ( name1 { ... } name2 { ... } );`
    // (the `models` entry of multiMode); false when absent or malformed
    bool dictList(const std::string& k, std::vector<std::pair<std::string, Dict>>& out) const;
};

// Throws rt::Error(RT_E_BAD_ARG) on malformed text.
Dict parseDict(const std::string& text);

// If the dictionary holds a `parameters` (or other single) sub-dictionary and
// no `type`, return that sub-dictionary; otherwise the dictionary itself.
const Dict& parametersOf(const Dict& d);

}  // namespace rt
