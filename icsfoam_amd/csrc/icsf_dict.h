// Minimal reader for OpenFOAM dictionary syntax (fvSchemes, fvSolution,
// thermophysicalProperties): nested `name { ... }` blocks and
// `keyword token token ... ;` entries, C/C++ comments, FoamFile headers.
// Only the keywords the implicit coupled step reads are interpreted.
#pragma once
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace icsf {

struct Dict {
    std::map<std::string, std::vector<std::string>> entries;
    std::map<std::string, std::shared_ptr<Dict>> subs;
    std::vector<std::string> order;  // keyword order as read

    const Dict* sub(const std::string& name) const;
    bool has(const std::string& key) const { return entries.count(key) > 0; }
    // entry lookup by "a/b/c" path; returns nullptr when absent
    const std::vector<std::string>* find(const std::string& path) const;
    std::string word(const std::string& path, const std::string& def) const;
    double scalar(const std::string& path, double def) const;
    bool sw(const std::string& path, bool def) const;  // OpenFOAM Switch
};

// Parse; returns an empty string on success, otherwise the error text.
std::string parseDict(const std::string& text, Dict& out);

}  // namespace icsf
