#include "icsf_dict.h"

#include <cctype>
#include <cstdlib>

namespace icsf {

namespace {

struct Lexer {
    const std::string& s;
    size_t i = 0;
    explicit Lexer(const std::string& t) : s(t) {}
    void skip() {
        for (;;) {
            while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
            if (i + 1 < s.size() && s[i] == '/' && s[i + 1] == '/') {
                while (i < s.size() && s[i] != '\n') ++i;
            } else if (i + 1 < s.size() && s[i] == '/' && s[i + 1] == '*') {
                i += 2;
                while (i + 1 < s.size() && !(s[i] == '*' && s[i + 1] == '/')) ++i;
                i = i + 2 <= s.size() ? i + 2 : s.size();
            } else {
                return;
            }
        }
    }
    // tokens: '{' '}' ';' or a word (parenthesised groups kept whole, e.g.
    // "reconstruct(rho)", "(0 0 1)")
    bool next(std::string& tok) {
        skip();
        if (i >= s.size()) return false;
        const char c = s[i];
        if (c == '{' || c == '}' || c == ';') {
            tok = std::string(1, c);
            ++i;
            return true;
        }
        if (c == '"') {
            size_t j = s.find('"', i + 1);
            if (j == std::string::npos) j = s.size();
            tok = s.substr(i + 1, j - i - 1);
            i = j + 1;
            return true;
        }
        size_t j = i;
        int depth = 0;
        while (j < s.size()) {
            const char d = s[j];
            if (d == '(') ++depth;
            else if (d == ')') { if (depth > 0) --depth; }
            else if (depth == 0 && (std::isspace((unsigned char)d) || d == ';' || d == '{' || d == '}')) break;
            ++j;
        }
        tok = s.substr(i, j - i);
        i = j;
        return true;
    }
};

std::string parseBody(Lexer& lx, Dict& d, bool top) {
    std::string tok;
    while (lx.next(tok)) {
        if (tok == "}") {
            if (top) return "unexpected '}'";
            return "";
        }
        if (tok == ";") continue;
        const std::string key = tok;
        std::vector<std::string> vals;
        bool done = false;
        while (!done) {
            if (!lx.next(tok)) return "unexpected end of input after '" + key + "'";
            if (tok == "{") {
                auto sub = std::make_shared<Dict>();
                std::string err = parseBody(lx, *sub, false);
                if (!err.empty()) return err;
                d.subs[key] = sub;
                d.order.push_back(key);
                done = true;
            } else if (tok == ";") {
                d.entries[key] = vals;
                d.order.push_back(key);
                done = true;
            } else if (tok == "}") {
                return "missing ';' after '" + key + "'";
            } else {
                vals.push_back(tok);
            }
        }
    }
    return top ? "" : "missing '}'";
}

}  // namespace

const Dict* Dict::sub(const std::string& name) const {
    auto it = subs.find(name);
    return it == subs.end() ? nullptr : it->second.get();
}

const std::vector<std::string>* Dict::find(const std::string& path) const {
    const Dict* d = this;
    size_t s = 0;
    for (;;) {
        const size_t p = path.find('/', s);
        if (p == std::string::npos) {
            auto it = d->entries.find(path.substr(s));
            return it == d->entries.end() ? nullptr : &it->second;
        }
        d = d->sub(path.substr(s, p - s));
        if (!d) return nullptr;
        s = p + 1;
    }
}

std::string Dict::word(const std::string& path, const std::string& def) const {
    const auto* v = find(path);
    return (v && !v->empty()) ? (*v)[0] : def;
}

double Dict::scalar(const std::string& path, double def) const {
    const auto* v = find(path);
    if (!v || v->empty()) return def;
    char* end = nullptr;
    const double x = std::strtod((*v)[0].c_str(), &end);
    return (end && *end == 0) ? x : def;
}

bool Dict::sw(const std::string& path, bool def) const {
    const auto* v = find(path);
    if (!v || v->empty()) return def;
    const std::string& w = (*v)[0];
    if (w == "true" || w == "on" || w == "yes" || w == "y" || w == "t" || w == "1") return true;
    if (w == "false" || w == "off" || w == "no" || w == "n" || w == "f" || w == "0" || w == "none") return false;
    return def;
}

std::string parseDict(const std::string& text, Dict& out) {
    Lexer lx(text);
    return parseBody(lx, out, true);
}

}  // namespace icsf
