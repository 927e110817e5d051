// config.cpp -- see config.hpp (semantics of src/configreader.cpp).
#include "config.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace mth {

std::string first_word(const std::string &s) {
    const size_t b = s.find_first_not_of(" \t");
    if (b == std::string::npos) return "";
    const size_t e = s.find_first_of(" \t", b);
    return s.substr(b, e == std::string::npos ? std::string::npos : e - b);
}

std::string replace_first(const std::string &s, const std::string &replacement, const std::string &token) {
    const size_t p = s.find(token);
    if (p == std::string::npos) return s;
    return s.substr(0, p) + replacement + s.substr(p + token.size());
}

void Config::parse_file(const std::string &path, const std::vector<std::string> &overrides) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw ConfigError("ERROR: Parameters file '" + path + "' can not be found.");
    std::stringstream ss;
    ss << in.rdbuf();
    parse_text(ss.str(), overrides);
}

void Config::parse_text(const std::string &text, const std::vector<std::string> &overrides) {
    entries_.clear();
    size_t pos = 0;
    while (pos < text.size()) {
        size_t nl = text.find('\n', pos);
        const std::string line = text.substr(pos, nl == std::string::npos ? std::string::npos : nl - pos);
        pos = nl == std::string::npos ? text.size() : nl + 1;
        // an empty line is "\n" to fgets: skipped like a comment
        if (line.empty() || line[0] == '#' || line[0] == ' ' || line[0] == '\t') continue;
        const size_t d = line.find_first_of(" \t");
        if (d == std::string::npos) throw ConfigError("ERROR: configuration line without a value: '" + line + "'");
        const std::string name = line.substr(0, d);
        size_t v = d + 1;
        while (v < line.size() && line[v] == '#') v++;
        const size_t ve = line.find('#', v);
        const std::string value = line.substr(v, ve == std::string::npos ? std::string::npos : ve - v);
        if (value.empty()) throw ConfigError("ERROR: configuration line without a value: '" + line + "'");
        entries_.emplace_back(name, value);
    }
    apply_overrides(overrides);
}

void Config::apply_overrides(const std::vector<std::string> &overrides) {
    for (const std::string &o : overrides) {
        const size_t eq = o.find('=');
        if (eq == std::string::npos) continue;
        set(o.substr(0, eq), o.substr(eq + 1), true);
    }
}

bool Config::set(const std::string &name, const std::string &value, bool force) {
    for (auto &kv : entries_)
        if (kv.first == name) {
            kv.second = value;
            return true;
        }
    if (force) entries_.emplace_back(name, value);
    return force;
}

const std::string *Config::find(const std::string &name) const {
    for (const auto &kv : entries_)
        if (kv.first == name) return &kv.second;
    return nullptr;
}

std::string Config::lookup(const std::string &name, const std::string *def) const {
    if (const std::string *v = find(name)) return *v;
    if (def) return *def;
    throw ConfigError("Parameter '" + name + "' should be specified in a configuration file.");
}

std::string Config::apply_mask(const std::string &mask) const {
    std::string s = mask;
    for (const auto &kv : entries_) {
        const std::string key = "<" + kv.first + ">";
        if (s.find(key) == std::string::npos) continue;
        const std::string w = first_word(kv.second);
        if (w.empty()) throw ConfigError("ERROR: parameter '" + kv.first + "' has no value to substitute");
        s = replace_first(s, w, key);
    }
    return s;
}

std::string Config::masked_impl(const std::string &name, const std::string *def) const {
    const std::string raw = lookup(name, def);
    const std::string w = first_word(raw);
    if (w.empty()) throw ConfigError("ERROR: parameter '" + name + "' has an empty value");
    return apply_mask(w);
}

std::string Config::masked_replace(const std::string &name, const std::string &replacement,
                                   const std::string &token) const {
    return replace_first(masked(name), replacement, token);
}

int Config::int_impl(const std::string &name, const int *def) const {
    const std::string d = def ? std::to_string(*def) : std::string();
    const std::string v = masked_impl(name, def ? &d : nullptr);
    const int r = std::atoi(v.c_str());
    if (r == 0 && v != "0")
        throw ConfigError("ERROR: Wrong value of " + name + " in a configuration file ('" + v +
                          "'). Should be integer.");
    return r;
}

long long Config::long_impl(const std::string &name, const long *def) const {
    const std::string d = def ? std::to_string(*def) : std::string();
    const std::string v = masked_impl(name, def ? &d : nullptr);
    const long long r = std::atol(v.c_str());
    if (r == 0 && v != "0")
        throw ConfigError("ERROR: Wrong value of " + name + " in a configuration file ('" + v +
                          "'). Should be long integer.");
    return r;
}

float Config::float_impl(const std::string &name, const float *def) const {
    std::string d;
    if (def) {
        char buf[64];
        std::snprintf(buf, sizeof buf, "%f", (double)*def);
        d = buf;
    }
    const std::string v = masked_impl(name, def ? &d : nullptr);
    const float r = (float)std::atof(v.c_str());
    if (r == 0.0f && v != "0" && v != "0.0" && v != "0.0f" && v != "0.000000")
        throw ConfigError("ERROR: Wrong value of " + name + " in a configuration file ('" + v +
                          "'). Should be float.");
    return r;
}

bool Config::yesno_impl(const std::string &name, const bool *def) const {
    const std::string d = def ? (*def ? "YES" : "NO") : "";
    const std::string v = masked_impl(name, def ? &d : nullptr);
    static const char *yes[] = {"YES", "Yes", "yes", "Y", "y", "ON", "On", "on", "TRUE", "True", "true"};
    for (const char *y : yes)
        if (v == y) return true;
    return false;
}

}  // namespace mth
