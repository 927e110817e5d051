// config.hpp -- the reference's key/value configuration files, read with the
// same semantics as src/configreader.cpp (the format its users' files are
// written in):
//   * a line is an entry unless it starts with '#', ' ', '\t' or is empty;
//     the name runs to the first space or tab, the value is the rest of the
//     line up to '#' or the end (leading '#' characters skipped, as strtok
//     with "#\n" does, configreader.cpp:47-50);
//   * each parsed file replaces the table (:31), then "name=value" overrides
//     from the command line are applied to it (:60-74);
//   * lookups return the first entry of that name (:121-129);
//   * "masked" values are the first word of the value with every "<key>"
//     replaced (first occurrence, table order) by that key's first word
//     (:240-249, :314-325);
//   * integer/float parses reject values that parse to zero unless spelled
//     as zero (:139-141, :195-198); yes/no accepts only the textual forms of
//     :217-220.
// Where the reference dies (exit), this reader throws ConfigError.
#pragma once

#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace mth {

struct ConfigError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

class Config {
public:
    void parse_file(const std::string &path, const std::vector<std::string> &overrides = {});
    void parse_text(const std::string &text, const std::vector<std::string> &overrides = {});

    bool has(const std::string &name) const { return find(name) != nullptr; }
    // setParameter: first entry of that name, else appended when `force`.
    bool set(const std::string &name, const std::string &value, bool force);

    std::string str(const std::string &name) const { return lookup(name, nullptr); }
    std::string str(const std::string &name, const std::string &def) const { return lookup(name, &def); }
    std::string masked(const std::string &name) const { return masked_impl(name, nullptr); }
    std::string masked(const std::string &name, const std::string &def) const { return masked_impl(name, &def); }
    // getMaskedParameterWithReplacement: masked value with the first
    // occurrence of `token` replaced (the reference's "<run>" file masks).
    std::string masked_replace(const std::string &name, const std::string &replacement,
                               const std::string &token) const;

    int integer(const std::string &name) const { return int_impl(name, nullptr); }
    int integer(const std::string &name, int def) const { return int_impl(name, &def); }
    long long long_integer(const std::string &name) const { return long_impl(name, nullptr); }
    long long long_integer(const std::string &name, long def) const { return long_impl(name, &def); }
    float real(const std::string &name) const { return float_impl(name, nullptr); }
    float real(const std::string &name, float def) const { return float_impl(name, &def); }
    bool yesno(const std::string &name) const { return yesno_impl(name, nullptr); }
    bool yesno(const std::string &name, bool def) const { return yesno_impl(name, &def); }

    const std::vector<std::pair<std::string, std::string>> &entries() const { return entries_; }

private:
    const std::string *find(const std::string &name) const;
    std::string lookup(const std::string &name, const std::string *def) const;
    std::string masked_impl(const std::string &name, const std::string *def) const;
    std::string apply_mask(const std::string &mask) const;
    int int_impl(const std::string &name, const int *def) const;
    long long long_impl(const std::string &name, const long *def) const;
    float float_impl(const std::string &name, const float *def) const;
    bool yesno_impl(const std::string &name, const bool *def) const;
    void apply_overrides(const std::vector<std::string> &overrides);

    std::vector<std::pair<std::string, std::string>> entries_;
};

// First space/tab-separated word of s ("" when there is none).
std::string first_word(const std::string &s);
// replaceString: the first occurrence of `token` in s replaced.
std::string replace_first(const std::string &s, const std::string &replacement, const std::string &token);

}  // namespace mth
