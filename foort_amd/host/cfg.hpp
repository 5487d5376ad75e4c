// cfg.hpp -- reader for the configuration-file syntax FOORT uses (the
// libconfig grammar: groups { }, settings `name = value;` or `name: value,`,
// integers, floats, booleans, strings, // # and /* */ comments).  Lookups are
// typed the way FOORT's lookups behave (FOORT/Config.cpp and its
// documentation): a float setting is only read as a float ("1000" is not a
// valid SphereRadius, "1000.0" is), an integer only as an integer, a boolean
// only as true/false, a string only as a string; a lookup of the wrong type
// fails and the caller keeps its default.  Integer literals that do not fit in
// 32 bits, or carry an L suffix, are 64-bit integers; an `int` lookup does not
// accept them.  Counter lookups follow Config::lookupValuelargecounter
// (FOORT/Config.cpp:44-88) over the library's typed getters: a 64-bit integer
// is taken when non-negative, a 32-bit integer through the unsigned getter,
// which turns negative values into 0.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace foort_host {

struct Setting {
    enum Type { Group, Int, Int64, Float, Bool, String, List } type = Group;
    long long i = 0;
    double f = 0;
    bool b = false;
    std::string s;
    std::vector<std::pair<std::string, std::unique_ptr<Setting>>> children; // group members in file order
    std::vector<std::unique_ptr<Setting>> items;                            // list / array elements

    const Setting *child(const std::string &name) const;
    bool exists(const std::string &name) const { return child(name) != nullptr; }
    // typed lookups: return false (value untouched) when absent or of another type
    bool lookup(const std::string &name, double &v) const;
    bool lookup(const std::string &name, bool &v) const;
    bool lookup(const std::string &name, std::string &v) const;
    bool lookup(const std::string &name, int &v) const;
    bool lookup_counter(const std::string &name, unsigned long &v) const;
};

struct ParseError {
    int line;
    std::string message;
};

// Parse a configuration text; throws ParseError.
std::unique_ptr<Setting> parse_config(const std::string &text);
// Read and parse a file; throws ParseError (line 0 = the file could not be read).
std::unique_ptr<Setting> parse_config_file(const std::string &path);

} // namespace foort_host
