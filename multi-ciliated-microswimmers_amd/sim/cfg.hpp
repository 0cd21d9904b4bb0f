// Reader and writer for the configuration files the reference's `cilia`
// executable takes (src/main.cpp:40-100, libconfig++ `Config::readFile` /
// `writeFile`).  The grammar is the part of libconfig's that those files
// use, as written by pylibconfig2 (sr/sim_cfg.py:123-133) or by hand:
//
//   file     := setting*
//   setting  := name ('=' | ':') value [';' | ',']
//   value    := scalar | '[' scalars ']' | '(' values ')' | '{' setting* '}'
//   scalar   := bool | integer | float | string ("..." with C escapes;
//               adjacent strings are joined)
//   comments := '#' or '//' to the end of the line, '/* ... */'
//
// Lookups follow libconfig++'s behaviour where the reference relies on it:
// a missing setting is an error naming its path (SettingNotFoundException,
// src/main.cpp:88-92); an integer where a float is read is accepted (the
// reverse is an error, as without auto-conversion).
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

namespace mcm {

struct CfgError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct Setting {
    enum Kind { Group, Array, List, Int, Float, Bool, String };
    Kind kind = Group;
    std::string name;  // empty for elements of arrays and lists
    std::string path;  // dotted path from the root, for messages
    long long ival = 0;
    double fval = 0.0;
    bool bval = false;
    std::string sval;
    std::vector<Setting> items;  // members of a group, elements of an array or list

    bool is_number() const { return kind == Int || kind == Float; }
    bool is_aggregate() const { return kind == Group || kind == Array || kind == List; }
    int length() const { return (int)items.size(); }

    bool exists(const std::string &key) const;
    const Setting &operator[](const std::string &key) const;  // group member; error if missing
    const Setting &operator[](int i) const;                   // element; error if out of range
    Setting *find(const std::string &key);

    double as_double() const;
    long long as_int() const;
    bool as_bool() const;
    const std::string &as_string() const;

    // lookupValue: false (and `out` untouched) when missing or of another type
    bool lookup(const std::string &key, double &out) const;
    bool lookup(const std::string &key, std::string &out) const;

    void set_double(const std::string &key, double v);  // replace a member's value; error if missing
};

// Parses a file; a syntax error names the file and line (ParseException,
// src/main.cpp:49-53).
Setting cfg_read_file(const std::string &path);
Setting cfg_parse(const std::string &text, const std::string &origin = "<string>");
// Writes the tree back in libconfig syntax, floats with 17 significant digits.
std::string cfg_format(const Setting &root);
void cfg_write_file(const Setting &root, const std::string &path);

}  // namespace mcm
