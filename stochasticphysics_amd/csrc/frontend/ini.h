// ini.h -- reader of the reference's parameter files (SerAr ConfigFile syntax).
//
// Syntax accepted (Examples/NeelSpherical/*.ini): [Section] and dotted
// subsections [A.B]; key = value; strings in single quotes; vectors {x, y, z};
// TRUE/FALSE; comments introduced by // or % (outside quotes).
#pragma once
#include <map>
#include <string>
#include <vector>

namespace sp::ini {

struct File {
    // section path -> key -> raw value text (quotes removed for strings)
    std::map<std::string, std::map<std::string, std::string>> sections;
    std::string path;   // directory of the file (for relative particle files)

    bool has_section(const std::string& s) const { return sections.count(s) != 0; }
    // value of key in section or any of its subsections (first match in map order)
    const std::string* find(const std::string& section, const std::string& key) const;
    std::vector<std::string> subsections(const std::string& section) const;
};

File parse_file(const std::string& filename);          // throws std::runtime_error
File parse_text(const std::string& text, const std::string& dir = ".");

double to_double(const std::string& v);
bool to_bool(const std::string& v);
std::vector<double> to_vector(const std::string& v);    // "{a, b, c}" or scalar

}  // namespace sp::ini
