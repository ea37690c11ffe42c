// ini.cpp -- parser of the reference's .ini parameter files.
#include "ini.h"

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace sp::ini {

static std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

static std::string strip_comment(const std::string& line) {
    bool in_q = false;
    for (size_t i = 0; i < line.size(); ++i) {
        const char c = line[i];
        if (c == '\'' || c == '"') in_q = !in_q;
        if (in_q) continue;
        if (c == '%') return line.substr(0, i);
        if (c == '/' && i + 1 < line.size() && line[i + 1] == '/') return line.substr(0, i);
    }
    return line;
}

File parse_text(const std::string& text, const std::string& dir) {
    File f;
    f.path = dir;
    std::istringstream in(text);
    std::string line, section;
    int lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        if (lineno == 1 && line.size() >= 3 && (unsigned char)line[0] == 0xEF) line = line.substr(3);   // UTF-8 BOM
        std::string s = trim(strip_comment(line));
        if (s.empty()) continue;
        if (s.front() == '[') {
            const size_t e = s.find(']');
            if (e == std::string::npos) throw std::runtime_error("ini: unterminated section header at line " + std::to_string(lineno));
            section = trim(s.substr(1, e - 1));
            f.sections[section];
            continue;
        }
        const size_t eq = s.find('=');
        if (eq == std::string::npos) throw std::runtime_error("ini: expected key = value at line " + std::to_string(lineno));
        const std::string key = trim(s.substr(0, eq));
        std::string val = trim(s.substr(eq + 1));
        if (val.size() >= 2 && (val.front() == '\'' || val.front() == '"') && val.back() == val.front())
            val = val.substr(1, val.size() - 2);
        f.sections[section][key] = val;
    }
    return f;
}

File parse_file(const std::string& filename) {
    std::ifstream in(filename, std::ios::binary);
    if (!in) throw std::runtime_error("ini: cannot open " + filename);
    std::stringstream ss;
    ss << in.rdbuf();
    const size_t slash = filename.find_last_of('/');
    return parse_text(ss.str(), slash == std::string::npos ? "." : filename.substr(0, slash));
}

const std::string* File::find(const std::string& section, const std::string& key) const {
    auto it = sections.find(section);
    if (it != sections.end()) {
        auto k = it->second.find(key);
        if (k != it->second.end()) return &k->second;
    }
    const std::string prefix = section + ".";
    for (auto& [name, kv] : sections) {
        if (name.compare(0, prefix.size(), prefix) != 0) continue;
        auto k = kv.find(key);
        if (k != kv.end()) return &k->second;
    }
    return nullptr;
}

std::vector<std::string> File::subsections(const std::string& section) const {
    std::vector<std::string> out;
    const std::string prefix = section + ".";
    for (auto& [name, kv] : sections)
        if (name.compare(0, prefix.size(), prefix) == 0) out.push_back(name.substr(prefix.size()));
    return out;
}

double to_double(const std::string& v) {
    char* end = nullptr;
    const double d = std::strtod(v.c_str(), &end);
    if (end == v.c_str()) throw std::runtime_error("ini: not a number: '" + v + "'");
    return d;
}

bool to_bool(const std::string& v) {
    std::string u;
    for (char c : v) u += (char)std::toupper((unsigned char)c);
    if (u == "TRUE" || u == "1") return true;
    if (u == "FALSE" || u == "0") return false;
    throw std::runtime_error("ini: not a boolean: '" + v + "'");
}

std::vector<double> to_vector(const std::string& v) {
    std::string s = trim(v);
    if (!s.empty() && s.front() == '{') {
        const size_t e = s.find('}');
        s = s.substr(1, (e == std::string::npos ? s.size() : e) - 1);
    }
    std::vector<double> out;
    std::stringstream ss(s);
    std::string item;
    while (std::getline(ss, item, ',')) {
        item = trim(item);
        if (!item.empty()) out.push_back(to_double(item));
    }
    return out;
}

}  // namespace sp::ini
