// file_manager.cpp -- see file_manager.hpp.
#include "file_manager.hpp"

#include <algorithm>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>

namespace st_host {

std::string catch_parameter(const std::string &target, const std::string &file) {
    std::ifstream in(file);
    std::string line, value;
    if (in.is_open()) {
        while (std::getline(in, line)) {
            if (line.empty() || line.front() == '#' || line.front() == ' ') continue;
            if (line.find(target) != 0) continue;
            const size_t eq = line.find('=');
            if (eq == std::string::npos) break;
            for (size_t i = eq + 2; i < line.size(); i++) {
                if (line[i] == '#') break;
                value += line[i];
            }
            break;
        }
    }
    // the reference strips quotes (paths) and brackets (coordinate pairs)
    // before removing spaces; the effective result is the value without any
    // of the three characters
    value.erase(std::remove_if(value.begin(), value.end(),
                               [](char c) { return c == '"' || c == '[' || c == ']' || c == ' ' ||
                                                   c == '\t' || c == '\r'; }),
                value.end());
    return value;
}

bool str2bool(std::string s) {
    std::transform(s.begin(), s.end(), s.begin(), ::tolower);
    bool b = false;
    std::istringstream(s) >> std::boolalpha >> b;
    return b;
}

void read_binary_float(const std::string &path, float *array, long n) {
    std::ifstream f(path, std::ios::in | std::ios::binary);
    if (!f.is_open()) throw std::invalid_argument("Error: file could not be opened: " + path);
    f.read(reinterpret_cast<char *>(array), n * sizeof(float));
}

void write_binary_float(const std::string &path, const float *array, long n) {
    std::ofstream f(path, std::ios::out | std::ios::binary);
    if (!f.is_open()) throw std::invalid_argument("Error: file could not be opened: " + path);
    f.write(reinterpret_cast<const char *>(array), n * sizeof(float));
    std::cout << "Binary file " << path << " was succesfully written." << std::endl;
}

void read_text_file(const std::string &path, std::vector<std::string> &elements) {
    std::ifstream f(path);
    if (!f.is_open()) throw std::invalid_argument("Error: file could not be opened: " + path);
    std::string line;
    while (std::getline(f, line)) elements.push_back(line);
}

std::vector<std::string> split(const std::string &s, char delimiter) {
    std::vector<std::string> tokens;
    std::string token;
    std::istringstream ts(s);
    while (std::getline(ts, token, delimiter)) tokens.push_back(token);
    return tokens;
}

}  // namespace st_host

// C entry for tests: catch_parameter into a caller buffer, returns the length
extern "C" int sth_catch_parameter(const char *target, const char *file, char *out, int cap) {
    const std::string v = st_host::catch_parameter(target, file);
    const int n = (int)std::min<size_t>(v.size(), (size_t)(cap > 0 ? cap - 1 : 0));
    std::copy(v.begin(), v.begin() + n, out);
    if (cap > 0) out[n] = 0;
    return (int)v.size();
}
