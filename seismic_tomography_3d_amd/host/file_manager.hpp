// file_manager.hpp -- parameter file and raw FP32 volume I/O with the
// reference's formats (src/utils/file_manager/file_manager.hpp:7-17).
#pragma once

#include <string>
#include <vector>

namespace st_host {

// "key = value # comment" lines (file_manager.cpp:74-120): the first line
// that starts with `target` and not with '#' or ' '; the value runs from two
// characters after '=' up to '#'; quotes, brackets and spaces are removed.
std::string catch_parameter(const std::string &target, const std::string &file);
bool str2bool(std::string s);                                    // :9-17
void read_binary_float(const std::string &path, float *array, long n);   // :19-33
void write_binary_float(const std::string &path, const float *array, long n);  // :35-51
void read_text_file(const std::string &path, std::vector<std::string> &elements);  // :53-72
std::vector<std::string> split(const std::string &s, char delimiter);  // :122-132

}  // namespace st_host
