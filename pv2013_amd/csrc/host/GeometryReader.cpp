#include "GeometryReader.h"

#include <fstream>
#include <iostream>
#include <stdexcept>

#include "lbm_capi.h"

static_assert(sizeof(Triangle) == 12 * sizeof(double), "Triangle must be 12 packed doubles");

std::vector<Triangle> Reader::readSTLASCII(const std::string& filename) {
    std::vector<Triangle> t;
    const int64_t n = lbm_read_stl_ascii(filename.c_str(), nullptr, 0);
    if (n < 0) {
        std::cerr << "Error: cannot open file " << filename << std::endl;  // GeometryReader.cpp:37
        return t;
    }
    t.resize((size_t)n);
    lbm_read_stl_ascii(filename.c_str(), reinterpret_cast<double*>(t.data()), n);
    return t;
}

std::vector<Triangle> Reader::readSTLBinary(std::string filename) {
    const int64_t n = lbm_read_stl_binary(filename.c_str(), nullptr, 0);
    if (n < 0) throw std::runtime_error("Failed to open file: " + filename);  // GeometryReader.cpp:124
    std::vector<Triangle> t((size_t)n);
    lbm_read_stl_binary(filename.c_str(), reinterpret_cast<double*>(t.data()), n);
    return t;
}

std::vector<Triangle> Reader::readSTL(const std::string& filename) {
    std::ifstream in(filename, std::ios::binary);
    char head[6] = {0};
    in.read(head, 5);
    if (std::string(head) == "solid") {
        std::vector<Triangle> t = readSTLASCII(filename);
        if (!t.empty()) return t;
    }
    return readSTLBinary(filename);
}
