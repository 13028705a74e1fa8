// STL input (GeometryReader.h:19-20 interface) over lbm_read_stl_*.
#pragma once

#include <string>
#include <vector>

struct Triangle {  // Struct.h:77-82 field order
    double x1, y1, z1;
    double x2, y2, z2;
    double x3, y3, z3;
    double nx, ny, nz;
};

class Reader {
public:
    std::vector<Triangle> readSTLASCII(const std::string& filename);
    std::vector<Triangle> readSTLBinary(std::string filename);  // throws std::runtime_error
    // ASCII when the file starts with "solid" and parses as ASCII, else binary
    std::vector<Triangle> readSTL(const std::string& filename);
};
