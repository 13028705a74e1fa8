// Velocity sets in the reference's index order, usable on host and device.
// D3Q19: LatticeModel.cpp:51-106 (rest population last, index 18).
// D3Q27: LatticeModel.cpp:208-266 (rest population first, index 0).
// Weights are written as the reference writes them (2.0/36.0 etc.) so the
// constant-folded doubles are the same.
#pragma once

#include <hip/hip_runtime.h>

#define LBM_HD __host__ __device__ __forceinline__

template <int Q>
struct Lat;

template <>
struct Lat<19> {
    static constexpr int rest = 18;
    LBM_HD static constexpr int c(int d, int a) {
        constexpr int t[19][3] = {{1, 0, 0},   {-1, 0, 0},  {0, 1, 0},   {0, -1, 0},  {0, 0, 1},
                                  {0, 0, -1},  {1, 1, 0},   {1, -1, 0},  {1, 0, 1},   {1, 0, -1},
                                  {-1, 1, 0},  {-1, -1, 0}, {-1, 0, 1},  {-1, 0, -1}, {0, 1, 1},
                                  {0, 1, -1},  {0, -1, 1},  {0, -1, -1}, {0, 0, 0}};
        return t[d][a];
    }
    LBM_HD static constexpr int opp(int d) {
        constexpr int t[19] = {1, 0, 3, 2, 5, 4, 11, 10, 13, 12, 7, 6, 9, 8, 17, 16, 15, 14, 18};
        return t[d];
    }
    LBM_HD static constexpr double w(int d) {
        return d < 6 ? 2.0 / 36.0 : (d < 18 ? 1.0 / 36.0 : 12.0 / 36.0);
    }
};

template <>
struct Lat<27> {
    static constexpr int rest = 0;
    LBM_HD static constexpr int c(int d, int a) {
        constexpr int t[27][3] = {
            {0, 0, 0},   {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},   {0, -1, 0},  {0, 0, 1},   {0, 0, -1},
            {1, 1, 0},   {1, -1, 0},  {-1, 1, 0}, {-1, -1, 0}, {1, 0, 1},   {-1, 0, 1},  {1, 0, -1},
            {-1, 0, -1}, {0, 1, 1},   {0, 1, -1}, {0, -1, 1},  {0, -1, -1}, {1, 1, 1},   {-1, -1, -1},
            {1, 1, -1},  {-1, -1, 1}, {1, -1, 1}, {-1, 1, -1}, {1, -1, -1}, {-1, 1, 1}};
        return t[d][a];
    }
    LBM_HD static constexpr int opp(int d) {
        constexpr int t[27] = {0,  2,  1,  4,  3,  6,  5,  10, 9,  8,  7,  14, 13, 12,
                               11, 18, 17, 16, 15, 20, 19, 22, 21, 24, 23, 26, 25};
        return t[d];
    }
    LBM_HD static constexpr double w(int d) {
        return d == 0 ? 8.0 / 27.0 : (d < 7 ? 2.0 / 27.0 : (d < 19 ? 1.0 / 54.0 : 1.0 / 216.0));
    }
};

// Host-side runtime accessors (domain build, face tables).
inline int lat_c(int q, int d, int a) { return q == 19 ? Lat<19>::c(d, a) : Lat<27>::c(d, a); }
inline int lat_opp(int q, int d) { return q == 19 ? Lat<19>::opp(d) : Lat<27>::opp(d); }
inline double lat_w(int q, int d) { return q == 19 ? Lat<19>::w(d) : Lat<27>::w(d); }
inline int lat_rest(int q) { return q == 19 ? 18 : 0; }
