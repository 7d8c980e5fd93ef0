// Global.hpp -- compile-time constants and small helpers of the host classes
// (the subset of the reference's includes/Global.hpp:20-69 they use).
#pragma once
#include <cmath>

#define NDIMS 2
#define NVARS 4

template <typename T>
void setarr(int i1, int i2, T val, T* arr) {
    for (int i = i1; i < i2; i++) arr[i] = val;
}

inline double dot(const double x1[NDIMS], const double x2[NDIMS]) { return x1[0] * x2[0] + x1[1] * x2[1]; }
