// Reference-signature host shim: run parameters and small helpers.
//
// Same names, types and meaning as the reference's module/dataset.h:8-52.
#pragma once

#include <array>
#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

inline int ista, iend, iout, istep, nlstep, rstflag, nmat, npropmax;
inline double dt, dti, real_time, dlstep, facl = 1.0e0;
inline double mtol, dtol;
inline std::string parafile, gridfile, outfile, pointfile;
inline std::vector<int> npxye(3);
inline std::vector<double> bb(3), matflag, iprop, nprop;
inline std::vector<std::vector<double>> mat_prop;
// Voigt position of tensor entry (i, j): xx yy zz yz xz xy
inline std::array<std::array<int, 3>, 3> idm = {{{{0, 5, 4}}, {{5, 1, 3}}, {{4, 3, 2}}}};
inline std::array<std::array<double, 3>, 3> eye_mat = {
    {{{1.0e0, 0.0e0, 0.0e0}}, {{0.0e0, 1.0e0, 0.0e0}}, {{0.0e0, 0.0e0, 1.0e0}}}};

enum class MapScheme { PIC, FLIP, TPIC, APIC };

inline MapScheme ParseMapScheme(const std::string &s) {
    static const char *names[4] = {"PIC", "FLIP", "TPIC", "APIC"};
    for (int i = 0; i < 4; i++)
        if (s == names[i]) return static_cast<MapScheme>(i);
    throw std::invalid_argument("ParseMapScheme: unknown mapping scheme '" + s + "'");
}

// size entries, each equal to init_value.
template <typename T> void VectorAssign(int size, std::vector<T> &variable, T init_value = T()) {
    variable.assign(size, init_value);
}

// size value-initialised (zero) fixed-size arrays.
template <typename T, std::size_t N> void VectorAssign(int size, std::vector<std::array<T, N>> &variable) {
    variable.assign(size, std::array<T, N>{});
}
