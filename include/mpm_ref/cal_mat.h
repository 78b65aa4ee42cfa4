// Reference-signature host shim: 3-vector and 3x3 matrix helpers with the
// signatures of the reference's module/cal_mat.h:10-77.
//
// The determinant is the cofactor expansion along the first row with the
// reference's grouping of the 2x2 minors (cal_mat.h:38-46), and the inverse
// is the adjugate times 1/det without a singularity check (:48-66): results
// are bit-identical to the reference's on the same compiler flags
// (-ffp-contract=off).
#pragma once

#include <array>
#include <cmath>
#include <utility>

#include "dataset.h"
#include "mesh.h"

template <typename T> T Sign(T val) { return (T(0) < val) - (val < T(0)); }

template <typename T> double NormVec3(const std::array<T, 3> &vec) {
    double s = 0.0e0;
    for (const T &v : vec) s += v * v;
    return std::sqrt(s);
}

template <typename T> double TraceMat3(const std::array<std::array<T, 3>, 3> &mat) {
    double s = 0.0e0;
    for (int i = 0; i < 3; i++) s += mat[i][i];
    return s;
}

namespace mpm_ref_detail {
// 2x2 minor a*d - b*c of rows r0, r1 and columns c0, c1 as written there.
template <typename T>
inline double minor2(const std::array<std::array<T, 3>, 3> &m, int r0, int c0, int r1, int c1) {
    return m[r0][c0] * m[r1][c1] - m[r0][c1] * m[r1][c0];
}
}  // namespace mpm_ref_detail

template <typename T> double DetMat3(const std::array<std::array<T, 3>, 3> &mat) {
    using mpm_ref_detail::minor2;
    // a00 (a11 a22 - a12 a21) + a01 (a12 a20 - a10 a22) + a02 (a10 a21 - a11 a20)
    return mat[0][0] * minor2(mat, 1, 1, 2, 2) + mat[0][1] * minor2(mat, 1, 2, 2, 0) +
           mat[0][2] * minor2(mat, 1, 0, 2, 1);
}

template <typename T> std::array<std::array<double, 3>, 3> InvMat3(const std::array<std::array<T, 3>, 3> &mat) {
    using mpm_ref_detail::minor2;
    const double rdet = 1.0e0 / DetMat3(mat);
    std::array<std::array<double, 3>, 3> inv;
    // adjugate entry (i, j) = cofactor (j, i), as minors of the rows and
    // columns other than j and i, in the reference's operand order
    inv[0][0] = minor2(mat, 1, 1, 2, 2) * rdet;
    inv[0][1] = minor2(mat, 0, 2, 2, 1) * rdet;
    inv[0][2] = minor2(mat, 0, 1, 1, 2) * rdet;
    inv[1][0] = minor2(mat, 1, 2, 2, 0) * rdet;
    inv[1][1] = minor2(mat, 0, 0, 2, 2) * rdet;
    inv[1][2] = minor2(mat, 0, 2, 1, 0) * rdet;
    inv[2][0] = minor2(mat, 1, 0, 2, 1) * rdet;
    inv[2][1] = minor2(mat, 0, 1, 2, 0) * rdet;
    inv[2][2] = minor2(mat, 0, 0, 1, 1) * rdet;
    return inv;
}

template <typename T> void TransMat3(std::array<std::array<T, 3>, 3> &mat) {
    std::swap(mat[0][1], mat[1][0]);
    std::swap(mat[0][2], mat[2][0]);
    std::swap(mat[1][2], mat[2][1]);
}
