// Reference-signature host shim: shape functions.
//
// MakSf with the signature of the reference's module/shape_function.h:10-11
// (definition module/shape_function.cpp:8-119): for the particle at xyg in
// cell m, the stencil's control-point ids nc, weights sf and gradients dsf
// (resized to (idimc0+1)(idimc1+1)(idimc2+1), nenode of them), stencil
// order k, j, i; idimc 1 = tent functions, 2 = open quadratic B-spline.
// Reads the globals xyelem, dxy, xymin, aelemmin, xynodecw of mesh.h.
#pragma once

#include <array>
#include <vector>

#include "cal_mat.h"
#include "dataset.h"
#include "mesh.h"

void MakSf(const int m, const std::array<double, 3> &xyg, const std::vector<int> &idimc,
           const std::vector<int> &xynodec, std::vector<int> &nc, int &nenode, std::vector<double> &sf,
           std::vector<std::array<double, 3>> &dsf);
