// Shared device-side definitions of the MI355X explicit-MPM engine.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpm {

// Background grid as seen by the kernels (single rank: xymin == xyminw).
// Node/cell numbering follows the reference: cell = ix + iy*nex + iz*nex*ney
// (module/material_point.cpp:65), node = i + nxc*j + nxc*nyc*k
// (module/shape_function.cpp:34).
struct GridGeom {
    double xmin[3];
    double xmax[3];
    double dx[3];
    int nex[3];
    int nxc[3];
    int idimc[3];
    int nelem;
    int nodec;
};

// Dirichlet BCs in dense form: bit c of mask[n] set when component c of node
// n is prescribed; val[c*nodec + n] is the prescribed value before facl.
struct BcDense {
    const uint8_t *mask;
    const double *val;
};

struct MaterialDev {
    int model;
    float dt;  // time step of the particles of this material when per-material steps are on, else 0
    double prop[8];
};

constexpr int kMaxMat = 1024;

// materials in device memory, indexed by particle matid
struct MaterialTable {
    const MaterialDev *m;
    int nmat;
};

// Device-side error flags (bitwise OR of these).
enum : uint32_t {
    kErrOutOfGrid = 1u,
    kErrNonFinite = 2u,
    kErrJacobian = 4u,
    kErrCellOverflow = 8u,  // more than 4096 particles in one cell
    kErrBinCounts = 16u,    // per-cell counts add up to more than the particle count
};

// Cells per binning tile (the unit of the per-tile totals of the cell counts).
constexpr int kBinTileShift = 12;
constexpr int kBinTile = 1 << kBinTileShift;

// Node tiles of the node passes (x fastest); a tile is processed only when a
// cell touching one of its nodes holds particles.
constexpr int kNodeTileX = 64;
constexpr int kNodeTileY = 2;
constexpr int kNodeTileZ = 2;

__host__ __device__ inline int cdiv(int a, int b) { return (a + b - 1) / b; }
__host__ __device__ inline int64_t cdiv64(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Cell key of a position with the reference's FP64 formula
// (material_point.cpp:59-65); returns nelem (sentinel) outside the grid.
__device__ inline uint32_t cell_key_f64(const GridGeom &g, double x, double y, double z) {
    double xx = x - g.xmin[0];
    double yy = y - g.xmin[1];
    double zz = z - g.xmin[2];
    double fx = floor(xx / g.dx[0]);
    double fy = floor(yy / g.dx[1]);
    double fz = floor(zz / g.dx[2]);
    // compare in double so that huge or NaN coordinates cannot wrap an int
    bool inside = (fx >= 0.0) && (fy >= 0.0) && (fz >= 0.0) && (fx < (double)g.nex[0]) &&
                  (fy < (double)g.nex[1]) && (fz < (double)g.nex[2]);
    if (!inside) return (uint32_t)g.nelem;
    int iex = (int)fx, iey = (int)fy, iez = (int)fz;
    return (uint32_t)(iex + iey * g.nex[0] + iez * g.nex[0] * g.nex[1]);
}

}  // namespace mpm

#define MPM_HIP_CHECK(expr)                                   \
    do {                                                      \
        hipError_t _e = (expr);                               \
        if (_e != hipSuccess) return ::mpm::hip_fail(_e, #expr, __FILE__, __LINE__); \
    } while (0)

namespace mpm {
int hip_fail(hipError_t e, const char *what, const char *file, int line);
}
