// Device layout of one tile's fields in HBM.
//
// Every 3-D field of a tile (cell, vertex and the three face classes) uses the
// same strides: x-fastest rows of `ld` elements, `ny+5` rows per plane, and
// `nz+5` planes, i.e. the vertex extents of the reference (build_field.f90:
// 41-147) for every class.  Sharing one stride set means a single offset per
// thread addresses every array, and neighbours are +-1, +-ld, +-plane.  Rows
// are padded so that the first interior element (j = x_min) of every row
// starts on a 32-element boundary (256 bytes of FP64, 128 of FP32): a
// wavefront's 64 interior elements then fill whole 128-byte cache lines
// instead of sharing partial lines with the neighbouring workgroups.  Indices are the reference's tile-local
// Fortran indices (x_min = 1, halo down to x_min-2).
#pragma once

#include <cstdint>

namespace clover {

constexpr int kRowAlign = 32;   // elements
constexpr int kFrontPad = 30;   // (x_min-2) sits 30 elements into the slab so x_min is aligned

struct Lay {
    int x_min, x_max, y_min, y_max, z_min, z_max;
    int ld;       // row pitch (elements)
    int plane;    // ld * (ny + 5)
    int base;     // offset of (x_min-2, y_min-2, z_min-2)
    // faces whose reflective halo the fused kernels do not read for cell and
    // vertex fields: they read the interior cell or node the halo would
    // mirror instead (bits 1, 2: the x_min and x_max faces; 4, 8: y; 16, 32:
    // z; 0 where halos are exchanged).  The advection sweeps use the x bits
    // only.
    int mirror;
};

__host__ __device__ inline int idx(const Lay &g, int j, int k, int l)
{
    return g.base + (j - g.x_min + 2) + g.ld * (k - g.y_min + 2) + g.plane * (l - g.z_min + 2);
}

inline Lay make_lay(int x_min, int x_max, int y_min, int y_max, int z_min, int z_max)
{
    Lay g;
    g.x_min = x_min; g.x_max = x_max; g.y_min = y_min; g.y_max = y_max; g.z_min = z_min; g.z_max = z_max;
    int nx = x_max - x_min + 1, ny = y_max - y_min + 1;
    int need = kFrontPad + nx + 5;
    g.ld = (need + kRowAlign - 1) / kRowAlign * kRowAlign;
    g.plane = g.ld * (ny + 5);
    g.base = kFrontPad;
    g.mirror = 0;
    return g;
}

// elements to allocate for one 3-D field
inline int64_t lay_elems(const Lay &g)
{
    int nz = g.z_max - g.z_min + 1;
    return (int64_t)g.base + (int64_t)g.plane * (nz + 5);
}

// Array classes of the reference, with their Fortran extents beyond n.
enum ArrClass : int {
    AC_CELL = 0,  // (nx+4, ny+4, nz+4)
    AC_VERT = 1,  // (nx+5, ny+5, nz+5)
    AC_XF = 2,    // (nx+5, ny+4, nz+4)
    AC_YF = 3,    // (nx+4, ny+5, nz+4)
    AC_ZF = 4,    // (nx+4, ny+4, nz+5)
    AC_C1X = 5, AC_C1Y = 6, AC_C1Z = 7,  // 1-D, n+4
    AC_V1X = 8, AC_V1Y = 9, AC_V1Z = 10, // 1-D, n+5
};

inline void class_extents(int cls, int nx, int ny, int nz, int &ex, int &ey, int &ez)
{
    ex = nx + 4; ey = ny + 4; ez = nz + 4;
    switch (cls) {
    case AC_VERT: ex++; ey++; ez++; break;
    case AC_XF: ex++; break;
    case AC_YF: ey++; break;
    case AC_ZF: ez++; break;
    case AC_C1X: ey = ez = 1; break;
    case AC_C1Y: ex = ny + 4; ey = ez = 1; break;
    case AC_C1Z: ex = nz + 4; ey = ez = 1; break;
    case AC_V1X: ex = nx + 5; ey = ez = 1; break;
    case AC_V1Y: ex = ny + 5; ey = ez = 1; break;
    case AC_V1Z: ex = nz + 5; ey = ez = 1; break;
    default: break;
    }
}

inline bool class_is_1d(int cls) { return cls >= AC_C1X; }

// Field ids of the reference (data.f90:55-74), zero based here.
enum FieldId : int {
    F_DENSITY0 = 0, F_DENSITY1, F_ENERGY0, F_ENERGY1, F_PRESSURE, F_VISCOSITY, F_SOUNDSPEED,
    F_XVEL0, F_XVEL1, F_YVEL0, F_YVEL1, F_ZVEL0, F_ZVEL1,
    F_VOL_FLUX_X, F_VOL_FLUX_Y, F_VOL_FLUX_Z, F_MASS_FLUX_X, F_MASS_FLUX_Y, F_MASS_FLUX_Z,
    NUM_FIELDS
};

}  // namespace clover
