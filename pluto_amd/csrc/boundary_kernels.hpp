// Ghost-zone filling of PLUTO's ghosted boxes on the device
// (boundary_kernels.hip): the physical boundary conditions of boundary.c
// (outflow, reflective, axisymmetric, equatorial-symmetric, periodic) on the
// cell-centred primitives and the staggered field, the normal staggered
// field of the ghost faces from div.B = 0 (MHD/CT/ct_fill_mag_field.c) and
// the cell-centred normal field of outflow ghosts (ct_field_average.c).
#pragma once
#include <hip/hip_runtime.h>

namespace pgpu {

// boundary types (pluto.h:146-152)
constexpr int BOUND_OUTFLOW = 1;
constexpr int BOUND_REFLECTIVE = 2;
constexpr int BOUND_AXISYMMETRIC = 3;
constexpr int BOUND_EQTSYMMETRIC = 4;
constexpr int BOUND_PERIODIC = 5;
constexpr int BOUND_USERDEF = 7;   // user code on the host (boundary.c:258-292)

// A ghosted box: t[a] cells along axis a (NX1_TOT, NX2_TOT, NX3_TOT), the
// domain beg[a]..end[a] (IBEG..IEND), n[a] = end - beg + 1 interior cells.
// Device arrays are [k][j][i] blocks.  With uni == 0 (PLUTO's own arrays) a
// staggered array (and a face-area array) also holds the face plane -1 of
// its own axis, so its extent along that axis is t[a] + 1 and its element
// (k, j, i) sits at index + 1 there.  With uni == 1 (a resident domain)
// every array is (t[2]+1) x (t[1]+1) x (t[0]+1) and element (k, j, i) sits
// at (k+1, j+1, i+1): cell -1 is a spare plane whose upper face is face -1.
struct GhostBox {
  int t[3], beg[3], end[3], n[3];
  int uni = 0;
};

// Fill the ghosts of one side (0..5 = X1_BEG, X1_END, X2_BEG, ... X3_END)
// of boundary type `type`: vc holds the 8 primitive arrays, vs the three
// staggered arrays, area the three face-area arrays (MHD/CT/ct_fill_mag_field.c
// reads them; unused for PERIODIC), or null with area_c[a] the area of every
// face normal to axis a (a uniform Cartesian grid).  Launches on s; returns
// a hipError_t.
int launch_boundary_side(const GhostBox& b, int side, int type, double* const* vc,
                         double* const* vs, const double* const* area, const double* area_c,
                         hipStream_t s);

// The device part of a USERDEF side once the user's values are in place
// (boundary.c:285-290): FillMagneticField's div.B = 0 ghost faces and
// CT_AverageNormalMagField's cell-centred normal field.  (The transverse
// average, CT_AverageTransverseMagField, is done by the host caller.)
int launch_boundary_fill_normal(const GhostBox& b, int side, double* const* vc, double* const* vs,
                                const double* const* area, const double* area_c, hipStream_t s);

}  // namespace pgpu
