// Internal declarations shared by the HIP kernels and the host driver.
// Written for gfx950 (MI355X): 64-wide wavefronts, FP64 vector ALUs; every
// kernel here is an HBM-bandwidth-bound stencil or reduction.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace mps {

enum Face { LEFT = 0, RIGHT = 1, BOTTOM = 2, TOP = 3, BACK = 4, FRONT = 5 };
enum BcCode { BC_NONE = 0, BC_DIRC = 1, BC_ZGRD = 2, BC_FLUX = 3 };

// Operator forms.
//  OP_REF      reference face fluxes from cell coefficients, uniform
//              spacing, the reference's exact evaluation order
//              (mgridlib/discretepde.f90:620-1119).
//  OP_ALIGNED  same fluxes on the vertex-aligned coarse grids of the
//              m*2^k+1 hierarchy, whose boundary cells have width
//              (H + h0)/2 and sit h0/2 from the wall.
//  OP_STENCIL  explicit 7-point rows: c u + xm u(i-1) + xp u(i+1) + ...
//              (Scharfetter-Gummel operators and transposed operators).
//  OP_FACES    (a kernel variant, not a level's form) OP_STENCIL rows of
//              a symmetric operator read from their face form (Lev::fc):
//              four arrays instead of seven.
enum OpKind { OP_REF = 0, OP_ALIGNED = 1, OP_STENCIL = 2, OP_FACES = 3 };
// flux schemes as the kernels see them (values of MPS_FLUX_*)
enum { MPS_FLUX_UPWIND_K = 0, MPS_FLUX_SG_K = 1 };

// Stencil slots for OP_STENCIL.
enum { ST_C = 0, ST_XM, ST_XP, ST_YM, ST_YP, ST_ZM, ST_ZP, ST_N };

// One grid level as the kernels see it. Arrays share the ghosted layout
// (nx+2, ny+2, nz+2), x fastest: cell (i,j,k), 0 <= i <= nx+1, sits at
// i + sy*j + sz*k.
template <typename T>
struct Lev {
    int nx, ny, nz;
    int sy;
    long long sz;
    const T *D, *vx, *vy, *vz, *reac;   // cell coefficients (nullable)
    const T *st[ST_N];                  // explicit stencil rows (OP_STENCIL)
    // face form of the rows (nullable; conductance-coarsened levels): the
    // diagonal and the + face conductance of each cell per direction, G_d;
    // row c's x- entry is -(G_x[c - 1] * a) and its x+ entry -(G_x[c] * a),
    // a = axb on the first and last column, else ax (y, z alike); the
    // faces on the boundary hold 0
    const T *fc[4];
    T dx, dy, dz;                       // centre spacing
    T ax, ay, az;                       // 1/width of interior cells
    T axb, ayb, azb;                    // 1/width of boundary cells (OP_ALIGNED)
    T ix0, iy0, iz0;                    // 1/(distance centre -> wall) (OP_ALIGNED)
    int bc[6];
    int timederiv;
    T dt;
    // fast arithmetic (no convection): face weight = (D_L + D_R) * h*,
    // h* = abv/(2 d); DIRC boundary diagonal = D_c * g*; 1/dt precomputed.
    int fast;
    T hx, hy, hz, hxb, hyb, hzb;        // interior / boundary-cell face factors
    T gx, gy, gz, gxb, gyb, gzb;        // DIRC wall factors (interior unused)
    T dtinv;
    // over-relaxation of the fast sweeps (k_sweep_fast): u = u_old +
    // omega (u_GS - u_old); 1 = plain Gauss-Seidel
    T omega;
};

// Per-level, per-dimension restriction weights of the aligned hierarchy:
// coarse cell I draws on fine cells 2I-2, 2I-1, 2I with w[3*(I-1)+0..2].
struct AlignedW {
    const double *wx, *wy, *wz;
};

struct Reduce {
    double *partials;   // one per block
    int nblocks;
};

}  // namespace mps
