// HIP kernels for the multigrid field solve (gfx950).
//
// All kernels are bandwidth-bound 7-point stencils, grid transfers or
// reductions: no MFMA. One thread owns one cell (or one coarse cell for the
// transfers); threads are laid out over flattened (i,j) planes with
// blockIdx.y walking k, so consecutive lanes of a 64-wide wavefront touch
// consecutive x addresses.
//
// Reference-arithmetic rule: for T = double and OP_REF the diagonal and
// off-diagonal sums are accumulated in the order of the reference's find_AX
// (mgridlib/discretepde.f90:754-1106) and findcellcontributions
// (:620-670), and the library is built without FMA contraction, so single
// cells match the reference bit for bit.
#pragma once

#include "mps3d_internal.h"

namespace mps {

constexpr int kBlock = 256;
constexpr int kBottomMaxDense = 512;  // largest coarsest grid solved by a dense inverse

// ------------------------------------------------------------------------
// XCD-aware tile order
// ------------------------------------------------------------------------
// MI355X dispatches the workgroups of a launch round-robin over its 8 XCDs,
// each with its own L2: consecutive block ids (neighbouring tiles) land on
// different L2s, and every tile's halo is fetched from HBM once per XCD.
// Tile kernels renumber their blocks so that each XCD works through one
// contiguous range of tiles (x fastest, then y, then z); neighbouring tiles
// then share an L2. The map is a bijection for any grid size.
constexpr unsigned kXcds = 8;

__device__ __forceinline__ void xcd_tile(int &bx, int &by, int &bz)
{
    const unsigned gx = gridDim.x, gy = gridDim.y;
    const unsigned n = gx * gy * gridDim.z;
    const unsigned lin = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const unsigned x = lin % kXcds, m = lin / kXcds;
    const unsigned per = n / kXcds, rem = n % kXcds;
    const unsigned t = x * per + (x < rem ? x : rem) + m;
    bx = (int)(t % gx);
    const unsigned r = t / gx;
    by = (int)(r % gy);
    bz = (int)(r / gy);
}

// ------------------------------------------------------------------------
// raw buffer access
// ------------------------------------------------------------------------
// A buffer resource (base, size in bytes) lets a load or store take a
// per-lane 32-bit byte offset in a vector register and a uniform one in a
// scalar register, with no 64-bit address arithmetic per access. The
// hardware range-checks raw buffers: an access whose offset lies at or
// beyond the resource's size returns 0 (loads) or is dropped (stores), so a
// per-lane offset of kOob marks an invalid element without a select.
// Resources here are built over at most 2^31 - 1 bytes, so every valid
// offset is below kOob.
constexpr int kBufferWord3 = 0x00020000;  // raw buffer descriptor word 3 (gfx9 family)
constexpr int kOob = (int)0x80000000u;     // per-lane offset of an invalid element

template <typename T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const T *p, int nbytes)
{
    return __builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, nbytes, kBufferWord3);
}
template <typename T>
__device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    if constexpr (sizeof(T) == 4)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
template <typename T>
__device__ __forceinline__ void buf_store(T v, __amdgpu_buffer_rsrc_t r, int voff, int soff);
template <>
__device__ __forceinline__ void buf_store<float>(float v, __amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, 0);
}
template <>
__device__ __forceinline__ void buf_store<double>(double v, __amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, 0);
}

// A 32-bit buffer load into LDS: lane l of the wave writes its word at
// lds + l (lds is uniform over the wave), out-of-range offsets write 0. The
// write lands asynchronously and is counted by vmcnt like a register load.
__device__ __forceinline__ void buf_load_lds(__amdgpu_buffer_rsrc_t r, float *lds, int voff, int soff)
{
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (__attribute__((address_space(3))) void *)lds, 4, voff, soff, 0, 0);
}
// s_waitcnt vmcnt(N) (gfx9 encoding; expcnt and lgkmcnt left at their
// maxima)
template <int N>
__device__ __forceinline__ void wait_vmcnt()
{
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | 0x0F70);
}

// per-face multipliers of a staged ghost coefficient (fast sweeps): the
// ghost cell's D is staged as m_f times the adjacent interior cell's, so a
// face weight (D_c + D_ghost) h_f carries the boundary condition
template <typename T>
struct GhostD {
    T m[6];
};

// ------------------------------------------------------------------------
// cell addressing
// ------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ bool cell_of(const Lev<T> &L, int &i, int &j, int &k, long long &c)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= L.nx * L.ny) return false;
    j = t / L.nx;
    i = t - j * L.nx + 1;
    j += 1;
    k = blockIdx.y + 1;
    c = (long long)i + (long long)L.sy * j + L.sz * k;
    return true;
}

// (i,j) of this thread for kernels that loop over k themselves
template <typename T>
__device__ __forceinline__ bool cell_ij(const Lev<T> &L, int &i, int &j)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= L.nx * L.ny) return false;
    j = t / L.nx;
    i = t - j * L.nx + 1;
    j += 1;
    return true;
}

// ------------------------------------------------------------------------
// reference face contributions (findcellcontributions)
// ------------------------------------------------------------------------
// contribution of the face (m|c) to cell c (c is the right-hand cell)
template <typename T, bool VEL>
__device__ __forceinline__ void face_right(T ul, T kl, T kr, T vl, T vr, T abv, T d, T &dg, T &of)
{
    const T zero = T(0);
    T kh = T(0.5) * (kl + kr);
    T dr = zero + (kh / d) * abv;
    T orr = zero - ul * (kh / d) * abv;
    if (VEL) {
        T vh = T(0.5) * (vl + vr);
        if (vh > zero) {
            orr = orr - vh * ul * abv;
        } else {
            dr = dr - vh * abv;
        }
    }
    dg = dg + dr;
    of = of + orr;
}

// contribution of the face (c|p) to cell c (c is the left-hand cell)
template <typename T, bool VEL>
__device__ __forceinline__ void face_left(T ur, T kl, T kr, T vl, T vr, T abv, T d, T &dg, T &of)
{
    const T zero = T(0);
    T kh = T(0.5) * (kl + kr);
    T dl = zero + (kh / d) * abv;
    T ol = zero - ur * (kh / d) * abv;
    if (VEL) {
        T vh = T(0.5) * (vl + vr);
        if (vh > zero) {
            dl = dl + vh * abv;
        } else {
            ol = ol + vh * ur * abv;
        }
    }
    dg = dg + dl;
    of = of + ol;
}

// boundary row terms of find_AX (discretepde.f90:798-865 and the y/z
// copies): 'lo' is the x-/y-/z- face, 'hi' the + face.
template <typename T, int OP, bool VEL>
__device__ __forceinline__ void bc_row(int code, bool hi, T Dc, T vb, T h, T abv, T idist, T &dg)
{
    if (code == BC_DIRC) {
        if (OP == OP_REF) {
            dg = dg + T(2) * Dc / (h * h);
            if (VEL) {
                if (hi ? (vb > T(0)) : (vb < T(0))) dg = hi ? dg + vb / h : dg - vb / h;
            }
        } else {
            dg = dg + Dc * idist * abv;
            if (VEL) {
                if (hi ? (vb > T(0)) : (vb < T(0))) dg = hi ? dg + vb * abv : dg - vb * abv;
            }
        }
    } else if (code == BC_ZGRD) {
        if (VEL) {
            if (OP == OP_REF) dg = hi ? dg + vb / h : dg - vb / h;
            else dg = hi ? dg + vb * abv : dg - vb * abv;
        }
    }
}

// diag and offdiag of a cell without convection, from the six neighbour
// values and the seven diffusion coefficients (find_AX order; the neighbour
// values of faces that do not exist are ignored).
template <typename T, int OP, bool REAC>
__device__ __forceinline__ void row7(const Lev<T> &L, int i, int j, int k, T uxm, T uxp, T uym, T uyp, T uzm,
                                     T uzp, T Dc, T Dxm, T Dxp, T Dym, T Dyp, T Dzm, T Dzp, T rc, T &dg, T &of)
{
    if (L.fast) {
#pragma clang fp contract(fast)
        // same operator, precomputed factors: w = (D_L + D_R) * abv / (2 d)
        T d = T(0), o = T(0);
        {
            const bool edge = (OP == OP_ALIGNED) && ((i == 1) || (i == L.nx));
            const T h = edge ? L.hxb : L.hx;
            if (i > 1) { const T w = (Dxm + Dc) * h; d += w; o -= w * uxm; }
            if (i < L.nx) { const T w = (Dc + Dxp) * h; d += w; o -= w * uxp; }
            if (i == 1 && L.bc[LEFT] == BC_DIRC) d += Dc * (edge ? L.gxb : L.gx);
            if (i == L.nx && L.bc[RIGHT] == BC_DIRC) d += Dc * (edge ? L.gxb : L.gx);
        }
        {
            const bool edge = (OP == OP_ALIGNED) && ((j == 1) || (j == L.ny));
            const T h = edge ? L.hyb : L.hy;
            if (j > 1) { const T w = (Dym + Dc) * h; d += w; o -= w * uym; }
            if (j < L.ny) { const T w = (Dc + Dyp) * h; d += w; o -= w * uyp; }
            if (j == 1 && L.bc[BOTTOM] == BC_DIRC) d += Dc * (edge ? L.gyb : L.gy);
            if (j == L.ny && L.bc[TOP] == BC_DIRC) d += Dc * (edge ? L.gyb : L.gy);
        }
        {
            const bool edge = (OP == OP_ALIGNED) && ((k == 1) || (k == L.nz));
            const T h = edge ? L.hzb : L.hz;
            if (k > 1) { const T w = (Dzm + Dc) * h; d += w; o -= w * uzm; }
            if (k < L.nz) { const T w = (Dc + Dzp) * h; d += w; o -= w * uzp; }
            if (k == 1 && L.bc[BACK] == BC_DIRC) d += Dc * (edge ? L.gzb : L.gz);
            if (k == L.nz && L.bc[FRONT] == BC_DIRC) d += Dc * (edge ? L.gzb : L.gz);
        }
        if (REAC) d -= rc;
        d += L.dtinv;
        dg = d;
        of = o;
        return;
    }
    T d = T(0), o = T(0);
    {
        const bool edge = (i == 1) || (i == L.nx);
        const T abv = (OP == OP_ALIGNED && edge) ? L.axb : L.ax;
        if (i > 1) face_right<T, false>(uxm, Dxm, Dc, T(0), T(0), abv, L.dx, d, o);
        if (i < L.nx) face_left<T, false>(uxp, Dc, Dxp, T(0), T(0), abv, L.dx, d, o);
        if (i == 1) bc_row<T, OP, false>(L.bc[LEFT], false, Dc, T(0), L.dx, abv, L.ix0, d);
        if (i == L.nx) bc_row<T, OP, false>(L.bc[RIGHT], true, Dc, T(0), L.dx, abv, L.ix0, d);
    }
    {
        const bool edge = (j == 1) || (j == L.ny);
        const T abv = (OP == OP_ALIGNED && edge) ? L.ayb : L.ay;
        if (j > 1) face_right<T, false>(uym, Dym, Dc, T(0), T(0), abv, L.dy, d, o);
        if (j < L.ny) face_left<T, false>(uyp, Dc, Dyp, T(0), T(0), abv, L.dy, d, o);
        if (j == 1) bc_row<T, OP, false>(L.bc[BOTTOM], false, Dc, T(0), L.dy, abv, L.iy0, d);
        if (j == L.ny) bc_row<T, OP, false>(L.bc[TOP], true, Dc, T(0), L.dy, abv, L.iy0, d);
    }
    {
        const bool edge = (k == 1) || (k == L.nz);
        const T abv = (OP == OP_ALIGNED && edge) ? L.azb : L.az;
        if (k > 1) face_right<T, false>(uzm, Dzm, Dc, T(0), T(0), abv, L.dz, d, o);
        if (k < L.nz) face_left<T, false>(uzp, Dc, Dzp, T(0), T(0), abv, L.dz, d, o);
        if (k == 1) bc_row<T, OP, false>(L.bc[BACK], false, Dc, T(0), L.dz, abv, L.iz0, d);
        if (k == L.nz) bc_row<T, OP, false>(L.bc[FRONT], true, Dc, T(0), L.dz, abv, L.iz0, d);
    }
    if (REAC) d = d - rc;
    if (L.timederiv) d = d + T(1) / L.dt;
    dg = d;
    of = o;
}

// Gauss-Seidel update (b - off)/diag; the fast arithmetic mode uses the
// hardware reciprocal in FP32.
template <typename T>
__device__ __forceinline__ T gs_update(const Lev<T> &L, T b, T of, T dg)
{
    if constexpr (sizeof(T) == 4) {
        if (L.fast) return (b - of) * __builtin_amdgcn_rcpf(dg);
    }
    return (b - of) / dg;
}

// diag and offdiag of cell (i,j,k) at address c for the current u.
template <typename T, int OP, bool VEL, bool REAC>
__device__ __forceinline__ void row(const Lev<T> &L, const T *__restrict__ u, int i, int j, int k,
                                    long long c, T &dg, T &of)
{
    if (OP != OP_STENCIL && !VEL) {
        const long long sy = L.sy, sz = L.sz;
        const T *__restrict__ D = L.D;
        row7<T, OP, REAC>(L, i, j, k, u[c - 1], u[c + 1], u[c - sy], u[c + sy], u[c - sz], u[c + sz], D[c],
                          D[c - 1], D[c + 1], D[c - sy], D[c + sy], D[c - sz], D[c + sz],
                          REAC ? L.reac[c] : T(0), dg, of);
        return;
    }
    if (OP == OP_STENCIL) {
        const long long sy = L.sy, sz = L.sz;
        T o = T(0);
        o = o + L.st[ST_XM][c] * u[c - 1];
        o = o + L.st[ST_XP][c] * u[c + 1];
        o = o + L.st[ST_YM][c] * u[c - sy];
        o = o + L.st[ST_YP][c] * u[c + sy];
        o = o + L.st[ST_ZM][c] * u[c - sz];
        o = o + L.st[ST_ZP][c] * u[c + sz];
        dg = L.st[ST_C][c];
        of = o;
        return;
    }
    const T *__restrict__ D = L.D;
    T d = T(0), o = T(0);
    const T Dc = D[c];
    // ---- x ----
    {
        const bool edge = (i == 1) || (i == L.nx);
        const T abv = (OP == OP_ALIGNED && edge) ? L.axb : L.ax;
        const T vc = VEL ? L.vx[c] : T(0);
        if (i > 1) face_right<T, VEL>(u[c - 1], D[c - 1], Dc, VEL ? L.vx[c - 1] : T(0), vc, abv, L.dx, d, o);
        if (i < L.nx) face_left<T, VEL>(u[c + 1], Dc, D[c + 1], vc, VEL ? L.vx[c + 1] : T(0), abv, L.dx, d, o);
        if (i == 1) bc_row<T, OP, VEL>(L.bc[LEFT], false, Dc, vc, L.dx, abv, L.ix0, d);
        if (i == L.nx) bc_row<T, OP, VEL>(L.bc[RIGHT], true, Dc, vc, L.dx, abv, L.ix0, d);
    }
    // ---- y ----
    {
        const long long s = L.sy;
        const bool edge = (j == 1) || (j == L.ny);
        const T abv = (OP == OP_ALIGNED && edge) ? L.ayb : L.ay;
        const T vc = VEL ? L.vy[c] : T(0);
        if (j > 1) face_right<T, VEL>(u[c - s], D[c - s], Dc, VEL ? L.vy[c - s] : T(0), vc, abv, L.dy, d, o);
        if (j < L.ny) face_left<T, VEL>(u[c + s], Dc, D[c + s], vc, VEL ? L.vy[c + s] : T(0), abv, L.dy, d, o);
        if (j == 1) bc_row<T, OP, VEL>(L.bc[BOTTOM], false, Dc, vc, L.dy, abv, L.iy0, d);
        if (j == L.ny) bc_row<T, OP, VEL>(L.bc[TOP], true, Dc, vc, L.dy, abv, L.iy0, d);
    }
    // ---- z ----
    {
        const long long s = L.sz;
        const bool edge = (k == 1) || (k == L.nz);
        const T abv = (OP == OP_ALIGNED && edge) ? L.azb : L.az;
        const T vc = VEL ? L.vz[c] : T(0);
        if (k > 1) face_right<T, VEL>(u[c - s], D[c - s], Dc, VEL ? L.vz[c - s] : T(0), vc, abv, L.dz, d, o);
        if (k < L.nz) face_left<T, VEL>(u[c + s], Dc, D[c + s], vc, VEL ? L.vz[c + s] : T(0), abv, L.dz, d, o);
        if (k == 1) bc_row<T, OP, VEL>(L.bc[BACK], false, Dc, vc, L.dz, abv, L.iz0, d);
        if (k == L.nz) bc_row<T, OP, VEL>(L.bc[FRONT], true, Dc, vc, L.dz, abv, L.iz0, d);
    }
    if (REAC) d = d - L.reac[c];
    if (L.timederiv) d = d + T(1) / L.dt;
    dg = d;
    of = o;
}

// The seven row coefficients (centre and the six neighbours) of an
// OP_REF / OP_ALIGNED row, for building explicit (e.g. transposed) stencils.
template <typename T, int OP, bool VEL, bool REAC>
__device__ void row_coeffs(const Lev<T> &L, int i, int j, int k, long long c, T a[ST_N])
{
    // the off-diagonal sum is linear in the neighbour values: probe it with
    // unit values, one neighbour at a time, through the same face formulas
    T dg, of;
    const T *D = L.D;
    const T Dc = D[c];
    for (int q = 0; q < ST_N; ++q) a[q] = T(0);
    T d = T(0), o;
    {
        const bool edge = (i == 1) || (i == L.nx);
        const T abv = (OP == OP_ALIGNED && edge) ? L.axb : L.ax;
        const T vc = VEL ? L.vx[c] : T(0);
        if (i > 1) { o = T(0); face_right<T, VEL>(T(1), D[c - 1], Dc, VEL ? L.vx[c - 1] : T(0), vc, abv, L.dx, d, o); a[ST_XM] = o; }
        if (i < L.nx) { o = T(0); face_left<T, VEL>(T(1), Dc, D[c + 1], vc, VEL ? L.vx[c + 1] : T(0), abv, L.dx, d, o); a[ST_XP] = o; }
        if (i == 1) bc_row<T, OP, VEL>(L.bc[LEFT], false, Dc, vc, L.dx, abv, L.ix0, d);
        if (i == L.nx) bc_row<T, OP, VEL>(L.bc[RIGHT], true, Dc, vc, L.dx, abv, L.ix0, d);
    }
    {
        const long long s = L.sy;
        const bool edge = (j == 1) || (j == L.ny);
        const T abv = (OP == OP_ALIGNED && edge) ? L.ayb : L.ay;
        const T vc = VEL ? L.vy[c] : T(0);
        if (j > 1) { o = T(0); face_right<T, VEL>(T(1), D[c - s], Dc, VEL ? L.vy[c - s] : T(0), vc, abv, L.dy, d, o); a[ST_YM] = o; }
        if (j < L.ny) { o = T(0); face_left<T, VEL>(T(1), Dc, D[c + s], vc, VEL ? L.vy[c + s] : T(0), abv, L.dy, d, o); a[ST_YP] = o; }
        if (j == 1) bc_row<T, OP, VEL>(L.bc[BOTTOM], false, Dc, vc, L.dy, abv, L.iy0, d);
        if (j == L.ny) bc_row<T, OP, VEL>(L.bc[TOP], true, Dc, vc, L.dy, abv, L.iy0, d);
    }
    {
        const long long s = L.sz;
        const bool edge = (k == 1) || (k == L.nz);
        const T abv = (OP == OP_ALIGNED && edge) ? L.azb : L.az;
        const T vc = VEL ? L.vz[c] : T(0);
        if (k > 1) { o = T(0); face_right<T, VEL>(T(1), D[c - s], Dc, VEL ? L.vz[c - s] : T(0), vc, abv, L.dz, d, o); a[ST_ZM] = o; }
        if (k < L.nz) { o = T(0); face_left<T, VEL>(T(1), Dc, D[c + s], vc, VEL ? L.vz[c + s] : T(0), abv, L.dz, d, o); a[ST_ZP] = o; }
        if (k == 1) bc_row<T, OP, VEL>(L.bc[BACK], false, Dc, vc, L.dz, abv, L.iz0, d);
        if (k == L.nz) bc_row<T, OP, VEL>(L.bc[FRONT], true, Dc, vc, L.dz, abv, L.iz0, d);
    }
    if (REAC) d = d - L.reac[c];
    if (L.timederiv) d = d + T(1) / L.dt;
    a[ST_C] = d;
    (void)dg;
    (void)of;
}

// ------------------------------------------------------------------------
// block reduction (deterministic: fixed tree inside a block)
// ------------------------------------------------------------------------
__device__ __forceinline__ double block_sum(double v)
{
    __shared__ double red[kBlock / 64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
    }
    return s;  // valid on thread 0
}

// ------------------------------------------------------------------------
// kernels: operator, smoother, residual
// ------------------------------------------------------------------------

// find_AX: AX = diag*u + off (interior arrays, each nullable)
template <typename T, int OP, bool VEL, bool REAC>
__global__ void __launch_bounds__(kBlock) k_apply(Lev<T> L, const T *__restrict__ u, T *ax, T *diag, T *off)
{
    int i, j, k;
    long long c;
    if (!cell_of(L, i, j, k, c)) return;
    T dg, of;
    row<T, OP, VEL, REAC>(L, u, i, j, k, c, dg, of);
    long long n = (long long)(i - 1) + (long long)L.nx * ((j - 1) + (long long)L.ny * (k - 1));
    if (ax) ax[n] = dg * u[c] + of;
    if (diag) diag[n] = dg;
    if (off) off[n] = of;
}

// one red (red=1) or black (red=0) pass: u = (b - off)/diag on one colour;
// red = odd (i+j+k), the reference's colouring (discretepde.f90:49-57)
// compact != 0: one thread per cell of the colour: row (j, k) holds the
// cells i = i0, i0 + 2, ... <= nx, i0 the first of the colour, and
// (nx + 1) / 2 threads per row cover them (colour_grid). compact == 0: one
// thread per cell (cells_grid), for small levels, where the compact grid
// would leave too few waves to hide the memory latency.
template <typename T, int OP, bool VEL, bool REAC>
__global__ void __launch_bounds__(kBlock) k_colour(Lev<T> L, T *__restrict__ u, const T *__restrict__ b, int red,
                                                  int compact)
{
    int i, j, k;
    long long c;
    if (compact) {
        const int hx = (L.nx + 1) / 2;
        const int t = blockIdx.x * blockDim.x + threadIdx.x;
        if (t >= hx * L.ny) return;
        j = t / hx + 1;
        k = blockIdx.y + 1;
        const int i0 = ((1 + j + k) & 1) == red ? 1 : 2;
        i = i0 + 2 * (t - (j - 1) * hx);
        if (i > L.nx) return;
        c = (long long)i + (long long)L.sy * j + L.sz * k;
    } else {
        if (!cell_of(L, i, j, k, c)) return;
        if (((i + j + k) & 1) != red) return;
    }
    T dg, of;
    row<T, OP, VEL, REAC>(L, u, i, j, k, c, dg, of);
    u[c] = gs_update(L, b[c], of, dg);
}

// nu red-black sweeps of a small level in one workgroup: k_colour's passes
// (red, then black, nu times) with block barriers between them instead of
// launches, each cell updated by the same arithmetic. Each lane takes the
// pass's colour cells (k_colour's compact numbering) two at a time and forms
// both values before it stores either, so their loads are in flight
// together. For levels of at most kSmallSmooth cells, where a pass is a few
// microseconds of launch.
constexpr int kSmallSmooth = 8192;
constexpr int kSmallBlock = 1024;
template <typename T, int OP, bool VEL, bool REAC>
__global__ void __launch_bounds__(kSmallBlock) k_colour_small(Lev<T> L, T *__restrict__ u, const T *__restrict__ b,
                                                             int nu)
{
    const int hx = (L.nx + 1) / 2, ncol = hx * L.ny * L.nz;
    // the new value of colour cell t (c = -1: none)
    auto cell = [&](int t, int red, long long &c, T &v) {
        c = -1;
        if (t >= ncol) return;
        const int rw = t / hx;   // (j - 1) + ny (k - 1)
        const int k = rw / L.ny + 1, j = rw - (k - 1) * L.ny + 1;
        const int i = (((1 + j + k) & 1) == red ? 1 : 2) + 2 * (t - rw * hx);
        if (i > L.nx) return;
        c = (long long)i + (long long)L.sy * j + L.sz * k;
        T dg, of;
        row<T, OP, VEL, REAC>(L, u, i, j, k, c, dg, of);
        v = gs_update(L, b[c], of, dg);
    };
    for (int s = 0; s < nu; ++s) {
        for (int red = 1; red >= 0; --red) {
            for (int t = threadIdx.x; t < ncol; t += 2 * kSmallBlock) {
                long long c0, c1;
                T v0 = T(0), v1 = T(0);
                cell(t, red, c0, v0);
                cell(t + kSmallBlock, red, c1, v1);
                if (c0 >= 0) u[c0] = v0;
                if (c1 >= 0) u[c1] = v1;
            }
            __syncthreads();
        }
    }
}

// r = b - (diag*u + off); writes r (type R, ghosted layout, nullable) and the
// block's partial sum of r^2 (nullable). When R is narrower than T, the
// stored r is scaled by rscale[0] (a power of two, nullable = 1) so that
// residuals of any magnitude land in R's range; the sum is of r itself.
template <typename T, typename R, int OP, bool VEL, bool REAC>
__global__ void __launch_bounds__(kBlock) k_residual(Lev<T> L, const T *__restrict__ u, const T *__restrict__ b,
                                                    R *r, double *partials, const double *rscale = nullptr)
{
    int i, j;
    double sq = 0.0;
    T rs = T(1);
    if constexpr (sizeof(R) < sizeof(T)) {
        if (rscale) rs = (T)rscale[0];
    }
    if (cell_ij(L, i, j)) {
        for (int k = blockIdx.y + 1; k <= L.nz; k += gridDim.y) {
            const long long c = (long long)i + (long long)L.sy * j + L.sz * k;
            T dg, of;
            row<T, OP, VEL, REAC>(L, u, i, j, k, c, dg, of);
            T rv = b[c] - (dg * u[c] + of);
            if constexpr (sizeof(R) < sizeof(T)) {
                if (r) r[c] = (R)(rv * rs);
            } else {
                if (r) r[c] = (R)rv;
            }
            sq += (double)rv * (double)rv;
        }
    }
    if (partials) {
        double s = block_sum(sq);
        if (threadIdx.x == 0) partials[blockIdx.x + (long long)gridDim.x * blockIdx.y] = s;
    }
}

// sums n partials in a fixed order with one block; out[0] = sum
__global__ void __launch_bounds__(kBlock) k_reduce_final(const double *partials, int n, double *out)
{
    double s = 0.0;
    for (int q = threadIdx.x; q < n; q += blockDim.x) s += partials[q];
    s = block_sum(s);
    if (threadIdx.x == 0) out[0] = s;
}

// ------------------------------------------------------------------------
// right-hand side (find_bvec) on the finest level, FP64
// ------------------------------------------------------------------------
struct Faces {
    const double *f[6];
};

__global__ void __launch_bounds__(kBlock) k_bvec(Lev<double> L, Faces F, const double *sterm /* interior, nullable */,
                                                const double *source /* ghosted, nullable */, int applybc, double *b)
{
    int i, j, k;
    long long c;
    if (!cell_of(L, i, j, k, c)) return;
    double v = 0.0;
    const double Dc = L.D[c];
    if (applybc) {
        const double vx = L.vx ? L.vx[c] : 0.0, vy = L.vy ? L.vy[c] : 0.0, vz = L.vz ? L.vz[c] : 0.0;
        const double dx = L.dx, dy = L.dy, dz = L.dz;
        if (i == 1) {
            double g = F.f[LEFT][(j - 1) + (long long)L.ny * (k - 1)];
            if (L.bc[LEFT] == BC_DIRC) {
                v = v + 2.0 * Dc * g / (dx * dx);
                if (vx > 0.0) v = v + vx * g / dx;
            } else if (L.bc[LEFT] == BC_FLUX) {
                v = v - g / dx;
            }
        }
        if (i == L.nx) {
            double g = F.f[RIGHT][(j - 1) + (long long)L.ny * (k - 1)];
            if (L.bc[RIGHT] == BC_DIRC) {
                v = v + 2.0 * Dc * g / (dx * dx);
                if (vx < 0.0) v = v - vx * g / dx;
            } else if (L.bc[RIGHT] == BC_FLUX) {
                v = v - g / dx;
            }
        }
        if (j == 1) {
            double g = F.f[BOTTOM][(i - 1) + (long long)L.nx * (k - 1)];
            if (L.bc[BOTTOM] == BC_DIRC) {
                v = v + 2.0 * Dc * g / (dy * dy);
                if (vy > 0.0) v = v + vy * g / dy;
            } else if (L.bc[BOTTOM] == BC_FLUX) {
                v = v - g / dy;
            }
        }
        if (j == L.ny) {
            double g = F.f[TOP][(i - 1) + (long long)L.nx * (k - 1)];
            if (L.bc[TOP] == BC_DIRC) {
                v = v + 2.0 * Dc * g / (dy * dy);
                if (vy < 0.0) v = v - vy * g / dy;
            } else if (L.bc[TOP] == BC_FLUX) {
                v = v - g / dy;
            }
        }
        if (k == 1) {
            double g = F.f[BACK][(i - 1) + (long long)L.nx * (j - 1)];
            if (L.bc[BACK] == BC_DIRC) {
                v = v + 2.0 * Dc * g / (dz * dz);
                if (vz > 0.0) v = v + vz * g / dz;
            } else if (L.bc[BACK] == BC_FLUX) {
                v = v - g / dz;
            }
        }
        if (k == L.nz) {
            double g = F.f[FRONT][(i - 1) + (long long)L.nx * (j - 1)];
            if (L.bc[FRONT] == BC_DIRC) {
                v = v + 2.0 * Dc * g / (dz * dz);
                if (vz < 0.0) v = v - vz * g / dz;
            } else if (L.bc[FRONT] == BC_FLUX) {
                v = v - g / dz;
            }
        }
    }
    long long n = (long long)(i - 1) + (long long)L.nx * ((j - 1) + (long long)L.ny * (k - 1));
    v = v + (sterm ? sterm[n] : 0.0);
    if (applybc && source) v = v + source[c];
    b[c] = v;
}

// ------------------------------------------------------------------------
// reference (FAS) transfers
// ------------------------------------------------------------------------

// restriction (mgridsteps.f90:11-53) of u and r: coarse cell sums its eight
// children in the reference order. Writes uold and u_c (= R u) and b_c
// (= R r; tau is added by k_tau).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_restrict_fas(Lev<T> F, Lev<T> C, const T *__restrict__ uf,
                                                        const T *__restrict__ rf, T *uold, T *uc, T *bc)
{
    int i, j, k;
    long long c;
    if (!cell_of(C, i, j, k, c)) return;
    const long long sy = F.sy, sz = F.sz;
    const long long f0 = (long long)(2 * i - 1) + sy * (2 * j - 1) + sz * (2 * k - 1);
    const long long o[8] = {0, 1, sy, sy + 1, sz, sz + 1, sz + sy, sz + sy + 1};
    T s = T(0), q = T(0);
    s = s + uf[f0 + o[0]] + uf[f0 + o[1]];
    s = s + uf[f0 + o[2]] + uf[f0 + o[3]];
    s = s + uf[f0 + o[4]] + uf[f0 + o[5]];
    s = s + uf[f0 + o[6]] + uf[f0 + o[7]];
    q = q + rf[f0 + o[0]] + rf[f0 + o[1]];
    q = q + rf[f0 + o[2]] + rf[f0 + o[3]];
    q = q + rf[f0 + o[4]] + rf[f0 + o[5]];
    q = q + rf[f0 + o[6]] + rf[f0 + o[7]];
    const T eighth = T(1.0 / 8.0);
    s = eighth * s;
    uold[c] = s;
    uc[c] = s;
    bc[c] = eighth * q;
}

// 8-child mean of a coefficient array (same rule; once per solve)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_restrict_mean(Lev<T> F, Lev<T> C, const T *__restrict__ f, T *cf)
{
    int i, j, k;
    long long c;
    if (!cell_of(C, i, j, k, c)) return;
    const long long sy = F.sy, sz = F.sz;
    const long long f0 = (long long)(2 * i - 1) + sy * (2 * j - 1) + sz * (2 * k - 1);
    T s = T(0);
    s = s + f[f0] + f[f0 + 1];
    s = s + f[f0 + sy] + f[f0 + sy + 1];
    s = s + f[f0 + sz] + f[f0 + sz + 1];
    s = s + f[f0 + sz + sy] + f[f0 + sz + sy + 1];
    cf[c] = T(1.0 / 8.0) * s;
}

// the coefficient fields of a level (null = absent) and their coarse copies
struct FieldSet5 {
    const double *src[5];
    double *dst[5];
};

// k_restrict_mean for up to five fields in one launch
__global__ void __launch_bounds__(kBlock) k_restrict_mean5(Lev<double> F, Lev<double> C, FieldSet5 S)
{
    int i, j, k;
    long long c;
    if (!cell_of(C, i, j, k, c)) return;
    const long long sy = F.sy, sz = F.sz;
    const long long f0 = (long long)(2 * i - 1) + sy * (2 * j - 1) + sz * (2 * k - 1);
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        const double *f = S.src[q];
        if (!f) continue;
        double s = 0.0;
        s = s + f[f0] + f[f0 + 1];
        s = s + f[f0 + sy] + f[f0 + sy + 1];
        s = s + f[f0 + sz] + f[f0 + sz + 1];
        s = s + f[f0 + sz + sy] + f[f0 + sz + sy + 1];
        S.dst[q][c] = (1.0 / 8.0) * s;
    }
}

// FAS right-hand side on the coarse level: b_c = A_c(R u) + R r
// (mgridsteps.f90:544-550); b_c holds R r on entry.
template <typename T, int OP, bool VEL, bool REAC>
__global__ void __launch_bounds__(kBlock) k_tau(Lev<T> C, const T *__restrict__ uold, T *bc)
{
    int i, j, k;
    long long c;
    if (!cell_of(C, i, j, k, c)) return;
    T dg, of;
    row<T, OP, VEL, REAC>(C, uold, i, j, k, c, dg, of);
    bc[c] = (dg * uold[c] + of) + bc[c];
}

template <typename T>
struct FacesT {
    const T *f[6];  // coarse BC faces for the ghost rule (nullable = 0)
};

// value of error_2h = unew - uold at a coarse index that may be a ghost,
// with update_boundary_ghostvalues applied to both (discretepde.f90:1144-1263);
// GIVEN: the ghost values as they stand in unew and uold (prolongation1 reads
// its argument's ghost layer, mgridsteps.f90:170-216)
template <typename T, bool GIVEN = false>
__device__ __forceinline__ T err_at(const Lev<T> &C, const T *unew, const T *uold, const FacesT<T> &G, int i, int j,
                                    int k)
{
    if (GIVEN) {
        const long long g = (long long)i + (long long)C.sy * j + C.sz * k;
        return unew[g] - uold[g];
    }
    int ii = i, jj = j, kk = k, face = -1;
    long long fidx = 0;
    if (i == 0) { ii = 1; face = LEFT; fidx = (j - 1) + (long long)C.ny * (k - 1); }
    else if (i == C.nx + 1) { ii = C.nx; face = RIGHT; fidx = (j - 1) + (long long)C.ny * (k - 1); }
    else if (j == 0) { jj = 1; face = BOTTOM; fidx = (i - 1) + (long long)C.nx * (k - 1); }
    else if (j == C.ny + 1) { jj = C.ny; face = TOP; fidx = (i - 1) + (long long)C.nx * (k - 1); }
    else if (k == 0) { kk = 1; face = BACK; fidx = (i - 1) + (long long)C.nx * (j - 1); }
    else if (k == C.nz + 1) { kk = C.nz; face = FRONT; fidx = (i - 1) + (long long)C.nx * (j - 1); }
    const long long c = (long long)ii + (long long)C.sy * jj + C.sz * kk;
    if (face < 0) return unew[c] - uold[c];
    if (C.bc[face] == BC_DIRC) {
        const T g2 = G.f[face] ? T(2) * G.f[face][fidx] : T(0);
        return (g2 - unew[c]) - (g2 - uold[c]);
    }
    return unew[c] - uold[c];
}

// prolongation1 + correction (mgridsteps.f90:170-216, :562-567)
template <typename T, bool GIVEN = false>
__global__ void __launch_bounds__(kBlock) k_prolong1_correct(Lev<T> F, Lev<T> C, T *uf, const T *__restrict__ unew,
                                                            const T *__restrict__ uold, FacesT<T> G)
{
    int i, j, k;
    long long c;
    if (!cell_of(C, i, j, k, c)) return;
    const T e = unew[c] - uold[c];
    const T sx = T(0.5) * (err_at<T, GIVEN>(C, unew, uold, G, i + 1, j, k) - err_at<T, GIVEN>(C, unew, uold, G, i - 1, j, k));
    const T sy_ = T(0.5) * (err_at<T, GIVEN>(C, unew, uold, G, i, j + 1, k) - err_at<T, GIVEN>(C, unew, uold, G, i, j - 1, k));
    const T sz_ = T(0.5) * (err_at<T, GIVEN>(C, unew, uold, G, i, j, k + 1) - err_at<T, GIVEN>(C, unew, uold, G, i, j, k - 1));
    const long long sy = F.sy, sz = F.sz;
    const long long f0 = (long long)(2 * i - 1) + sy * (2 * j - 1) + sz * (2 * k - 1);
    const T q = T(0.25);
    uf[f0] = uf[f0] + (e + q * (-sx - sy_ - sz_));
    uf[f0 + 1] = uf[f0 + 1] + (e + q * (+sx - sy_ - sz_));
    uf[f0 + sy] = uf[f0 + sy] + (e + q * (-sx + sy_ - sz_));
    uf[f0 + sy + 1] = uf[f0 + sy + 1] + (e + q * (+sx + sy_ - sz_));
    uf[f0 + sz] = uf[f0 + sz] + (e + q * (-sx - sy_ + sz_));
    uf[f0 + sz + 1] = uf[f0 + sz + 1] + (e + q * (+sx - sy_ + sz_));
    uf[f0 + sz + sy] = uf[f0 + sz + sy] + (e + q * (-sx + sy_ + sz_));
    uf[f0 + sz + sy + 1] = uf[f0 + sz + sy + 1] + (e + q * (+sx + sy_ + sz_));
}

// ------------------------------------------------------------------------
// aligned (m*2^k+1) transfers
// ------------------------------------------------------------------------

// volume-weighted full weighting: coarse I draws on fine 2I-2, 2I-1, 2I
template <typename T>
__global__ void __launch_bounds__(kBlock) k_restrict_fw(Lev<T> F, Lev<T> C, AlignedW W, const T *__restrict__ f,
                                                       T *cf)
{
    int i, j, k;
    long long c;
    if (!cell_of(C, i, j, k, c)) return;
    double acc = 0.0;
    for (int dk = -1; dk <= 1; ++dk) {
        const int fk = 2 * k - 1 + dk;
        const double wk = W.wz[3 * (k - 1) + dk + 1];
        if (wk == 0.0) continue;
        for (int dj = -1; dj <= 1; ++dj) {
            const int fj = 2 * j - 1 + dj;
            const double wjk = wk * W.wy[3 * (j - 1) + dj + 1];
            if (wjk == 0.0) continue;
            const long long base = (long long)F.sy * fj + F.sz * fk;
            double row = 0.0;
            for (int di = -1; di <= 1; ++di) {
                const double wi = W.wx[3 * (i - 1) + di + 1];
                if (wi != 0.0) row += wi * (double)f[base + 2 * i - 1 + di];
            }
            acc += wjk * row;
        }
    }
    cf[c] = (T)acc;
}

// trilinear prolongation + correction on the aligned hierarchy: fine cell f
// takes coarse (f+1)/2 when f is odd, the mean of f/2 and f/2+1 when even
template <typename T>
__global__ void __launch_bounds__(kBlock) k_prolong_tri_correct(Lev<T> F, Lev<T> C, T *uf, const T *__restrict__ uc)
{
    int i, j, k;
    long long c;
    if (!cell_of(F, i, j, k, c)) return;
    const int ia = (i + 1) >> 1, ja = (j + 1) >> 1, ka = (k + 1) >> 1;
    const int ib = (i & 1) ? ia : ia + 1, jb = (j & 1) ? ja : ja + 1, kb = (k & 1) ? ka : ka + 1;
    const long long sy = C.sy, sz = C.sz;
    auto at = [&](int a, int b_, int d) { return uc[(long long)a + sy * b_ + sz * d]; };
    T e;
    if (i & 1) {
        if (j & 1) {
            e = (k & 1) ? at(ia, ja, ka) : T(0.5) * (at(ia, ja, ka) + at(ia, ja, kb));
        } else {
            T a0 = T(0.5) * (at(ia, ja, ka) + at(ia, jb, ka));
            e = (k & 1) ? a0 : T(0.5) * (a0 + T(0.5) * (at(ia, ja, kb) + at(ia, jb, kb)));
        }
    } else {
        T r00 = T(0.5) * (at(ia, ja, ka) + at(ib, ja, ka));
        if (j & 1) {
            e = (k & 1) ? r00 : T(0.5) * (r00 + T(0.5) * (at(ia, ja, kb) + at(ib, ja, kb)));
        } else {
            T r10 = T(0.5) * (at(ia, jb, ka) + at(ib, jb, ka));
            T p0 = T(0.5) * (r00 + r10);
            if (k & 1) {
                e = p0;
            } else {
                T r01 = T(0.5) * (at(ia, ja, kb) + at(ib, ja, kb));
                T r11 = T(0.5) * (at(ia, jb, kb) + at(ib, jb, kb));
                e = T(0.5) * (p0 + T(0.5) * (r01 + r11));
            }
        }
    }
    uf[c] = uf[c] + e;
}

// ------------------------------------------------------------------------
// element-wise helpers (graph-safe replacements for memset/copies)
// ------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kBlock) k_fill(T *a, long long n, T v)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; t < n; t += stride) a[t] = v;
}

template <typename Ti, typename To>
__global__ void __launch_bounds__(kBlock) k_convert(const Ti *in, To *out, long long n)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; t < n; t += stride) out[t] = (To)in[t];
}

// interior correction u += e * escale[1] (defect correction: FP64 u, FP32 e
// computed for a right-hand side scaled by escale[0] = 1 / escale[1], both
// powers of two; escale nullable = no scaling)
template <typename T, typename E>
__global__ void __launch_bounds__(kBlock) k_add_interior(Lev<T> L, T *u, const E *e, const double *escale)
{
    int i, j, k;
    long long c;
    if (!cell_of(L, i, j, k, c)) return;
    const T s = escale ? (T)escale[1] : T(1);
    u[c] = u[c] + (T)e[c] * s;
}

// the defect-correction scale pair (2^-e, 2^e) with 2^e the binade of the rms
// sqrt(sumsq / n); (1, 1) for a zero or non-finite sum
__global__ void k_dc_scale(const double *sumsq, double n, double *scale)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double rms = sqrt(sumsq[0] / n);
    if (rms > 0.0 && isfinite(rms)) {
        const int e = ilogb(rms);
        scale[0] = ldexp(1.0, -e);
        scale[1] = ldexp(1.0, e);
    } else {
        scale[0] = 1.0;
        scale[1] = 1.0;
    }
}

// interior copy between a caller's interior array (nx,ny,nz) and the
// ghosted layout, either direction
// interior of a ghosted array into another ghosted array of the same level
template <typename T>
__global__ void __launch_bounds__(kBlock) k_copy_interior(Lev<T> L, const T *__restrict__ src, T *dst)
{
    int i, j, k;
    long long c;
    if (!cell_of(L, i, j, k, c)) return;
    dst[c] = src[c];
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_interior_in(Lev<T> L, const T *src, T *dst_ghosted)
{
    int i, j, k;
    long long c;
    if (!cell_of(L, i, j, k, c)) return;
    long long n = (long long)(i - 1) + (long long)L.nx * ((j - 1) + (long long)L.ny * (k - 1));
    dst_ghosted[c] = src[n];
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_interior_out(Lev<T> L, const T *src_ghosted, T *dst)
{
    int i, j, k;
    long long c;
    if (!cell_of(L, i, j, k, c)) return;
    long long n = (long long)(i - 1) + (long long)L.nx * ((j - 1) + (long long)L.ny * (k - 1));
    dst[n] = src_ghosted[c];
}

// update_boundary_ghostvalues (discretepde.f90:1121-1267) on one level;
// one thread per face element, faces in the reference order
__global__ void __launch_bounds__(kBlock) k_ghosts(Lev<double> L, double *u, Faces G)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nyz = (long long)L.ny * L.nz, nxz = (long long)L.nx * L.nz, nxy = (long long)L.nx * L.ny;
    const long long sy = L.sy, sz = L.sz;
    if (t < nyz) {
        int j = (int)(t % L.ny) + 1, k = (int)(t / L.ny) + 1;
        long long c0 = sy * j + sz * k;
        double g = G.f[LEFT] ? G.f[LEFT][t] : 0.0;
        u[c0] = L.bc[LEFT] == BC_DIRC ? 2.0 * g - u[c0 + 1] : u[c0 + 1];
        g = G.f[RIGHT] ? G.f[RIGHT][t] : 0.0;
        u[c0 + L.nx + 1] = L.bc[RIGHT] == BC_DIRC ? 2.0 * g - u[c0 + L.nx] : u[c0 + L.nx];
        return;
    }
    t -= nyz;
    if (t < nxz) {
        int i = (int)(t % L.nx) + 1, k = (int)(t / L.nx) + 1;
        long long c0 = i + sz * k;
        double g = G.f[BOTTOM] ? G.f[BOTTOM][t] : 0.0;
        u[c0] = L.bc[BOTTOM] == BC_DIRC ? 2.0 * g - u[c0 + sy] : u[c0 + sy];
        g = G.f[TOP] ? G.f[TOP][t] : 0.0;
        u[c0 + sy * (L.ny + 1)] = L.bc[TOP] == BC_DIRC ? 2.0 * g - u[c0 + sy * L.ny] : u[c0 + sy * L.ny];
        return;
    }
    t -= nxz;
    if (t < nxy) {
        int i = (int)(t % L.nx) + 1, j = (int)(t / L.nx) + 1;
        long long c0 = i + sy * j;
        double g = G.f[BACK] ? G.f[BACK][t] : 0.0;
        u[c0] = L.bc[BACK] == BC_DIRC ? 2.0 * g - u[c0 + sz] : u[c0 + sz];
        g = G.f[FRONT] ? G.f[FRONT][t] : 0.0;
        u[c0 + sz * (L.nz + 1)] = L.bc[FRONT] == BC_DIRC ? 2.0 * g - u[c0 + sz * L.nz] : u[c0 + sz * L.nz];
    }
}

// update_efield (plasma_pde/poissoneq.f90:565-680) on the interior: the
// central difference 0.5 (phi(+1) - phi(-1)) / d per direction, the minmod
// limiter (minmodlimiter, :549-563) when limiter != 0, E = -E * scale; phi's
// ghost values set beforehand. Same operation order as the reference.
__device__ __forceinline__ double minmod_phi(double ntr, double dtr)
{
    if (dtr != 0.0) {
        const double r = ntr / dtr;
        const double m = r < 1.0 ? r : 1.0;
        return m > 0.0 ? m : 0.0;
    }
    return 1.0;
}

//
// One thread per ghosted cell, so every row is written whole: an interior
// cell gets its own value; a face ghost cell (one index outside the
// interior) the value of the interior cell next to it, which is the
// zero-gradient ghost rule update_efield applies to each component
// (efield_bc_codes = 'ZGRD', :580-583, 670-678); edge and corner ghosts are
// not written (the ghost rule sets faces only).
__global__ void __launch_bounds__(kBlock) k_efield(Lev<double> L, const double *__restrict__ phi, double *ex,
                                                  double *ey, double *ez, double scale, int limiter)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int W = L.nx + 2;
    if (t >= W * (L.ny + 2)) return;
    const int j = t / W, i = t - j * W, k = blockIdx.y;
    const int ci = min(max(i, 1), L.nx), cj = min(max(j, 1), L.ny), ck = min(max(k, 1), L.nz);
    if ((ci != i) + (cj != j) + (ck != k) > 1) return;
    const long long o = (long long)t + L.sz * k;
    const long long c = (long long)ci + (long long)L.sy * cj + L.sz * ck;
    const long long st[3] = {1, L.sy, L.sz};
    const double d[3] = {L.dx, L.dy, L.dz};
    double *E[3] = {ex, ey, ez};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const double a = phi[c + st[q]], b = phi[c - st[q]], p0 = phi[c];
        double e = 0.5 * (a - b) / d[q];
        if (limiter) e = minmod_phi(p0 - b, a - p0) * e;
        E[q][o] = -e * scale;
    }
}

// explicit stencil of an OP_REF / OP_ALIGNED operator; transpose != 0 writes
// the transposed rows (A^T: the coefficient of u(i-1) in row i becomes the
// coefficient of u(i+1) in row i-1, and so on)
template <typename T, int OP, bool VEL, bool REAC>
__global__ void __launch_bounds__(kBlock) k_build_stencil(Lev<T> L, T *st0, T *st1, T *st2, T *st3, T *st4, T *st5,
                                                         T *st6, int transpose)
{
    int i, j, k;
    long long c;
    if (!cell_of(L, i, j, k, c)) return;
    T *st[ST_N] = {st0, st1, st2, st3, st4, st5, st6};
    T a[ST_N];
    row_coeffs<T, OP, VEL, REAC>(L, i, j, k, c, a);
    if (!transpose) {
        for (int q = 0; q < ST_N; ++q) st[q][c] = a[q];
        return;
    }
    // row c of A couples to c-1 with a[XM]: in A^T that is row c-1's
    // coefficient of c, i.e. slot XP of row c-1, and so on
    st[ST_C][c] = a[ST_C];
    if (i > 1) st[ST_XP][c - 1] = a[ST_XM];
    if (i < L.nx) st[ST_XM][c + 1] = a[ST_XP];
    if (j > 1) st[ST_YP][c - L.sy] = a[ST_YM];
    if (j < L.ny) st[ST_YM][c + L.sy] = a[ST_YP];
    if (k > 1) st[ST_ZP][c - L.sz] = a[ST_ZM];
    if (k < L.nz) st[ST_ZM][c + L.sz] = a[ST_ZP];
}

}  // namespace mps

namespace mps {

// ------------------------------------------------------------------------
// conductance coarsening for the aligned hierarchy (rows without
// convection)
// ------------------------------------------------------------------------
// face means of the finest level: f_d[c] = (D[c] + D[c + s_d]) / 2 on the +
// face of cell c in direction d (the reference face rule), 0 on the
// boundary
__global__ void __launch_bounds__(kBlock) k_face_means(Lev<double> L, double *fx, double *fy, double *fz)
{
    int i, j, k;
    long long c;
    if (!cell_of(L, i, j, k, c)) return;
    const double *D = L.D;
    fx[c] = i < L.nx ? 0.5 * (D[c] + D[c + 1]) : 0.0;
    fy[c] = j < L.ny ? 0.5 * (D[c] + D[c + L.sy]) : 0.0;
    fz[c] = k < L.nz ? 0.5 * (D[c] + D[c + L.sz]) : 0.0;
}

__device__ __forceinline__ double series(double a, double b) { return (a > 0.0 && b > 0.0) ? 2.0 * a * b / (a + b) : 0.0; }

// Coarse face coefficients: the + face of coarse cell I in x joins coarse
// centres that sit on fine centres 2I-1 and 2I+1, so it spans the fine
// faces (2I-1|2I) and (2I|2I+1) in series (the harmonic mean of their face
// coefficients: equal conductance over twice the distance); across the face
// the fine rows 2J-2..2J and planes 2K-2..2K conduct in parallel, weighted
// like the restriction. For a coefficient jump on a fine face this keeps
// the coarse conductance exact, where restricting cell coefficients smears
// the jump over the coarse cell.
__global__ void __launch_bounds__(kBlock) k_coarsen_faces(Lev<double> F, Lev<double> C, AlignedW W, const double *ffx,
                                                         const double *ffy, const double *ffz, double *cfx,
                                                         double *cfy, double *cfz)
{
    int I, J, K;
    long long c;
    if (!cell_of(C, I, J, K, c)) return;
    const long long sy = F.sy, sz = F.sz;
    auto at = [&](int a, int b, int d) { return (long long)a + sy * b + sz * d; };
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int p = -1; p <= 1; ++p) {
        for (int q = -1; q <= 1; ++q) {
            if (I < C.nx) {  // x face: transverse (j, k)
                const double w = W.wy[3 * (J - 1) + p + 1] * W.wz[3 * (K - 1) + q + 1];
                if (w != 0.0) {
                    const int fj = 2 * J - 1 + p, fk = 2 * K - 1 + q;
                    ax += w * series(ffx[at(2 * I - 1, fj, fk)], ffx[at(2 * I, fj, fk)]);
                }
            }
            if (J < C.ny) {  // y face: transverse (i, k)
                const double w = W.wx[3 * (I - 1) + p + 1] * W.wz[3 * (K - 1) + q + 1];
                if (w != 0.0) {
                    const int fi = 2 * I - 1 + p, fk = 2 * K - 1 + q;
                    ay += w * series(ffy[at(fi, 2 * J - 1, fk)], ffy[at(fi, 2 * J, fk)]);
                }
            }
            if (K < C.nz) {  // z face: transverse (i, j)
                const double w = W.wx[3 * (I - 1) + p + 1] * W.wy[3 * (J - 1) + q + 1];
                if (w != 0.0) {
                    const int fi = 2 * I - 1 + p, fj = 2 * J - 1 + q;
                    az += w * series(ffz[at(fi, fj, 2 * K - 1)], ffz[at(fi, fj, 2 * K)]);
                }
            }
        }
    }
    cfx[c] = ax;
    cfy[c] = ay;
    cfz[c] = az;
}

// 7-point rows of an aligned coarse level from its face coefficients: face
// weight (f / d) * abv, with abv the reciprocal width of the row's cell;
// DIRC walls D_c * idist * abv (the cell coefficient at the wall); then
// -reac and +1/dt, as the cell-coefficient rows of the aligned grids. Also
// the rows' face form: the face conductances f / d per direction (gx, gy,
// gz; the diagonal is st0)
__global__ void __launch_bounds__(kBlock) k_faces_stencil(Lev<double> L, const double *fx, const double *fy,
                                                         const double *fz, double *st0, double *st1, double *st2,
                                                         double *st3, double *st4, double *st5, double *st6,
                                                         double *gx, double *gy, double *gz)
{
    int i, j, k;
    long long c;
    if (!cell_of(L, i, j, k, c)) return;
    gx[c] = fx[c] / L.dx;
    gy[c] = fy[c] / L.dy;
    gz[c] = fz[c] / L.dz;
    double d = 0.0, w;
    const double Dc = L.D[c];
    {
        const double abv = (i == 1 || i == L.nx) ? L.axb : L.ax;
        w = i > 1 ? fx[c - 1] / L.dx * abv : 0.0;
        st1[c] = -w;
        d += w;
        w = i < L.nx ? fx[c] / L.dx * abv : 0.0;
        st2[c] = -w;
        d += w;
        if (i == 1 && L.bc[LEFT] == BC_DIRC) d += Dc * L.ix0 * abv;
        if (i == L.nx && L.bc[RIGHT] == BC_DIRC) d += Dc * L.ix0 * abv;
    }
    {
        const long long s = L.sy;
        const double abv = (j == 1 || j == L.ny) ? L.ayb : L.ay;
        w = j > 1 ? fy[c - s] / L.dy * abv : 0.0;
        st3[c] = -w;
        d += w;
        w = j < L.ny ? fy[c] / L.dy * abv : 0.0;
        st4[c] = -w;
        d += w;
        if (j == 1 && L.bc[BOTTOM] == BC_DIRC) d += Dc * L.iy0 * abv;
        if (j == L.ny && L.bc[TOP] == BC_DIRC) d += Dc * L.iy0 * abv;
    }
    {
        const long long s = L.sz;
        const double abv = (k == 1 || k == L.nz) ? L.azb : L.az;
        w = k > 1 ? fz[c - s] / L.dz * abv : 0.0;
        st5[c] = -w;
        d += w;
        w = k < L.nz ? fz[c] / L.dz * abv : 0.0;
        st6[c] = -w;
        d += w;
        if (k == 1 && L.bc[BACK] == BC_DIRC) d += Dc * L.iz0 * abv;
        if (k == L.nz && L.bc[FRONT] == BC_DIRC) d += Dc * L.iz0 * abv;
    }
    if (L.reac) d -= L.reac[c];
    if (L.timederiv) d += 1.0 / L.dt;
    st0[c] = d;
}

// The off-diagonal rows of a level from their face form (L.fc), in T's
// arithmetic: x- entry -(G_x[c - 1] * a), x+ entry -(G_x[c] * a), with a the
// row's reciprocal cell width (y, z alike). For FP64 these are k_faces_stencil's
// values; for the FP32 copy of a level, the values the face-form sweep
// (k_sweep_fast<T, OP_FACES>) forms, so every kernel on the level applies the
// same FP32 operator.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_faces_rows(Lev<T> L, T *st1, T *st2, T *st3, T *st4, T *st5, T *st6)
{
    int i, j, k;
    long long c;
    if (!cell_of(L, i, j, k, c)) return;
    const T ax = (i == 1 || i == L.nx) ? L.axb : L.ax;
    const T ay = (j == 1 || j == L.ny) ? L.ayb : L.ay;
    const T az = (k == 1 || k == L.nz) ? L.azb : L.az;
    st1[c] = -(L.fc[1][c - 1] * ax);
    st2[c] = -(L.fc[1][c] * ax);
    st3[c] = -(L.fc[2][c - L.sy] * ay);
    st4[c] = -(L.fc[2][c] * ay);
    st5[c] = -(L.fc[3][c - L.sz] * az);
    st6[c] = -(L.fc[3][c] * az);
}

// ------------------------------------------------------------------------
// Krylov vector kernels (FP64): fused multi-dot and multi-axpy over the
// interior of ghosted-layout vectors, with block partial sums reduced in a
// fixed order by k_reduce_cols.
// ------------------------------------------------------------------------
constexpr int kMaxVec = 64;

struct VecList {
    const double *v[kMaxVec];
};

// partials[q*nblocks + block] = sum over the block's cells of w*v_q
__global__ void __launch_bounds__(kBlock) k_multidot(Lev<double> L, const double *__restrict__ w, VecList V, int nv,
                                                    double *partials)
{
    int i, j;
    const bool ok = cell_ij(L, i, j);
    const int nblocks = gridDim.x * gridDim.y;
    const int bid = blockIdx.x + gridDim.x * blockIdx.y;
    for (int q = 0; q < nv; ++q) {
        double acc = 0.0;
        if (ok) {
            for (int k = blockIdx.y + 1; k <= L.nz; k += gridDim.y) {
                const long long c = (long long)i + (long long)L.sy * j + L.sz * k;
                acc += w[c] * V.v[q][c];
            }
        }
        double s = block_sum(acc);
        if (threadIdx.x == 0) partials[(long long)q * nblocks + bid] = s;
        __syncthreads();
    }
}

// out[q] = sum_b partials[q*nblocks + b]; one block per column
__global__ void __launch_bounds__(kBlock) k_reduce_cols(const double *partials, int nblocks, double *out)
{
    const int q = blockIdx.x;
    double s = 0.0;
    for (int t = threadIdx.x; t < nblocks; t += blockDim.x) s += partials[(long long)q * nblocks + t];
    s = block_sum(s);
    if (threadIdx.x == 0) out[q] = s;
}

// w = w - sum_q h[q] v_q over the interior (q < nv, and q < *nvd when nvd
// is given); block partial of |w|^2
__global__ void __launch_bounds__(kBlock) k_multiaxpy(Lev<double> L, double *w, VecList V, const double *h, int nv,
                                                     double *partials, const int *nvd = nullptr)
{
    int i, j;
    double sq = 0.0;
    if (nvd) nv = min(nv, *nvd);
    if (cell_ij(L, i, j)) {
        for (int k = blockIdx.y + 1; k <= L.nz; k += gridDim.y) {
            const long long c = (long long)i + (long long)L.sy * j + L.sz * k;
            double x = w[c];
            for (int q = 0; q < nv; ++q) x = x - h[q] * V.v[q][c];
            w[c] = x;
            sq += x * x;
        }
    }
    if (partials) {
        double s = block_sum(sq);
        if (threadIdx.x == 0) partials[blockIdx.x + (long long)gridDim.x * blockIdx.y] = s;
    }
}

// out = A u over the interior (ghosted output)
template <typename T, int OP, bool VEL, bool REAC>
__global__ void __launch_bounds__(kBlock) k_matvec(Lev<T> L, const T *__restrict__ u, T *out)
{
    int i, j, k;
    long long c;
    if (!cell_of(L, i, j, k, c)) return;
    T dg, of;
    row<T, OP, VEL, REAC>(L, u, i, j, k, c, dg, of);
    out[c] = dg * u[c] + of;
}

// v = w / s[0] over the interior (ghosts of v stay zero); nothing when s[0]
// is not positive (an Arnoldi breakdown or NaN leaves v as it is)
__global__ void __launch_bounds__(kBlock) k_scale(Lev<double> L, const double *w, double *v, const double *s)
{
    int i, j, k;
    long long c;
    if (!cell_of(L, i, j, k, c)) return;
    const double d = s[0];
    if (!(d > 0.0)) return;
    v[c] = w[c] / d;
}

// ------------------------------------------------------------------------
// FGMRES(m) Hessenberg and Givens on the device (one lane): the host queues
// the Arnoldi steps and reads the residual once per restart.
// ------------------------------------------------------------------------
struct ArnoldiDev {
    double *H;      // (m + 1) x m Hessenberg, column-major, leading dimension ld
    double *cs, *sn, *gv, *y;
    double *hn;     // |w| of the last step (the next basis vector's scale)
    int *state;     // [0] done, [1] columns kk, [2] 1 after a NaN or Inf
};

// start of a restart: gv = (beta, 0, ...), beta = sqrt(*beta2)
__global__ void k_arnoldi_begin(ArnoldiDev A, const double *beta2)
{
    if (threadIdx.x | blockIdx.x) return;
    A.state[0] = 0;
    A.state[1] = 0;
    A.state[2] = 0;
    A.gv[0] = sqrt(beta2[0]);
}

// Arnoldi step j after both Gram-Schmidt passes: column j of H is h + h2,
// H(j+1, j) = |w| = sqrt(*wn2); the previous rotations and a new one are
// applied, and the step marks the restart done when |gv(j+1)| <= tolb or
// |w| <= 1e-300 (breakdown), or flags a NaN. Nothing once done (steps
// queued past the end of a restart leave H and gv as they are).
__global__ void k_arnoldi_step(ArnoldiDev A, int j, int ld, const double *h, const double *h2, const double *wn2,
                               double tolb)
{
    if (threadIdx.x | blockIdx.x) return;
    if (A.state[0]) return;
    double *Hj = A.H + (long long)ld * j;
    for (int i = 0; i <= j; ++i) Hj[i] = h[i] + h2[i];
    const double hn = sqrt(wn2[0]);
    Hj[j + 1] = hn;
    A.hn[0] = hn;
    if (!isfinite(hn)) {
        A.state[2] = 1;
        A.state[0] = 1;
        return;
    }
    for (int i = 0; i < j; ++i) {
        const double t = A.cs[i] * Hj[i] + A.sn[i] * Hj[i + 1];
        Hj[i + 1] = -A.sn[i] * Hj[i] + A.cs[i] * Hj[i + 1];
        Hj[i] = t;
    }
    const double rr = Hj[j], ss = Hj[j + 1];
    const double den = hypot(rr, ss);
    A.cs[j] = den == 0.0 ? 1.0 : rr / den;
    A.sn[j] = den == 0.0 ? 0.0 : ss / den;
    Hj[j] = A.cs[j] * rr + A.sn[j] * ss;
    Hj[j + 1] = 0.0;
    A.gv[j + 1] = -A.sn[j] * A.gv[j];
    A.gv[j] = A.cs[j] * A.gv[j];
    A.state[1] = j + 1;
    if (fabs(A.gv[j + 1]) <= tolb || hn <= 1e-300) A.state[0] = 1;
}

// y = H(0:kk, 0:kk)^-1 gv by back substitution; ny = -y (the coefficients
// k_multiaxpy subtracts); nothing after a NaN
__global__ void k_arnoldi_solve(ArnoldiDev A, int ld, double *ny)
{
    if (threadIdx.x | blockIdx.x) return;
    if (A.state[2]) {
        A.state[1] = 0;
        return;
    }
    const int kk = A.state[1];
    for (int i = kk - 1; i >= 0; --i) {
        double s = A.gv[i];
        for (int jj = i + 1; jj < kk; ++jj) s -= A.H[i + (long long)ld * jj] * A.y[jj];
        A.y[i] = s / A.H[i + (long long)ld * i];
    }
    for (int i = 0; i < kk; ++i) ny[i] = -A.y[i];
}

}  // namespace mps
