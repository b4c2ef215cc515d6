// Fused red-black Gauss-Seidel sweep (gfx950).
//
// One launch performs a whole red-then-black sweep: the red update reads the
// old black values from `uin`, the black update reads the new red values,
// and both colours land in `uout` (ping-pong buffers, so no workgroup ever
// reads a value another workgroup is rewriting). The result is bit-identical
// to the two in-place colour passes of perform_gseidel_smoothing
// (mgridlib/discretepde.f90:202-228): every cell is evaluated by row7 with
// the same neighbour values.
//
// Blocking (2.5D): a workgroup owns a TX x TY column of cells and marches up
// a chunk of kz planes. Every global stream (u and D with a 2-cell xy halo,
// b and the reaction term with a 1-cell halo) is loaded into registers one
// plane ahead and staged into LDS rings, so no global-memory latency sits
// on the critical path of a plane. Red cells are computed on the tile plus
// a 1-cell ring (the ring is recomputed by the neighbouring tiles), one red
// cell per lane; black cells one plane behind on the tile. Ring depths
// (u 4, D 5, b 4, red 4) leave every slot that the previous plane still
// reads untouched, so each plane needs two workgroup barriers; two register
// sets keep two planes of loads in flight. Each sweep streams u, D and b
// once from HBM and writes u once.
#pragma once

#include "mps3d_kernels.h"

#ifndef MPS_SWEEP_ONEBAR
#define MPS_SWEEP_ONEBAR 0
#endif
#ifndef MPS_SWEEP_DMA
#define MPS_SWEEP_DMA 1  // LDS-direct plane staging: 0 off, 1 face-form rows, 2 all fast rows
#endif

namespace mps {

template <typename T, int OP, bool REAC, int TX, int TY>
__global__ void __launch_bounds__(TX *TY) k_sweep(Lev<T> L, const T *__restrict__ uin, T *__restrict__ uout,
                                                 const T *__restrict__ b, int kz)
{
    constexpr int W2 = TX + 4, H2 = TY + 4, P2 = W2 * H2;
    constexpr int W1 = TX + 2, H1 = TY + 2, P1 = W1 * H1;
    constexpr int NT = TX * TY;
    constexpr int N2 = (P2 + NT - 1) / NT;  // E2 slots per lane
    constexpr int N1 = (P1 + NT - 1) / NT;  // E1 slots per lane
    constexpr int RW = W1 / 2;              // red cells per E1 row
    constexpr int NR = RW * H1;             // red cells per E1 plane
    constexpr int RR = (NR + NT - 1) / NT;  // red cells per lane
    static_assert(W1 % 2 == 0, "tile width must be even");
    constexpr int DR = 5;                   // D ring depth
    __shared__ T su[4][P2];
    __shared__ T sD[DR][P2];
    __shared__ T sb[4][P1];
    __shared__ T sq[REAC ? 4 : 1][REAC ? P1 : 1];
    __shared__ T sr[4][P1];

    const int tid = threadIdx.x;
    int bx, by, bz;
    xcd_tile(bx, by, bz);
    const int i0 = bx * TX + 1, j0 = by * TY + 1;
    const int kb = bz * kz + 1;
    const int ke = min(kb + kz - 1, L.nz);
    const long long sy = L.sy, sz = L.sz;
    const T *__restrict__ Dg = L.D;
    const T *__restrict__ Rg = L.reac;

    // per-lane global offsets (within a plane) of the staged elements, -1
    // when outside the ghosted array
    long long o2[N2], o1[N1];
#pragma unroll
    for (int r = 0; r < N2; ++r) {
        const int q = tid + r * NT;
        o2[r] = -1;
        if (q < P2) {
            const int gi = i0 - 2 + q % W2, gj = j0 - 2 + q / W2;
            if (gi >= 0 && gi <= L.nx + 1 && gj >= 0 && gj <= L.ny + 1) o2[r] = (long long)gi + sy * gj;
        }
    }
#pragma unroll
    for (int r = 0; r < N1; ++r) {
        const int q = tid + r * NT;
        o1[r] = -1;
        if (q < P1) {
            const int gi = i0 - 1 + q % W1, gj = j0 - 1 + q / W1;
            if (gi >= 1 && gi <= L.nx && gj >= 1 && gj <= L.ny) o1[r] = (long long)gi + sy * gj;
        }
    }
    // two register sets: while plane s+1 is staged from one set, plane s+2
    // is in flight in the other, and the freed set then loads plane s+3
    T pu[2][N2], pD[2][N2], pb[2][N1], pq[2][N1];
    auto gload = [&](int kp, int set) {
        const bool okp = kp >= 0 && kp <= L.nz + 1;
        const bool okb = kp >= 1 && kp <= L.nz;
#pragma unroll
        for (int r = 0; r < N2; ++r) {
            T a = T(0), d = T(0);
            if (okp && o2[r] >= 0) {
                a = uin[o2[r] + sz * kp];
                d = Dg[o2[r] + sz * kp];
            }
            if (set) { pu[1][r] = a; pD[1][r] = d; } else { pu[0][r] = a; pD[0][r] = d; }
        }
#pragma unroll
        for (int r = 0; r < N1; ++r) {
            T a = T(0), d = T(0);
            if (okb && o1[r] >= 0) {
                a = b[o1[r] + sz * kp];
                if (REAC) d = Rg[o1[r] + sz * kp];
            }
            if (set) { pb[1][r] = a; pq[1][r] = d; } else { pb[0][r] = a; pq[0][r] = d; }
        }
    };
    auto lstore = [&](int kp, int set) {
        const int s4 = kp & 3, s5 = (kp + 5) % DR;
#pragma unroll
        for (int r = 0; r < N2; ++r) {
            const int q = tid + r * NT;
            if (q < P2) {
                su[s4][q] = set ? pu[1][r] : pu[0][r];
                sD[s5][q] = set ? pD[1][r] : pD[0][r];
            }
        }
#pragma unroll
        for (int r = 0; r < N1; ++r) {
            const int q = tid + r * NT;
            if (q < P1) {
                sb[s4][q] = set ? pb[1][r] : pb[0][r];
                if (REAC) sq[s4][q] = set ? pq[1][r] : pq[0][r];
            }
        }
    };

    // black-phase lane constants
    const int tx = tid % TX, ty = tid / TX;
    const int bi = i0 + tx, bj = j0 + ty;
    const bool bin = bi <= L.nx && bj <= L.ny;
    const int q1c = (tx + 1) + W1 * (ty + 1);
    const int e2c = (tx + 2) + W2 * (ty + 2);

    gload(kb - 2, 0);
    lstore(kb - 2, 0);
    gload(kb - 1, 1);
    lstore(kb - 1, 1);
    gload(kb, 0);
    gload(kb + 1, 1);
    for (int s = kb - 1; s <= ke + 1; ++s) {
        const int set = (s - kb + 1) & 1;  // the set that holds plane s+1
        lstore(s + 1, set);
        __syncthreads();
        if (s + 3 <= ke + 2) gload(s + 3, set);
        // red cells of plane s on the tile and its 1-cell ring, one per lane
        if (s >= 1 && s <= L.nz) {
            const T *um = su[(s - 1) & 3], *u0 = su[s & 3], *up = su[(s + 1) & 3];
            const T *Dm = sD[(s - 1 + DR) % DR], *D0 = sD[s % DR], *Dp = sD[(s + 1) % DR];
            const T *b0 = sb[s & 3];
            const T *q0 = sq[REAC ? (s & 3) : 0];
            T *rs = sr[s & 3];
#pragma unroll
            for (int r = 0; r < RR; ++r) {
                const int t = tid + r * NT;
                if (t >= NR) break;
                const int jj = t / RW;
                const int par = (i0 + j0 + jj + s + 1) & 1;  // (i + j + s) odd
                const int ii = 2 * (t - jj * RW) + par;
                const int i = i0 - 1 + ii, j = j0 - 1 + jj;
                if (i < 1 || i > L.nx || j < 1 || j > L.ny) continue;
                const int q = ii + W1 * jj;
                const int e = (ii + 1) + W2 * (jj + 1);
                T dg, of;
                row7<T, OP, REAC>(L, i, j, s, u0[e - 1], u0[e + 1], u0[e - W2], u0[e + W2], um[e], up[e], D0[e],
                                  D0[e - 1], D0[e + 1], D0[e - W2], D0[e + W2], Dm[e], Dp[e],
                                  REAC ? q0[q] : T(0), dg, of);
                rs[q] = gs_update(L, b0[q], of, dg);
            }
        }
        __syncthreads();
        // black cells of plane s-1 on the tile; plane s-1 is complete
        const int k = s - 1;
        if (k >= kb && k <= ke && bin) {
            const long long c = (long long)bi + sy * bj + sz * k;
            T val;
            if ((bi + bj + k) & 1) {
                val = sr[k & 3][q1c];
            } else {
                const T *rm = sr[(k - 1) & 3], *r0 = sr[k & 3], *rp = sr[(k + 1) & 3];
                const T *Dm = sD[(k - 1 + DR) % DR], *D0 = sD[k % DR], *Dp = sD[(k + 1) % DR];
                T dg, of;
                row7<T, OP, REAC>(L, bi, bj, k, r0[q1c - 1], r0[q1c + 1], r0[q1c - W1], r0[q1c + W1], rm[q1c],
                                  rp[q1c], D0[e2c], D0[e2c - 1], D0[e2c + 1], D0[e2c - W2], D0[e2c + W2], Dm[e2c],
                                  Dp[e2c], REAC ? sq[REAC ? (k & 3) : 0][q1c] : T(0), dg, of);
                val = gs_update(L, sb[k & 3][q1c], of, dg);
            }
            uout[c] = val;
        }
    }
}

// Dense coarsest-level solve u = A^-1 b with a precomputed inverse (row
// major, n x n, FP64), one workgroup.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_bottom_dense(Lev<T> L, T *u, const T *__restrict__ b,
                                                        const double *__restrict__ Ainv)
{
    __shared__ double sbv[kBottomMaxDense];
    const int nx = L.nx, ny = L.ny, n = L.nx * L.ny * L.nz;
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        const int i = q % nx + 1, j = (q / nx) % ny + 1, k = q / (nx * ny) + 1;
        sbv[q] = (double)b[(long long)i + (long long)L.sy * j + L.sz * k];
    }
    __syncthreads();
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        const double *row_ = Ainv + (long long)q * n;
        double s = 0.0;
        for (int p = 0; p < n; ++p) s += row_[p] * sbv[p];
        const int i = q % nx + 1, j = (q / nx) % ny + 1, k = q / (nx * ny) + 1;
        u[(long long)i + (long long)L.sy * j + L.sz * k] = (T)s;
    }
}

}  // namespace mps

namespace mps {

// reciprocal for the Gauss-Seidel update in fast arithmetic: the hardware
// estimate in FP32; in FP64 the estimate refined by two Newton steps
__device__ __forceinline__ float fast_rcp(float d) { return __builtin_amdgcn_rcpf(d); }
__device__ __forceinline__ double fast_rcp(double d)
{
    double r = __builtin_amdgcn_rcp(d);
    r = r * (2.0 - d * r);
    r = r * (2.0 - d * r);
    return r;
}

// compile-time register-set index
template <int I>
using SlotC = std::integral_constant<int, I>;

// Fused red-black sweep in fast arithmetic (rows without convection).
//
// Same schedule as k_sweep: plane s's red cells (on the tile and its ring,
// from old black values) and plane s-1's black cells (from new red values)
// per step, two barriers per step. Lanes are assigned by colour so that no
// wave runs both paths: the first (TX+2)(TY+2)/2 lanes each own one red cell
// of the tile-plus-ring row pair (its x position alternates with the plane
// parity) and write it out when it lies in the tile; the first TX*TY/2
// lanes own one black cell of the tile in the same way. Rows are
// branch-free: each lane precomputes the face factors abv/(2d) of its cells
// (zero on the physical boundary) and their DIRC wall factors, so a row is
// w_f = (D_c + D_nb) h_f,  diag = sum w_f + D_c g + 1/dt - reac,
// u = (b + sum w_f u_nb) / diag.  Values outside the interior are staged as
// zeros, so masked faces multiply zeros. LDS holds rings of four u planes
// and eight D planes of the tile with a 2-cell halo and four planes of red
// results; each lane's own b values travel in registers. Ring depths: step
// s stores u/D plane s+1 before its first barrier, and its red pass writes
// red plane s after it. The u slot it reuses (plane s-3) was last read by
// the red pass of step s-2 and the red slot (plane s-4) by the black pass
// of step s-2, both behind a barrier; the D slot (plane s-7) was last read
// by the black pass of step s-6 (the black pass of step s-1, which follows
// no barrier before step s's store, reads D planes s-3..s-1, so the D ring
// needs at least five). The depths are powers of two and the march is
// unrolled by eight, so every ring slot is a compile-time constant and
// every LDS access a per-lane base plus an immediate offset. Global loads
// and stores are raw buffer accesses with per-lane offsets fixed for the
// launch (out-of-interior positions load 0 and drop their stores) and run
// one step ahead in two alternating register sets.
// (MPS_SWEEP_WAVES, a build-time tuning knob, sets the waves per SIMD the
// register allocation aims at)
#ifdef MPS_SWEEP_WAVES
#define MPS_SWEEP_OCC __attribute__((amdgpu_waves_per_eu(MPS_SWEEP_WAVES, MPS_SWEEP_WAVES)))
#else
#define MPS_SWEEP_OCC
#endif
template <typename T, int OP, bool REAC, int TX, int TY, bool ZG>
__global__ void __launch_bounds__(TX *TY) MPS_SWEEP_OCC k_sweep_fast(Lev<T> L, const T *__restrict__ uin, T *__restrict__ uout,
                                                      const T *__restrict__ b, int kz, T zg0, T zg1)
{
#pragma clang fp contract(fast)
    constexpr int W2 = TX + 4, H2 = TY + 4, P2 = W2 * H2;
    constexpr int W1 = TX + 2, H1 = TY + 2, P1 = W1 * H1;
    constexpr int NT = TX * TY;
    constexpr int N2 = (P2 + NT - 1) / NT;
    constexpr int N1 = (P1 + NT - 1) / NT;
    constexpr int RW = W1 / 2;   // red cells per ring row
    constexpr int NR = RW * H1;  // red lanes
    static_assert(TX % 2 == 0, "tile width must be even");
    static_assert(NR <= NT, "one red cell per lane");
    static_assert((TX / 2 + 1) * (TY + 1) <= NT, "one black cell per lane, extended tile included");
    // OB (FP32, MPS_SWEEP_ONEBAR): one barrier per z-step, the black pass two
    // planes behind the red one in the same barrier interval (so the u and
    // reac rings hold 8 planes: the next plane is staged while other waves
    // still read planes s - 4 .. s); else two barriers, black one plane
    // behind
    constexpr bool OB = MPS_SWEEP_ONEBAR && sizeof(T) == 4;
    constexpr int RU = OB ? 8 : 4, RD = 8, RR = 4;   // ring depths (powers of two: slots are t & (R - 1))
    constexpr bool STN = OP == OP_STENCIL || OP == OP_FACES;  // explicit 7-point rows (no D tiles)
    // OP_FACES: the rows' face form (L.fc): per cell the diagonal and the
    // conductances of its six faces (the - face of a direction is the +
    // face of the neighbour), so four arrays are streamed instead of seven;
    // G_x and G_y are staged in LDS with each plane (a cell's - faces are
    // its LDS neighbours), the diagonal and G_z are read per cell. Each
    // coefficient is formed as -(G * a), which is the value k_faces_rows
    // stores for the level's other kernels
    constexpr bool FCS = OP == OP_FACES;
    // ZGHOST (OP_REF rows): the z faces of planes 1 and nz take the interior
    // factor hz like every other plane, and the staged D of the ghost plane
    // carries the boundary instead: zg * D of the adjacent interior plane,
    // with zg = -1 on a no-flux face (the face weight (D_c + D_g) hz is 0)
    // and gz/hz - 1 on a DIRC face (the weight is D_c gz, the wall term of
    // the diagonal); the ghost plane of u is staged as 0.
    constexpr bool ZGHOST = OP == OP_REF;
    // DL (FP32 face-form rows; with MPS_SWEEP_DMA=2 also the face-factor
    // rows, where it measured 2 % slower at 257^3): the staged planes go
    // from HBM straight into their LDS ring slots (buffer loads
    // with the LDS flag: lane l of a wave writes word l after the wave's
    // uniform LDS base), issued one step ahead like the register sets; each
    // wave waits for its own before the barrier that publishes them. Rows
    // are padded to whole waves (the padding takes the zeros that lanes past
    // the halo plane load). Planes outside 1..nz are staged with LDS stores
    // instead (zeros, or ZGHOST's scaled D).
    constexpr bool DL = sizeof(T) == 4 && !OB && !REAC && ((MPS_SWEEP_DMA >= 1 && FCS) || (MPS_SWEEP_DMA >= 2 && !STN));
    constexpr int P2L = DL ? (P2 + 63) / 64 * 64 : P2;
    static_assert(!DL || NT % 64 == 0, "whole waves");
    __shared__ T su[RU][P2L];
    // (OP_FACES stages its x and y face conductances with u: G_x in sD,
    // G_y in sGy, in u's ring slots, so a cell's - faces are its LDS
    // neighbours)
    __shared__ T sD[FCS ? RU : (STN ? 1 : RD)][(STN && !FCS) ? 1 : P2L];
    __shared__ T sGy[FCS ? RU : 1][FCS ? P2L : 1];
    __shared__ T sq[REAC ? RU : 1][REAC ? P1 : 1];
    __shared__ T sr[RR][P1 + W1];   // (one spare row: an extended tile's last black row reads one below)

    const int tid = threadIdx.x;
    int bx, by, bz;
    xcd_tile(bx, by, bz);
    const int i0 = bx * TX + 1, j0 = by * TY + 1;
    // the gridDim.z = ceil(nz / kz) chunks share the planes evenly (lengths
    // differ by at most one, none above kz), so no chunk is a short tail
    const int zb = L.nz / (int)gridDim.z, zr = L.nz % (int)gridDim.z;
    const int kb = bz * zb + min(bz, zr) + 1;
    const int ke = kb + zb - (bz < zr ? 0 : 1);
    const int sy = L.sy;
    const long long sz = L.sz;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    // Extended tiles: on a grid of k TX + 1 columns (the 2^m + 1 levels of
    // the aligned hierarchy) the host launches k tiles in x and the last one
    // also owns column nx, which is its ring column: its red cells there are
    // written out, and one more black pair per row covers it. Likewise the
    // last row. (The fields beyond nx and ny are outside the interior: zero
    // face factors, staged zeros.)
    const bool xext = bx == (int)gridDim.x - 1 && i0 + TX == nx;
    const bool yext = by == (int)gridDim.y - 1 && j0 + TY == ny;
    const int txe = TX + (xext ? 1 : 0), tye = TY + (yext ? 1 : 0);  // owned columns and rows
    // The red ring is zeroed once: its black-parity positions are never
    // written (each slot holds planes of one parity), and an extended
    // tile's black cells on column nx or row ny read one position past
    // their row or plane there, with a zero face factor.
    for (int q = tid; q < RR * (P1 + W1); q += NT) (&sr[0][0])[q] = T(0);

    // Every global access goes through buffer resources over the planes
    // this block touches, pl0..pl1 (its chunk and two planes either side,
    // clamped to the ghost planes; the host keeps that slab below 2^31
    // bytes): a per-lane in-plane byte offset fixed for the whole march
    // (kOob outside the interior: the load returns 0, the store is dropped)
    // and a uniform plane offset, so loads and stores need no address
    // arithmetic or selects per step.
    const int pl0 = max(kb - 2, 0), pl1 = min(ke + 2, nz + 1);
    const int szb = (int)sz * (int)sizeof(T);
    const int nrec = (pl1 - pl0 + 1) * szb;
    auto rsrc = [&](const T *p) { return buf_rsrc(p + (long long)pl0 * sz, nrec); };
    const __amdgpu_buffer_rsrc_t Ru = rsrc(uin), Ro = rsrc(uout), Rb = rsrc(b);
    const __amdgpu_buffer_rsrc_t Rd = rsrc(STN ? b : L.D), Rq = rsrc(REAC ? L.reac : b);
    __amdgpu_buffer_rsrc_t Rs[STN ? ST_N : 1];
#pragma unroll
    for (int q = 0; q < (FCS ? 4 : STN ? ST_N : 1); ++q) Rs[q] = rsrc(FCS ? L.fc[q] : STN ? L.st[q] : b);
    // byte offset of plane kp in the resources (kp clamped to pl0..pl1)
    auto pofs = [&](int kp) { return (min(max(kp, pl0), pl1) - pl0) * szb; };

    // staging offsets: in-plane bytes of interior positions, kOob elsewhere
    int v2[N2], v1[N1];
#pragma unroll
    for (int r = 0; r < N2; ++r) {
        const int q = tid + r * NT;
        v2[r] = kOob;
        if (q < P2) {
            const int gi = i0 - 2 + q % W2, gj = j0 - 2 + q / W2;
            if (gi >= 1 && gi <= nx && gj >= 1 && gj <= ny) v2[r] = (gi + sy * gj) * (int)sizeof(T);
        }
    }
#pragma unroll
    for (int r = 0; r < N1; ++r) {
        const int q = tid + r * NT;
        v1[r] = kOob;
        if (REAC && q < P1) {
            const int gi = i0 - 1 + q % W1, gj = j0 - 1 + q / W1;
            if (gi >= 1 && gi <= nx && gj >= 1 && gj <= ny) v1[r] = (gi + sy * gj) * (int)sizeof(T);
        }
    }

    // x/y face factors of a cell (masked at the physical boundary) and its
    // x/y DIRC wall factor
    auto xy_factors = [&](int i, int j, T &fxm, T &fxp, T &fym, T &fyp, T &gxy) {
        const bool ex = OP == OP_ALIGNED && (i == 1 || i == nx);
        const bool ey = OP == OP_ALIGNED && (j == 1 || j == ny);
        const T hx = ex ? L.hxb : L.hx, hy = ey ? L.hyb : L.hy;
        fxm = i > 1 ? hx : T(0);
        fxp = i < nx ? hx : T(0);
        fym = j > 1 ? hy : T(0);
        fyp = j < ny ? hy : T(0);
        const T gx = ex ? L.gxb : L.gx, gy = ey ? L.gyb : L.gy;
        T g = T(0);
        if (i == 1 && L.bc[LEFT] == BC_DIRC) g += gx;
        if (i == nx && L.bc[RIGHT] == BC_DIRC) g += gx;
        if (j == 1 && L.bc[BOTTOM] == BC_DIRC) g += gy;
        if (j == ny && L.bc[TOP] == BC_DIRC) g += gy;
        gxy = g;
    };
    auto z_factors = [&](int k, T &fzm, T &fzp, T &gz) {
        const bool ez = OP == OP_ALIGNED && (k == 1 || k == nz);
        const T hz = ez ? L.hzb : L.hz;
        fzm = k > 1 ? hz : T(0);
        fzp = k < nz ? hz : T(0);
        const T g1 = ez ? L.gzb : L.gz;
        T g = T(0);
        if (k == 1 && L.bc[BACK] == BC_DIRC) g += g1;
        if (k == nz && L.bc[FRONT] == BC_DIRC) g += g1;
        gz = g;
    };

    // Two-element per-lane quantities are scalar pairs (suffix 0/1 = the
    // cell of plane parity 0/1), not arrays: a select between array
    // elements may become a dynamic index, which moves the array out of
    // registers.
    //
    // red lane: ring row jr, pair mr; its cell sits at ring column 2mr + p,
    // p = (rpar + k) & 1 on plane k
    const bool red_lane = tid < NR;
    const int jr = tid / RW, mr = tid - jr * RW;
    const int jrg = j0 - 1 + jr;
    const int rpar = (i0 + j0 + jr + 1) & 1;
    const int ii0 = 2 * mr, ii1 = 2 * mr + 1;
    const int i_r0 = i0 - 1 + ii0, i_r1 = i0 - 1 + ii1;
    const int rq0 = ii0 + W1 * jr, rq1 = ii1 + W1 * jr;
    const int re0 = (ii0 + 1) + W2 * (jr + 1), re1 = (ii1 + 1) + W2 * (jr + 1);
    const bool jin = red_lane && jrg >= 1 && jrg <= ny;
    const int ro0 = (jin && i_r0 >= 1 && i_r0 <= nx) ? i_r0 + sy * jrg : -1;
    const int ro1 = (jin && i_r1 >= 1 && i_r1 <= nx) ? i_r1 + sy * jrg : -1;
    // the red cell lies in this block's tile (it is written out by this lane)
    const bool rowin = jr >= 1 && jr <= tye;
    const bool rt0 = ro0 >= 0 && rowin && ii0 >= 1 && ii0 <= txe;
    const bool rt1 = ro1 >= 0 && rowin && ii1 >= 1 && ii1 <= txe;
    T rfxm0, rfxp0, rgxy0, rfxm1, rfxp1, rgxy1, rfym, rfyp;
    xy_factors(i_r0, jrg, rfxm0, rfxp0, rfym, rfyp, rgxy0);
    xy_factors(i_r1, jrg, rfxm1, rfxp1, rfym, rfyp, rgxy1);
    // black lane: tile row yb, pair mb; its cell sits at tile column
    // 2mb + pb, pb = (bpar + k) & 1 on plane k
    const int bw = (txe + 1) / 2;  // black pairs per row
    const bool blk_lane = tid < bw * tye;
    const int yb = tid / bw, mb = tid - yb * bw;
    const int bj = j0 + yb;
    const int bpar = (i0 + bj) & 1;
    const int tb0 = 2 * mb, tb1 = 2 * mb + 1;
    const int bi0 = i0 + tb0, bi1 = i0 + tb1;
    const bool bjin = blk_lane && bj <= ny;
    const int cb0 = (bjin && bi0 <= nx) ? bi0 + sy * bj : -1;
    const int cb1 = (bjin && bi1 <= nx) ? bi1 + sy * bj : -1;
    const int bq0 = (tb0 + 1) + W1 * (yb + 1), bq1 = (tb1 + 1) + W1 * (yb + 1);
    const int be0 = (tb0 + 2) + W2 * (yb + 2), be1 = (tb1 + 2) + W2 * (yb + 2);
    T bfxm0, bfxp0, bgxy0, bfxm1, bfxp1, bgxy1, bfym, bfyp;
    xy_factors(bi0, bj, bfxm0, bfxp0, bfym, bfyp, bgxy0);
    xy_factors(bi1, bj, bfxm1, bfxp1, bfym, bfyp, bgxy1);

    // The loop runs step s in register slot (s - kb + 1) & 1, so each slot
    // always sees the same plane parity, and each lane the same red and
    // black cell: resolve the column parity per slot once, here, instead of
    // selecting on every step. Slot y's red cell is column parity
    // (rpar + kb - 1 + y) & 1, its black cell (bpar + kb - 2 + y) & 1.
    struct Cell {
        int q, e, o;    // LDS index (ring plane), LDS index (halo-2 plane), field offset (-1: none)
        int vb, vs;     // in-plane bytes of the cell for loads and for the store (kOob: none)
        T fxm, fxp, g;  // x face factors and x/y DIRC factor
        T ax, ay;       // OP_FACES: the row's reciprocal cell widths in x and y
    };
    auto vbytes = [](int o) { return o >= 0 ? o * (int)sizeof(T) : kOob; };
    auto cell = [&](int q, int e, int o, bool own, T fxm, T fxp, T g, int i, int j) {
        const T ax = (i == 1 || i == nx) ? L.axb : L.ax, ay = (j == 1 || j == ny) ? L.ayb : L.ay;
        return Cell{q, e, o, vbytes(o), own ? vbytes(o) : kOob, fxm, fxp, g, ax, ay};
    };
    auto red_cell = [&](int pc) {
        return pc ? cell(rq1, re1, ro1, rt1, rfxm1, rfxp1, rgxy1, i_r1, jrg)
                  : cell(rq0, re0, ro0, rt0, rfxm0, rfxp0, rgxy0, i_r0, jrg);
    };
    auto blk_cell = [&](int pc) {
        return pc ? cell(bq1, be1, cb1, true, bfxm1, bfxp1, bgxy1, bi1, bj)
                  : cell(bq0, be0, cb0, true, bfxm0, bfxp0, bgxy0, bi0, bj);
    };
    const int pr = (rpar + kb - 1) & 1, pbk = (bpar + kb - 2 - (OB ? 1 : 0)) & 1;
    const Cell R0 = red_cell(pr), R1 = red_cell(pr ^ 1);
    const Cell B0 = blk_cell(pbk), B1 = blk_cell(pbk ^ 1);

    // gload(kp, X): into register set X, the u and D tiles of plane kp (for
    // LDS), the red-cell b of plane kp - 1 and the black-cell b of plane
    // kp - 2: everything step kp - 1 consumes. Step s consumes gload(s + 1)
    // and issues gload(s + 2), so loads have a step to arrive; the two sets
    // alternate with the step parity (the march is unrolled, so no set is
    // copied while its loads are in flight). Positions outside
    // the interior load 0; b of a cell outside the interior or of a plane
    // outside 1..nz is never consumed.
    T pu[2][N2], pD[2][N2], pq[2][N1], nbr0, nbr1, nbb0, nbb1;
    T pG[FCS ? 2 : 1][FCS ? N2 : 1];
    // stencil rows: the 7 coefficients of the slot's red and black cell
    T rs0[STN ? ST_N : 1], rs1[STN ? ST_N : 1], bs0[STN ? ST_N : 1], bs1[STN ? ST_N : 1];
    bool okp0 = false, okp1 = false;
    T dm0 = T(0), dm1 = T(0);
    // DL: plane kp of u and D (FCS: u, G_x and G_y) into u slot US and D
    // slot DS, as direct LDS loads, or, outside 1..nz, as LDS stores of
    // zeros (ZGHOST: D as zg0 / zg1 times the first / last interior plane,
    // a register load waited for in place, at most twice per block)
    auto stage = [&](int kp, auto US, auto DS) {
        constexpr int us = decltype(US)::value, ds = FCS ? us : decltype(DS)::value;
        T *const du = su[us];
        T *const dd = sD[ds];
        T *const dg = sGy[FCS ? us : 0];
        if (kp >= 1 && kp <= nz) {
            const int pk = pofs(kp);
            const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
#pragma unroll
            for (int r = 0; r < N2; ++r) {
                const int q0 = r * NT + 64 * w;
                if (q0 < P2) {
                    if constexpr (!ZG) buf_load_lds(Ru, (float *)(du + q0), v2[r], pk);
                    if constexpr (FCS) {
                        buf_load_lds(Rs[1], (float *)(dd + q0), v2[r], pk);
                        buf_load_lds(Rs[2], (float *)(dg + q0), v2[r], pk);
                    } else {
                        buf_load_lds(Rd, (float *)(dd + q0), v2[r], pk);
                    }
                }
            }
        } else {
            const T dmv = (ZGHOST && kp == 0) ? zg0 : (ZGHOST && kp == nz + 1) ? zg1 : T(0);
            const int pkd = ZGHOST ? pofs(min(max(kp, 1), nz)) : pofs(kp);
#pragma unroll
            for (int r = 0; r < N2; ++r) {
                const int q = tid + r * NT;
                if (r * NT + NT <= P2L || q < P2L) {
                    if constexpr (!ZG) du[q] = T(0);
                    if constexpr (FCS) {
                        dd[q] = T(0);
                        dg[q] = T(0);
                    } else {
                        dd[q] = ZGHOST ? buf_load<T>(Rd, v2[r], pkd) * dmv : T(0);
                    }
                }
            }
        }
    };
    // DL: the register loads gload issues after a stage (the b values and
    // the face-form row's diagonal and z faces): a wait for at most that
    // many outstanding vector-memory operations completes the stage
    constexpr int KREG = FCS ? 8 : 2;
    auto gload = [&](int kp, auto X) {
        constexpr int x = decltype(X)::value;
        const int pk = pofs(kp);
        // OP_REF: the z ghost planes of D are staged as zg0 / zg1 times the
        // first / last interior plane (see ZGHOST below)
        const int pkd = ZGHOST ? pofs(min(max(kp, 1), nz)) : pk;
        if constexpr (!DL) {
#pragma unroll
            for (int r = 0; r < N2; ++r) {
                pu[x][r] = ZG ? T(0) : buf_load<T>(Ru, v2[r], pk);  // ZG: sweep from a zero guess (uin unread)
                if constexpr (FCS) {
                    pD[x][r] = buf_load<T>(Rs[1], v2[r], pk);
                    pG[x][r] = buf_load<T>(Rs[2], v2[r], pk);
                } else if (!STN) {
                    pD[x][r] = buf_load<T>(Rd, v2[r], pkd);
                }
            }
        }
        if (REAC && !STN) {
#pragma unroll
            for (int r = 0; r < N1; ++r) pq[x][r] = buf_load<T>(Rq, v1[r], pk);
        }
        // (gload(kp) fills set (kp - kb) & 1, the slot of step kp - 1 that
        // consumes it; outside that pattern the b values are not consumed)
        // (the black plane of step kp - 1 is kp - 2, or kp - 3 with OB)
        const int pr = pofs(kp - 1), pb = pofs(kp - 2 - (OB ? 1 : 0));
        const int vr_ = x ? R1.vb : R0.vb;
        const int vb_ = x ? B1.vb : B0.vb;
        const T vr = buf_load<T>(Rb, vr_, pr);
        const T vb = buf_load<T>(Rb, vb_, pb);
        if constexpr (FCS) {
            // the diagonal and the z face conductances of the cell (the x
            // and y ones are staged in LDS with the plane)
            const int pr2 = pofs(kp - 2), pb2 = pofs(kp - 3 - (OB ? 1 : 0));
            T *ra = x ? rs1 : rs0, *ba = x ? bs1 : bs0;
            ra[ST_C] = buf_load<T>(Rs[0], vr_, pr);
            ra[ST_ZM] = buf_load<T>(Rs[3], vr_, pr2);
            ra[ST_ZP] = buf_load<T>(Rs[3], vr_, pr);
            ba[ST_C] = buf_load<T>(Rs[0], vb_, pb);
            ba[ST_ZM] = buf_load<T>(Rs[3], vb_, pb2);
            ba[ST_ZP] = buf_load<T>(Rs[3], vb_, pb);
        } else if constexpr (STN) {
#pragma unroll
            for (int q = 0; q < ST_N; ++q) {
                const T sr_ = buf_load<T>(Rs[q], vr_, pr), sb_ = buf_load<T>(Rs[q], vb_, pb);
                if (x) {
                    rs1[q] = sr_;
                    bs1[q] = sb_;
                } else {
                    rs0[q] = sr_;
                    bs0[q] = sb_;
                }
            }
        }
        const bool ok = kp >= 1 && kp <= nz;
        const T dm = ok ? T(1) : (ZGHOST && kp == 0) ? zg0 : (ZGHOST && kp == nz + 1) ? zg1 : T(0);
        if (x) {
            nbr1 = vr, nbb1 = vb, okp1 = ok, dm1 = dm;
        } else {
            nbr0 = vr, nbb0 = vb, okp0 = ok, dm0 = dm;
        }
    };
    // stores plane kp (u into ring slot s3, D into s5, reac into s4)
    auto lstore = [&](int s3, int s5, int s4, auto X) {
        constexpr int x = decltype(X)::value;
        const bool ok = x ? okp1 : okp0;
        const T dm = x ? dm1 : dm0;
#pragma unroll
        for (int r = 0; r < N2; ++r) {
            const int q = tid + r * NT;
            if (r * NT + NT <= P2 || q < P2) {
                su[s3][q] = ok ? pu[x][r] : T(0);
                if constexpr (FCS) {
                    sD[s5][q] = pD[x][r];
                    sGy[s5][q] = pG[x][r];
                } else if (!STN) {
                    sD[s5][q] = pD[x][r] * dm;
                }
            }
        }
        if (REAC && !STN) {
#pragma unroll
            for (int r = 0; r < N1; ++r) {
                const int q = tid + r * NT;
                if (q < P1) sq[s4][q] = ok ? pq[x][r] : T(0);
            }
        }
    };
    // one fast row: returns the Gauss-Seidel value
    auto fast_row = [&](T Dc, T Dxm, T Dxp, T Dym, T Dyp, T Dzm, T Dzp, T uxm, T uxp, T uym, T uyp, T uzm, T uzp,
                        T fxm, T fxp, T fym, T fyp, T fzm, T fzp, T g, T rc, T bv) {
        const T wxm = (Dxm + Dc) * fxm, wxp = (Dc + Dxp) * fxp;
        const T wym = (Dym + Dc) * fym, wyp = (Dc + Dyp) * fyp;
        const T wzm = (Dzm + Dc) * fzm, wzp = (Dc + Dzp) * fzp;
        T d = ((wxm + wxp) + (wym + wyp)) + (wzm + wzp);
        d = d + Dc * g + L.dtinv;
        if (REAC) d = d - rc;
        T s = bv + wxm * uxm;
        s = s + wxp * uxp;
        s = s + wym * uym;
        s = s + wyp * uyp;
        s = s + wzm * uzm;
        s = s + wzp * uzp;
        return s * fast_rcp(d);
    };

    // one stencil row (the arithmetic of row() and gs_update for explicit
    // rows): neighbours in the order x-, x+, y-, y+, z-, z+
    auto stencil_row = [&](const T *st, T uxm, T uxp, T uym, T uyp, T uzm, T uzp, T bv) {
        T o = T(0);
        o = o + st[ST_XM] * uxm;
        o = o + st[ST_XP] * uxp;
        o = o + st[ST_YM] * uym;
        o = o + st[ST_YP] * uyp;
        o = o + st[ST_ZM] * uzm;
        o = o + st[ST_ZP] * uzp;
        return gs_update(L, bv, o, st[ST_C]);
    };
    // the same row from its face form: g holds the diagonal and the face
    // conductances in ST_* order, each coefficient is -(G * a)
    auto faces_row = [&](const T *g, T ax, T ay, T az, T uxm, T uxp, T uym, T uyp, T uzm, T uzp, T bv) {
        T o = T(0);
        o = o + (-(g[ST_XM] * ax)) * uxm;
        o = o + (-(g[ST_XP] * ax)) * uxp;
        o = o + (-(g[ST_YM] * ay)) * uym;
        o = o + (-(g[ST_YP] * ay)) * uyp;
        o = o + (-(g[ST_ZM] * az)) * uzm;
        o = o + (-(g[ST_ZP] * az)) * uzp;
        return gs_update(L, bv, o, g[ST_C]);
    };
    // z face factors of interior planes; planes 1 and nz get their own
    T fzm_i, fzp_i, gz_i;
    z_factors(2, fzm_i, fzp_i, gz_i);
    if (nz < 3) z_factors(nz, fzm_i, fzp_i, gz_i);
    auto zf = [&](int k, T &fzm, T &fzp, T &gz) {
        if (k == 1 || k == nz) {
            z_factors(k, fzm, fzp, gz);
        } else {
            fzm = fzm_i, fzp = fzp_i, gz = gz_i;
        }
    };
    // Ring slots are compile-time: the march is unrolled by eight steps and
    // the rings hold RU = 4 planes (u, red results, reac) and RD = 8 (D), so
    // step t of a group (plane s, t = (s - kb + 1) mod 8) finds plane s in
    // u/red/reac slot (t + 1) & 3 and D slot (t + 1) & 7, and every LDS
    // access is a per-lane base plus an immediate offset. Plane kb - 2 is
    // staged into slots 0, plane kb - 1 into slots 1.
    if constexpr (DL) {
        // (ZG: the u ring stays zero)
        if constexpr (ZG)
            for (int q = tid; q < RU * P2L; q += NT) (&su[0][0])[q] = T(0);
        stage(kb - 2, SlotC<0>{}, SlotC<0>{});
        stage(kb - 1, SlotC<1>{}, SlotC<1>{});
        stage(kb, SlotC<2>{}, SlotC<2>{});
        __builtin_amdgcn_sched_barrier(0);
    } else {
        gload(kb - 2, SlotC<0>{});
        lstore(0, 0, 0, SlotC<0>{});
        gload(kb - 1, SlotC<0>{});
        lstore(1, 1, 1, SlotC<0>{});
    }
    // set of gload(kp): (kp - kb) & 1
    gload(kb, SlotC<0>{});
    // (DL: the compiler may drop prologue loads whose values the first
    // steps provably skip, so the steps' count of loads issued after a stage
    // does not hold here: wait for everything)
    if constexpr (DL) wait_vmcnt<0>();
    auto step = [&](int s, auto Tc) {
        constexpr int t = decltype(Tc)::value;
        constexpr int x = t & 1;
        // u / reac slots of planes s + 1, s, s - 1
        constexpr int up_ = (t + 2) & (RU - 1), u0_ = (t + 1) & (RU - 1), um_ = t & (RU - 1);
        // red-result slots of planes s, s - 1, s - 2, s - 3
        constexpr int r0_ = (t + 1) & (RR - 1), rm_ = t & (RR - 1), rmm_ = (t + 3) & (RR - 1),
                      rmmm_ = (t + 2) & (RR - 1);
        // D slots of planes s + 1, s, s - 1, s - 2, s - 3
        constexpr int dp_ = (t + 2) & (RD - 1), d0_ = (t + 1) & (RD - 1), dm_ = t & (RD - 1),
                      dmm_ = (t + 7) & (RD - 1), dmmm_ = (t + 6) & (RD - 1);
        // the black pass's plane k = s - 1 (OB: s - 2): its red neighbours'
        // planes k - 1, k, k + 1, its D planes, and its own u / reac slot
        constexpr int bq_m = OB ? rmmm_ : rmm_, bq_0 = OB ? rmm_ : rm_, bq_p = OB ? rm_ : r0_;
        constexpr int bd_m = OB ? dmmm_ : dmm_, bd_0 = OB ? dmm_ : dm_, bd_p = OB ? dm_ : d0_;
        constexpr int bu_ = OB ? (t + 7) & (RU - 1) : um_;
        if constexpr (DL) {
            wait_vmcnt<KREG>();  // this wave's stage of plane s + 1 has landed
        } else {
            lstore(up_, FCS ? up_ : dp_, up_, SlotC<x>{});
        }
        const T br = x ? nbr1 : nbr0;
        const T bk = x ? nbb1 : nbb0;
        // z face factors of the red plane s and the black plane s - 1
        // (uniform: formed outside the lanes' branches; with ZGHOST the
        // interior ones on every plane)
        T rfzm = L.hz, rfzp = L.hz, rgz = T(0), bfzm = L.hz, bfzp = L.hz, bgz = T(0);
        if constexpr (!ZGHOST) {
            zf(s, rfzm, rfzp, rgz);
            zf(s - 1 - (OB ? 1 : 0), bfzm, bfzp, bgz);
        }
        __syncthreads();
        // the prefetch is issued on every step (past the last plane this
        // block needs it re-reads the last planes, L2 hits: pofs clamps them
        // to the slab), so the loop body is straight-line and the compiler
        // keeps the register sets in place. With OB the last step's black
        // plane ke takes its b from gload(ke + 3).
        // DL: plane s + 2 goes to u slot (t + 3) & 3 and D slot (t + 3) & 7
        // (those of planes s - 2 and s - 6, whose last readers, the passes
        // of step s - 1, are behind the barrier above); none is needed past
        // plane ke + 2, so the last step stages nothing and no LDS load is
        // left in flight when the block ends
        if constexpr (DL) {
            if (s + 2 <= ke + 2) stage(s + 2, SlotC<(t + 3) & (RU - 1)>{}, SlotC<(t + 3) & (RD - 1)>{});
            __builtin_amdgcn_sched_barrier(0);
        }
        gload(min(s + 2, ke + 2 + (OB ? 1 : 0)), SlotC<1 - x>{});
        // red cells of plane s; red cells outside the interior (or planes
        // outside 1..nz) hold 0 in the red ring, so the black pass reads its
        // neighbours without masking them (their face factors are zero)
        if (red_lane) {
            const Cell &C = x ? R1 : R0;
            T val = T(0);
            if (s >= 1 && s <= nz && C.o >= 0) {
                const int e = C.e;
                const T *um = su[um_], *u0 = su[u0_], *up = su[up_];
                if constexpr (FCS) {
                    const T az = (s == 1 || s == nz) ? L.azb : L.az;
                    const T *gr = x ? rs1 : rs0, *Gx = sD[u0_], *Gy = sGy[u0_];
                    const T g[ST_N] = {gr[ST_C], Gx[e - 1], Gx[e], Gy[e - W2], Gy[e], gr[ST_ZM], gr[ST_ZP]};
                    val = faces_row(g, C.ax, C.ay, az, u0[e - 1], u0[e + 1], u0[e - W2], u0[e + W2], um[e], up[e], br);
                } else if constexpr (STN) {
                    val = stencil_row(x ? rs1 : rs0, u0[e - 1], u0[e + 1], u0[e - W2], u0[e + W2], um[e], up[e], br);
                } else {
                    const T *Dm = sD[dm_], *D0 = sD[d0_], *Dp = sD[dp_];
                    const T rc = REAC ? sq[REAC ? u0_ : 0][C.q] : T(0);
                    val = fast_row(D0[e], D0[e - 1], D0[e + 1], D0[e - W2], D0[e + W2], Dm[e], Dp[e], u0[e - 1],
                                   u0[e + 1], u0[e - W2], u0[e + W2], um[e], up[e], C.fxm, C.fxp, rfym, rfyp, rfzm,
                                   rfzp, ZGHOST ? C.g : C.g + rgz, rc, br);
                }
                if (L.omega != T(1)) val = u0[e] + L.omega * (val - u0[e]);
                if (s >= kb && s <= ke) buf_store<T>(val, Ro, C.vs, pofs(s));
            }
            sr[r0_][C.q] = val;
        }
        // DL: the barrier after the red pass only publishes the red results'
        // LDS stores; __syncthreads' fence would also wait for the stage of
        // plane s + 2 in flight, which no wave reads before the next step's
        // barrier. The memory clobber keeps LDS accesses on their side.
        if constexpr (DL) {
            asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
        } else if constexpr (!OB) {
            __syncthreads();
        }
        const int k = s - 1 - (OB ? 1 : 0);
        const Cell &Bc = x ? B1 : B0;
        if (k >= kb && k <= ke && Bc.o >= 0) {
            const int q = Bc.q, e = Bc.e;
            // red planes k-1, k, k+1
            const T *rm = sr[bq_m], *r0 = sr[bq_0], *rp = sr[bq_p];
            T v;
            if constexpr (FCS) {
                const T az = (k == 1 || k == nz) ? L.azb : L.az;
                const T *gb = x ? bs1 : bs0, *Gx = sD[bu_], *Gy = sGy[bu_];
                const T g[ST_N] = {gb[ST_C], Gx[e - 1], Gx[e], Gy[e - W2], Gy[e], gb[ST_ZM], gb[ST_ZP]};
                v = faces_row(g, Bc.ax, Bc.ay, az, r0[q - 1], r0[q + 1], r0[q - W1], r0[q + W1], rm[q], rp[q], bk);
            } else if constexpr (STN) {
                v = stencil_row(x ? bs1 : bs0, r0[q - 1], r0[q + 1], r0[q - W1], r0[q + W1], rm[q], rp[q], bk);
            } else {
                const T *Dm = sD[bd_m], *D0 = sD[bd_0], *Dp = sD[bd_p];
                const T rc = REAC ? sq[REAC ? bu_ : 0][q] : T(0);
                v = fast_row(D0[e], D0[e - 1], D0[e + 1], D0[e - W2], D0[e + W2], Dm[e], Dp[e], r0[q - 1], r0[q + 1],
                             r0[q - W1], r0[q + W1], rm[q], rp[q], Bc.fxm, Bc.fxp, bfym, bfyp, bfzm, bfzp,
                             ZGHOST ? Bc.g : Bc.g + bgz, rc, bk);
            }
            if (L.omega != T(1)) v = su[bu_][e] + L.omega * (v - su[bu_][e]);
            buf_store<T>(v, Ro, Bc.vs, pofs(k));
        }
    };
    // (OB: one step more, for the last black plane)
    const int se = ke + 1 + (OB ? 1 : 0);
    for (int s = kb - 1; s <= se; s += 8) {
        step(s, SlotC<0>{});
        if (s + 1 <= se) step(s + 1, SlotC<1>{});
        if (s + 2 <= se) step(s + 2, SlotC<2>{});
        if (s + 3 <= se) step(s + 3, SlotC<3>{});
        if (s + 4 <= se) step(s + 4, SlotC<4>{});
        if (s + 5 <= se) step(s + 5, SlotC<5>{});
        if (s + 6 <= se) step(s + 6, SlotC<6>{});
        if (s + 7 <= se) step(s + 7, SlotC<7>{});
    }
}

}  // namespace mps

namespace mps {

// Residual r = b - A u of rows without convection in fast arithmetic (and
// the block's sum of r^2), marching the z-planes of a TX x (TY*RB) tile:
// each lane owns RB cells of its x column (rows ty, ty+TY, ...), keeps their
// z neighbours in registers and reads x/y neighbours from a
// (TX+2) x (TY*RB+2) plane in LDS (the halo ring, corners excluded, is
// loaded by the first lanes). The next plane's loads are in flight while
// the current plane is evaluated; two LDS planes alternate, so each plane
// costs one barrier, and RB cells per lane put RB times the loads in flight
// per barrier. Values outside the interior are staged as zeros and their
// face factors are zero. R is the stored type of r (float for the FP32
// defect of the FP64 iterate, stored scaled by rscale[0], a power of two;
// rscale nullable = 1); r is nullable.
template <typename T, typename R, int OP, bool REAC, int TX, int TY, int RB>
__global__ void __launch_bounds__(TX *TY) k_residual_planes(Lev<T> L, const T *__restrict__ u,
                                                           const T *__restrict__ b, R *__restrict__ r,
                                                           double *partials, int kz, const double *rscale)
{
#pragma clang fp contract(fast)
    constexpr int HY = TY * RB;
    constexpr int W = TX + 2, H = HY + 2, NT = TX * TY;
    constexpr int NH = 2 * (TX + HY);
    constexpr int NHL = (NH + NT - 1) / NT;  // halo cells per lane
    static_assert(NT <= kBlock, "block_sum holds kBlock / 64 waves");
    __shared__ T su[2][W * H];
    __shared__ T sD[2][W * H];

    const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
    int bx, by, bz;
    xcd_tile(bx, by, bz);
    const int i0 = bx * TX + 1, j0 = by * HY + 1;
    const int kb = bz * kz + 1, ke = min(kb + kz - 1, L.nz);
    const int nx = L.nx, ny = L.ny, nz = L.nz, sy = L.sy;
    const long long sz = L.sz;
    const int i = i0 + tx;
    T rs = T(1);
    if constexpr (sizeof(R) < sizeof(T)) {
        if (rscale) rs = (T)rscale[0];
    }

    int oc[RB], q[RB];
    bool in[RB];
    T fxm[RB], fxp[RB], fym[RB], fyp[RB], gxy[RB];
#pragma unroll
    for (int m = 0; m < RB; ++m) {
        const int yy = ty + m * TY, j = j0 + yy;
        in[m] = i <= nx && j <= ny;
        oc[m] = in[m] ? i + sy * j : -1;
        q[m] = (tx + 1) + W * (yy + 1);
        const bool ex = OP == OP_ALIGNED && (i == 1 || i == nx);
        const bool ey = OP == OP_ALIGNED && (j == 1 || j == ny);
        const T hx = ex ? L.hxb : L.hx, hy = ey ? L.hyb : L.hy;
        fxm[m] = i > 1 ? hx : T(0);
        fxp[m] = i < nx ? hx : T(0);
        fym[m] = j > 1 ? hy : T(0);
        fyp[m] = j < ny ? hy : T(0);
        const T gx = ex ? L.gxb : L.gx, gy = ey ? L.gyb : L.gy;
        T g = T(0);
        if (i == 1 && L.bc[LEFT] == BC_DIRC) g += gx;
        if (i == nx && L.bc[RIGHT] == BC_DIRC) g += gx;
        if (j == 1 && L.bc[BOTTOM] == BC_DIRC) g += gy;
        if (j == ny && L.bc[TOP] == BC_DIRC) g += gy;
        gxy[m] = g;
    }
    int oh[NHL], qh[NHL];
#pragma unroll
    for (int h = 0; h < NHL; ++h) {
        const int t = tid + h * NT;
        oh[h] = -1;
        qh[h] = -1;
        if (t < NH) {
            int hi, hj;
            if (t < TX) {
                hi = t, hj = -1;
            } else if (t < 2 * TX) {
                hi = t - TX, hj = HY;
            } else if (t < 2 * TX + HY) {
                hi = -1, hj = t - 2 * TX;
            } else {
                hi = TX, hj = t - 2 * TX - HY;
            }
            const int gi = i0 + hi, gj = j0 + hj;
            qh[h] = (hi + 1) + W * (hj + 1);
            if (gi >= 1 && gi <= nx && gj >= 1 && gj <= ny) oh[h] = gi + sy * gj;
        }
    }

    T pu[RB], pD[RB], pb[RB], pq[RB], phu[NHL], phD[NHL];
    auto gload = [&](int k) {
        const bool ok = k >= 1 && k <= nz;
        const long long base = sz * k;
#pragma unroll
        for (int m = 0; m < RB; ++m) {
            pu[m] = pD[m] = pb[m] = pq[m] = T(0);
            if (ok && oc[m] >= 0) {
                pu[m] = u[base + oc[m]];
                pD[m] = L.D[base + oc[m]];
                pb[m] = b[base + oc[m]];
                if (REAC) pq[m] = L.reac[base + oc[m]];
            }
        }
#pragma unroll
        for (int h = 0; h < NHL; ++h) {
            phu[h] = phD[h] = T(0);
            if (ok && oh[h] >= 0) {
                phu[h] = u[base + oh[h]];
                phD[h] = L.D[base + oh[h]];
            }
        }
    };

    T um[RB], Dm[RB], uc[RB], Dc[RB], bc[RB], qc[RB], hu[NHL], hD[NHL];
    gload(kb - 1);
#pragma unroll
    for (int m = 0; m < RB; ++m) um[m] = pu[m], Dm[m] = pD[m];
    gload(kb);
#pragma unroll
    for (int m = 0; m < RB; ++m) uc[m] = pu[m], Dc[m] = pD[m], bc[m] = pb[m], qc[m] = pq[m];
#pragma unroll
    for (int h = 0; h < NHL; ++h) hu[h] = phu[h], hD[h] = phD[h];
    gload(kb + 1);
    double sq = 0.0;
    for (int k = kb; k <= ke; ++k) {
        const int s = k & 1;
#pragma unroll
        for (int m = 0; m < RB; ++m) {
            su[s][q[m]] = uc[m];
            sD[s][q[m]] = Dc[m];
        }
#pragma unroll
        for (int h = 0; h < NHL; ++h) {
            if (qh[h] >= 0) {
                su[s][qh[h]] = hu[h];
                sD[s][qh[h]] = hD[h];
            }
        }
        __syncthreads();
        T up[RB], Dp[RB], bn[RB], qn[RB], hun[NHL], hDn[NHL];
#pragma unroll
        for (int m = 0; m < RB; ++m) up[m] = pu[m], Dp[m] = pD[m], bn[m] = pb[m], qn[m] = pq[m];
#pragma unroll
        for (int h = 0; h < NHL; ++h) hun[h] = phu[h], hDn[h] = phD[h];
        if (k < ke) gload(k + 2);
        const bool ez = OP == OP_ALIGNED && (k == 1 || k == nz);
        const T hz = ez ? L.hzb : L.hz;
        const T fzm = k > 1 ? hz : T(0), fzp = k < nz ? hz : T(0);
        T gz = T(0);
        if (k == 1 && L.bc[BACK] == BC_DIRC) gz += ez ? L.gzb : L.gz;
        if (k == nz && L.bc[FRONT] == BC_DIRC) gz += ez ? L.gzb : L.gz;
        const T *U = su[s], *Dl = sD[s];
#pragma unroll
        for (int m = 0; m < RB; ++m) {
            if (in[m]) {
                const int qq = q[m];
                const T D0 = Dc[m];
                const T wxm = (Dl[qq - 1] + D0) * fxm[m], wxp = (D0 + Dl[qq + 1]) * fxp[m];
                const T wym = (Dl[qq - W] + D0) * fym[m], wyp = (D0 + Dl[qq + W]) * fyp[m];
                const T wzm = (Dm[m] + D0) * fzm, wzp = (D0 + Dp[m]) * fzp;
                T d = ((wxm + wxp) + (wym + wyp)) + (wzm + wzp);
                d = d + D0 * (gxy[m] + gz) + L.dtinv;
                if (REAC) d = d - qc[m];
                T nb = wxm * U[qq - 1];
                nb = nb + wxp * U[qq + 1];
                nb = nb + wym * U[qq - W];
                nb = nb + wyp * U[qq + W];
                nb = nb + wzm * um[m];
                nb = nb + wzp * up[m];
                const T rv = bc[m] - (d * uc[m] - nb);
                if constexpr (sizeof(R) < sizeof(T)) {
                    if (r) r[sz * k + oc[m]] = (R)(rv * rs);
                } else {
                    if (r) r[sz * k + oc[m]] = (R)rv;
                }
                sq += (double)rv * (double)rv;
            }
        }
#pragma unroll
        for (int m = 0; m < RB; ++m) {
            um[m] = uc[m], Dm[m] = Dc[m];
            uc[m] = up[m], Dc[m] = Dp[m], bc[m] = bn[m], qc[m] = qn[m];
        }
#pragma unroll
        for (int h = 0; h < NHL; ++h) hu[h] = hun[h], hD[h] = hDn[h];
    }
    if (partials) {
        const double t = block_sum(sq);
        if (tid == 0) partials[blockIdx.x + gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z)] = t;
    }
}

}  // namespace mps

namespace mps {

// Volume-weighted full weighting onto the aligned coarse grid, marching
// coarse z-planes: a block owns CX x CY coarse columns and stages, per
// coarse plane K, the fine planes 2K-1 and 2K of its (2CX+1) x (2CY+1) fine
// footprint in LDS; each lane forms the xy-weighted 3x3 sum of its coarse
// cell on each fine plane (FP64 accumulation) and combines three planes
// with the z weights, reusing plane 2K's sum as the next cell's plane 2K-2.
// Fine values outside the interior are staged as zeros (their weights are
// zero). The next step's fine planes are loaded while the current one is
// summed.
template <typename T, int CX, int CY>
__global__ void __launch_bounds__(CX *CY) k_restrict_fw_planes(Lev<T> F, Lev<T> C, AlignedW W,
                                                              const T *__restrict__ f, T *__restrict__ cf, int kz)
{
    constexpr int FW = 2 * CX + 1, FH = 2 * CY + 1, FP = FW * FH, NT = CX * CY;
    constexpr int NS = (FP + NT - 1) / NT;
    __shared__ T sf[2][FP];
    const int tid = threadIdx.x, tx = tid % CX, ty = tid / CX;
    int bx, by, bz;
    xcd_tile(bx, by, bz);
    const int I0 = bx * CX + 1, J0 = by * CY + 1;
    const int Kb = bz * kz + 1, Ke = min(Kb + kz - 1, C.nz);
    const int I = I0 + tx, J = J0 + ty;
    const bool in = I <= C.nx && J <= C.ny;
    int of[NS];
#pragma unroll
    for (int r = 0; r < NS; ++r) {
        const int q = tid + r * NT;
        of[r] = -1;
        if (q < FP) {
            const int gi = 2 * I0 - 2 + q % FW, gj = 2 * J0 - 2 + q / FW;
            if (gi >= 1 && gi <= F.nx && gj >= 1 && gj <= F.ny) of[r] = gi + F.sy * gj;
        }
    }
    double wx0 = 0, wx1 = 0, wx2 = 0, wy0 = 0, wy1 = 0, wy2 = 0;
    if (in) {
        wx0 = W.wx[3 * (I - 1)], wx1 = W.wx[3 * (I - 1) + 1], wx2 = W.wx[3 * (I - 1) + 2];
        wy0 = W.wy[3 * (J - 1)], wy1 = W.wy[3 * (J - 1) + 1], wy2 = W.wy[3 * (J - 1) + 2];
    }
    const int qc = (2 * tx + 1) + FW * (2 * ty + 1);
    T pa[NS], pb[NS];
    auto gload = [&](int fk, T(&p)[NS]) {
        const bool ok = fk >= 1 && fk <= F.nz;
        const T *base = f + F.sz * fk;
#pragma unroll
        for (int r = 0; r < NS; ++r) p[r] = (ok && of[r] >= 0) ? base[of[r]] : T(0);
    };
    auto lstore = [&](int slot, const T(&p)[NS]) {
#pragma unroll
        for (int r = 0; r < NS; ++r) {
            const int q = tid + r * NT;
            if (q < FP) sf[slot][q] = p[r];
        }
    };
    auto xysum = [&](int slot) {
        const T *s = sf[slot];
        const double r0 = (wx0 * (double)s[qc - FW - 1] + wx1 * (double)s[qc - FW]) + wx2 * (double)s[qc - FW + 1];
        const double r1 = (wx0 * (double)s[qc - 1] + wx1 * (double)s[qc]) + wx2 * (double)s[qc + 1];
        const double r2 = (wx0 * (double)s[qc + FW - 1] + wx1 * (double)s[qc + FW]) + wx2 * (double)s[qc + FW + 1];
        return (wy0 * r0 + wy1 * r1) + wy2 * r2;
    };

    gload(2 * Kb - 2, pa);
    lstore(0, pa);
    __syncthreads();
    double sprev = in ? xysum(0) : 0.0;
    gload(2 * Kb - 1, pa);
    gload(2 * Kb, pb);
    for (int K = Kb; K <= Ke; ++K) {
        __syncthreads();
        lstore(0, pa);
        lstore(1, pb);
        __syncthreads();
        if (K < Ke) {
            gload(2 * K + 1, pa);
            gload(2 * K + 2, pb);
        }
        if (in) {
            const double s1 = xysum(0), s2 = xysum(1);
            const double *wz = W.wz + 3 * (K - 1);
            cf[I + (long long)C.sy * J + C.sz * K] = (T)((wz[0] * sprev + wz[1] * s1) + wz[2] * s2);
            sprev = s2;
        }
    }
}

// Residual and full-weighting restriction in one pass (aligned hierarchy,
// rows without convection): the coarse right-hand side cf = R (b - A u) with
// the fine residual formed in LDS and never stored. A block owns a CX x CY
// tile of coarse cells and marches coarse planes K; step K forms the fine
// residual planes 2K-1 and 2K on the tile's (2CX+1) x (2CY+1) fine region
// from four u (and D) planes with a one-cell halo kept in a ring, then the
// x/y sums of the weights as k_restrict_fw_planes, carrying plane 2K's sum
// to step K+1. Rows are the fast face-factor rows of k_residual_planes, or
// explicit 7-point rows (OP_STENCIL, coefficients read per cell), with the
// same arithmetic. Fine residuals outside the interior are zero (their
// weights are zero). Global loads of step K+1 are issued before step K's
// residuals are formed. M lanes per coarse cell share the staging and the
// residuals (the first CX*CY lanes also form the coarse values).
template <typename T, int OP, bool REAC, int CX, int CY, int M>
__global__ void __launch_bounds__(CX *CY *M) k_resrestrict_planes(Lev<T> F, Lev<T> C, AlignedW W,
                                                                 const T *__restrict__ u, const T *__restrict__ b,
                                                                 T *__restrict__ cf, int kz)
{
#pragma clang fp contract(fast)
    constexpr bool STN = OP == OP_STENCIL;
    constexpr int FW = 2 * CX + 1, FH = 2 * CY + 1, FP = FW * FH;
    constexpr int UW = FW + 2, UH = FH + 2, UP = UW * UH;
    constexpr int NT = CX * CY * M;
    constexpr int NU = (UP + NT - 1) / NT;
    constexpr int NF = (FP + NT - 1) / NT;
    __shared__ T su[4][UP];
    __shared__ T sD[STN ? 1 : 4][STN ? 1 : UP];
    __shared__ T sr[2][FP];

    const int tid = threadIdx.x, tx = tid % CX, ty = tid / CX;
    int bx, by, bz;
    xcd_tile(bx, by, bz);
    const int I0 = bx * CX + 1, J0 = by * CY + 1;
    const int Kb = bz * kz + 1, Ke = min(Kb + kz - 1, C.nz);
    const int I = I0 + tx, J = J0 + ty;
    const bool in = tid < CX * CY && I <= C.nx && J <= C.ny;
    const int nx = F.nx, ny = F.ny, nz = F.nz, sy = F.sy;
    const int szi = (int)F.sz;

    // staging offsets of the u/D region (fine cells 2I0-3 .., 2J0-3 ..;
    // -1 outside the interior: staged as zeros)
    int ou[NU];
#pragma unroll
    for (int r = 0; r < NU; ++r) {
        const int q = tid + r * NT;
        ou[r] = -1;
        if (q < UP) {
            const int gi = 2 * I0 - 3 + q % UW, gj = 2 * J0 - 3 + q / UW;
            if (gi >= 1 && gi <= nx && gj >= 1 && gj <= ny) ou[r] = gi + sy * gj;
        }
    }
    // this lane's residual cells: fine region index tid + r NT, its place in
    // the u region, its field offset (-1: none) and x/y factors
    int of[NF], qu[NF];
    T fxm[NF], fxp[NF], fym[NF], fyp[NF], gxy[NF];
#pragma unroll
    for (int r = 0; r < NF; ++r) {
        const int q = tid + r * NT;
        of[r] = -1;
        qu[r] = UW + 1;
        fxm[r] = fxp[r] = fym[r] = fyp[r] = gxy[r] = T(0);
        if (q < FP) {
            const int fi = q % FW, fj = q / FW;
            const int gi = 2 * I0 - 2 + fi, gj = 2 * J0 - 2 + fj;
            qu[r] = (fi + 1) + UW * (fj + 1);
            if (gi >= 1 && gi <= nx && gj >= 1 && gj <= ny) {
                of[r] = gi + sy * gj;
                if constexpr (!STN) {
                    const bool ex = OP == OP_ALIGNED && (gi == 1 || gi == nx);
                    const bool ey = OP == OP_ALIGNED && (gj == 1 || gj == ny);
                    const T hx = ex ? F.hxb : F.hx, hy = ey ? F.hyb : F.hy;
                    fxm[r] = gi > 1 ? hx : T(0);
                    fxp[r] = gi < nx ? hx : T(0);
                    fym[r] = gj > 1 ? hy : T(0);
                    fyp[r] = gj < ny ? hy : T(0);
                    const T gx = ex ? F.gxb : F.gx, gy = ey ? F.gyb : F.gy;
                    T g = T(0);
                    if (gi == 1 && F.bc[LEFT] == BC_DIRC) g += gx;
                    if (gi == nx && F.bc[RIGHT] == BC_DIRC) g += gx;
                    if (gj == 1 && F.bc[BOTTOM] == BC_DIRC) g += gy;
                    if (gj == ny && F.bc[TOP] == BC_DIRC) g += gy;
                    gxy[r] = g;
                }
            }
        }
    }
    double wx0 = 0, wx1 = 0, wx2 = 0, wy0 = 0, wy1 = 0, wy2 = 0;
    if (in) {
        wx0 = W.wx[3 * (I - 1)], wx1 = W.wx[3 * (I - 1) + 1], wx2 = W.wx[3 * (I - 1) + 2];
        wy0 = W.wy[3 * (J - 1)], wy1 = W.wy[3 * (J - 1) + 1], wy2 = W.wy[3 * (J - 1) + 2];
    }
    const int qc = (2 * tx + 1) + FW * (2 * ty + 1);

    auto plane = [&](int p) { return min(max(p, 0), nz + 1) * szi; };
    auto slot = [](int p) { return (p + 4) & 3; };
    // register sets: u/D of plane p and b of plane p - 1 (the residual plane
    // that plane p completes); reac of plane p - 1 likewise
    T pu[2][NU], pD[2][STN ? 1 : NU], pb[2][NF], pq[2][REAC && !STN ? NF : 1];
    bool okp[2];
    auto gload = [&](int p, int x) {
        const int pk = plane(p), pkb = plane(p - 1);
        okp[x] = p >= 1 && p <= nz;
#pragma unroll
        for (int r = 0; r < NU; ++r) {
            const int o = pk + (ou[r] >= 0 ? ou[r] : 0);
            pu[x][r] = u[o];
            if constexpr (!STN) pD[x][r] = F.D[o];
        }
#pragma unroll
        for (int r = 0; r < NF; ++r) {
            const int o = pkb + (of[r] >= 0 ? of[r] : 0);
            pb[x][r] = b[o];
            if constexpr (REAC && !STN) pq[x][r] = F.reac[o];
        }
    };
    auto lstore = [&](int p, int x) {
        const int sl = slot(p);
#pragma unroll
        for (int r = 0; r < NU; ++r) {
            const int q = tid + r * NT;
            if (r * NT + NT <= UP || q < UP) {
                const bool on = okp[x] && ou[r] >= 0;
                su[sl][q] = on ? pu[x][r] : T(0);
                if constexpr (!STN) sD[sl][q] = on ? pD[x][r] : T(0);
            }
        }
    };
    // residual plane p into sr[dst], with b (and reac) of plane p from set x
    auto residual = [&](int p, int dst, const T(&bp)[NF], const T(&qp)[REAC && !STN ? NF : 1]) {
        const bool pin = p >= 1 && p <= nz;
        const T *Um = su[slot(p - 1)], *U0 = su[slot(p)], *Up = su[slot(p + 1)];
        T fzm = T(0), fzp = T(0), gz = T(0);
        if constexpr (!STN) {
            const bool ez = OP == OP_ALIGNED && (p == 1 || p == nz);
            const T hz = ez ? F.hzb : F.hz;
            fzm = p > 1 ? hz : T(0);
            fzp = p < nz ? hz : T(0);
            if (p == 1 && F.bc[BACK] == BC_DIRC) gz += ez ? F.gzb : F.gz;
            if (p == nz && F.bc[FRONT] == BC_DIRC) gz += ez ? F.gzb : F.gz;
        }
#pragma unroll
        for (int r = 0; r < NF; ++r) {
            const int qf = tid + r * NT;
            if (r * NT + NT <= FP || qf < FP) {
                T rv = T(0);
                if (pin && of[r] >= 0) {
                    const int q = qu[r];
                    if constexpr (STN) {
                        const long long c = (long long)plane(p) + of[r];
                        T o = T(0);
                        o = o + F.st[ST_XM][c] * U0[q - 1];
                        o = o + F.st[ST_XP][c] * U0[q + 1];
                        o = o + F.st[ST_YM][c] * U0[q - UW];
                        o = o + F.st[ST_YP][c] * U0[q + UW];
                        o = o + F.st[ST_ZM][c] * Um[q];
                        o = o + F.st[ST_ZP][c] * Up[q];
                        rv = bp[r] - (F.st[ST_C][c] * U0[q] + o);
                    } else {
                        const T *Dm = sD[slot(p - 1)], *D0 = sD[slot(p)], *Dp = sD[slot(p + 1)];
                        const T Dc = D0[q];
                        const T wxm = (D0[q - 1] + Dc) * fxm[r], wxp = (Dc + D0[q + 1]) * fxp[r];
                        const T wym = (D0[q - UW] + Dc) * fym[r], wyp = (Dc + D0[q + UW]) * fyp[r];
                        const T wzm = (Dm[q] + Dc) * fzm, wzp = (Dc + Dp[q]) * fzp;
                        T d = ((wxm + wxp) + (wym + wyp)) + (wzm + wzp);
                        d = d + Dc * (gxy[r] + gz) + F.dtinv;
                        if constexpr (REAC) d = d - qp[r];
                        T nb = wxm * U0[q - 1];
                        nb = nb + wxp * U0[q + 1];
                        nb = nb + wym * U0[q - UW];
                        nb = nb + wyp * U0[q + UW];
                        nb = nb + wzm * Um[q];
                        nb = nb + wzp * Up[q];
                        rv = bp[r] - (d * U0[q] - nb);
                    }
                }
                sr[dst][qf] = rv;
            }
        }
    };
    auto xysum = [&](int sl) {
        const T *s_ = sr[sl];
        const double r0 = (wx0 * (double)s_[qc - FW - 1] + wx1 * (double)s_[qc - FW]) + wx2 * (double)s_[qc - FW + 1];
        const double r1 = (wx0 * (double)s_[qc - 1] + wx1 * (double)s_[qc]) + wx2 * (double)s_[qc + 1];
        const double r2 = (wx0 * (double)s_[qc + FW - 1] + wx1 * (double)s_[qc + FW]) + wx2 * (double)s_[qc + FW + 1];
        return (wy0 * r0 + wy1 * r1) + wy2 * r2;
    };

    // warm-up: u/D planes 2Kb-3 .. 2Kb-1 and the residual of plane 2Kb-2
    const int f0 = 2 * Kb - 2;
    gload(f0 - 1, 0);
    lstore(f0 - 1, 0);
    gload(f0, 1);
    lstore(f0, 1);
    gload(f0 + 1, 0);
    lstore(f0 + 1, 0);
    T bw[NF], qw[REAC && !STN ? NF : 1];
#pragma unroll
    for (int r = 0; r < NF; ++r) bw[r] = pb[0][r];
    if constexpr (REAC && !STN) {
#pragma unroll
        for (int r = 0; r < NF; ++r) qw[r] = pq[0][r];
    }
    gload(f0 + 2, 0);
    gload(f0 + 3, 1);
    __syncthreads();
    residual(f0, 0, bw, qw);
    __syncthreads();
    double sprev = in ? xysum(0) : 0.0;
    for (int K = Kb; K <= Ke; ++K) {
        const int p1 = 2 * K - 1, p2 = 2 * K;
        lstore(p2, 0);
        lstore(p2 + 1, 1);
        T b1[NF], b2[NF], q1[REAC && !STN ? NF : 1], q2[REAC && !STN ? NF : 1];
#pragma unroll
        for (int r = 0; r < NF; ++r) b1[r] = pb[0][r], b2[r] = pb[1][r];
        if constexpr (REAC && !STN) {
#pragma unroll
            for (int r = 0; r < NF; ++r) q1[r] = pq[0][r], q2[r] = pq[1][r];
        }
        __syncthreads();
        if (K < Ke) {
            gload(p2 + 2, 0);
            gload(p2 + 3, 1);
        }
        residual(p1, 0, b1, q1);
        residual(p2, 1, b2, q2);
        __syncthreads();
        if (in) {
            const double s1 = xysum(0), s2 = xysum(1);
            const double *wz = W.wz + 3 * (K - 1);
            cf[I + (long long)C.sy * J + C.sz * K] = (T)((wz[0] * sprev + wz[1] * s1) + wz[2] * s2);
            sprev = s2;
        }
    }
}

// Trilinear prolongation + correction on the aligned hierarchy: each lane
// owns a fine column of its block's z-chunk (at most KZM planes, kz even,
// so the chunk starts on an odd fine plane). It interpolates the coarse
// grid in x and y once per coarse plane the chunk touches (fine cell f
// takes coarse (f+1)/2 when f is odd, the mean of f/2 and f/2+1 when even)
// and loads all of its fine values before the first store, so every load
// of the chunk is in flight at once; then fine plane k gets the coarse
// plane's value (k odd) or the mean of the two around it (k even). Same
// arithmetic as k_prolong_tri_correct.
template <typename T, int TX, int TY, int KZM>
__global__ void __launch_bounds__(TX *TY) k_prolong_tri_planes(Lev<T> F, Lev<T> C, T *__restrict__ uf,
                                                              const T *__restrict__ uc, int kz)
{
    // the block's coarse values first: for the tile's fine columns i0 ..
    // i0 + TX - 1 (i0 odd) the coarse columns I0 .. I0 + TX/2, rows likewise,
    // planes K0 .. K0 + KC - 1, staged in LDS with one coalesced load per
    // value (each is read by up to eight lanes); the interpolation below then
    // reads LDS with the arithmetic of the per-lane form
    static_assert(TX % 2 == 0 && TY % 2 == 0, "even tiles");
    constexpr int KC = KZM / 2 + 1;                  // coarse planes K0 .. K0 + KC - 1
    constexpr int CW = TX / 2 + 1, CH = TY / 2 + 1;  // coarse tile
    __shared__ T sc[KC * CH * CW];
    const int tid = threadIdx.x;
    int bx, by, bz;
    xcd_tile(bx, by, bz);
    const int i0 = bx * TX + 1, j0 = by * TY + 1;
    const int I0 = (i0 + 1) >> 1, J0 = (j0 + 1) >> 1;
    const int kb = bz * kz + 1, ke = min(kb + kz - 1, F.nz);
    const int K0 = (kb + 1) >> 1, nk = ke - kb + 1;
    for (int t = tid; t < KC * CH * CW; t += TX * TY) {
        const int q = t / (CH * CW), r = t - q * (CH * CW);
        const int J = J0 + r / CW, I = I0 + r % CW, K = K0 + q;
        // fine planes kb .. ke use coarse planes K0 .. (ke + 2) / 2 (zero
        // past the last coarse plane); columns and rows past the ghost layer
        // serve no lane
        const bool ok = 2 * q <= nk && K <= C.nz && I <= C.nx + 1 && J <= C.ny + 1;
        sc[t] = ok ? uc[I + (long long)C.sy * J + C.sz * K] : T(0);
    }
    __syncthreads();
    const int i = i0 + tid % TX, j = j0 + tid / TX;
    if (i > F.nx || j > F.ny) return;
    const int ia = (i + 1) >> 1, ja = (j + 1) >> 1;
    const int ib = (i & 1) ? ia : ia + 1, jb = (j & 1) ? ja : ja + 1;
    const bool xo = i & 1, yo = j & 1;
    const int oaa = (ia - I0) + CW * (ja - J0), oba = (ib - I0) + CW * (ja - J0);
    const int oab = (ia - I0) + CW * (jb - J0), obb = (ib - I0) + CW * (jb - J0);
    auto xy = [&](int q) {
        const T *p = sc + q * (CH * CW);
        if (xo) return yo ? p[oaa] : T(0.5) * (p[oaa] + p[oab]);
        const T r0 = T(0.5) * (p[oaa] + p[oba]);
        if (yo) return r0;
        const T r1 = T(0.5) * (p[oab] + p[obb]);
        return T(0.5) * (r0 + r1);
    };
    T ec[KC], fv[KZM];
    T *col = uf + i + (long long)F.sy * j + F.sz * kb;
#pragma unroll
    for (int q = 0; q < KC; ++q) ec[q] = (2 * q <= nk && K0 + q <= C.nz) ? xy(q) : T(0);
#pragma unroll
    for (int q = 0; q < KZM; ++q) fv[q] = q < nk ? col[F.sz * q] : T(0);
#pragma unroll
    for (int q = 0; q < KZM; ++q) {
        if (q >= nk) break;
        const T e = (q & 1) ? T(0.5) * (ec[q >> 1] + ec[(q >> 1) + 1]) : ec[q >> 1];
        col[F.sz * q] = fv[q] + e;
    }
}

}  // namespace mps
