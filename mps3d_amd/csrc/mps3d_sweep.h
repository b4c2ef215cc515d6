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
    const int i0 = blockIdx.x * TX + 1, j0 = blockIdx.y * TY + 1;
    const int kb = blockIdx.z * kz + 1;
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

// Fused red-black sweep in fast arithmetic (rows without convection).
//
// Same schedule and LDS rings as k_sweep. Rows are branch-free: each lane
// precomputes, for its cells, the six face factors abv/(2d) with the faces
// on the physical boundary masked to zero and the DIRC wall factor, so a row
// is  w_f = (D_c + D_nb) h_f,  diag = sum w_f + D_c g + 1/dt - reac,
// u = (b + sum w_f u_nb) / diag.  Values outside the interior are staged as
// zeros, so masked faces multiply zeros. Global addressing uses 32-bit lane
// offsets against per-plane base pointers.
template <typename T, int OP, bool REAC, int TX, int TY>
__global__ void __launch_bounds__(TX *TY) k_sweep_fast(Lev<T> L, const T *__restrict__ uin, T *__restrict__ uout,
                                                      const T *__restrict__ b, int kz)
{
#pragma clang fp contract(fast)
    constexpr int W2 = TX + 4, H2 = TY + 4, P2 = W2 * H2;
    constexpr int W1 = TX + 2, H1 = TY + 2, P1 = W1 * H1;
    constexpr int NT = TX * TY;
    constexpr int N2 = (P2 + NT - 1) / NT;
    constexpr int N1 = (P1 + NT - 1) / NT;
    constexpr int RW = W1 / 2;
    constexpr int NR = RW * H1;
    static_assert(W1 % 2 == 0, "tile width must be even");
    static_assert(NR <= NT, "one red cell per lane");
    constexpr int DR = 5;
    __shared__ T su[4][P2];
    __shared__ T sD[DR][P2];
    __shared__ T sb[4][P1];
    __shared__ T sq[REAC ? 4 : 1][REAC ? P1 : 1];
    __shared__ T sr[4][P1];

    const int tid = threadIdx.x;
    const int i0 = blockIdx.x * TX + 1, j0 = blockIdx.y * TY + 1;
    const int kb = blockIdx.z * kz + 1;
    const int ke = min(kb + kz - 1, L.nz);
    const int sy = L.sy;
    const long long sz = L.sz;
    const int nx = L.nx, ny = L.ny, nz = L.nz;

    // staging offsets (interior positions only; -1 elsewhere)
    int o2[N2], o1[N1];
#pragma unroll
    for (int r = 0; r < N2; ++r) {
        const int q = tid + r * NT;
        o2[r] = -1;
        if (q < P2) {
            const int gi = i0 - 2 + q % W2, gj = j0 - 2 + q / W2;
            if (gi >= 1 && gi <= nx && gj >= 1 && gj <= ny) o2[r] = gi + sy * gj;
        }
    }
#pragma unroll
    for (int r = 0; r < N1; ++r) {
        const int q = tid + r * NT;
        o1[r] = -1;
        if (q < P1) {
            const int gi = i0 - 1 + q % W1, gj = j0 - 1 + q / W1;
            if (gi >= 1 && gi <= nx && gj >= 1 && gj <= ny) o1[r] = gi + sy * gj;
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

    // red lane: the red cell of E1 row jr for both plane parities
    const bool red_lane = tid < NR;
    const int jr = tid / RW, mr = tid - jr * RW;
    const int jrg = j0 - 1 + jr;
    int rq[2], re[2];
    bool rin[2];
    T rfxm[2], rfxp[2], rfym, rfyp, rgxy[2];
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        const int ii = 2 * mr + p, i = i0 - 1 + ii;
        rq[p] = ii + W1 * jr;
        re[p] = (ii + 1) + W2 * (jr + 1);
        rin[p] = red_lane && i >= 1 && i <= nx && jrg >= 1 && jrg <= ny;
        T a, c2, d2;
        xy_factors(i, jrg, rfxm[p], rfxp[p], a, c2, rgxy[p]);
        rfym = a;
        rfyp = c2;
        (void)d2;
    }
    // black lane: its tile cell
    const int tx = tid % TX, ty = tid / TX;
    const int bi = i0 + tx, bj = j0 + ty;
    const bool bin = bi <= nx && bj <= ny;
    const int q1c = (tx + 1) + W1 * (ty + 1);
    const int e2c = (tx + 2) + W2 * (ty + 2);
    T bfxm, bfxp, bfym, bfyp, bgxy;
    xy_factors(bi, bj, bfxm, bfxp, bfym, bfyp, bgxy);
    const int cb = bi + sy * bj;

    T pu[N2], pD[N2], pb[N1], pq[N1];
    auto gload = [&](int kp) {
        const bool ok = kp >= 1 && kp <= nz;
        const T *ub = uin + sz * kp;
        const T *Db = L.D + sz * kp;
        const T *bb = b + sz * kp;
        const T *qb = REAC ? L.reac + sz * kp : nullptr;
#pragma unroll
        for (int r = 0; r < N2; ++r) {
            pu[r] = T(0);
            pD[r] = T(0);
            if (ok && o2[r] >= 0) {
                pu[r] = ub[o2[r]];
                pD[r] = Db[o2[r]];
            }
        }
#pragma unroll
        for (int r = 0; r < N1; ++r) {
            pb[r] = T(0);
            pq[r] = T(0);
            if (ok && o1[r] >= 0) {
                pb[r] = bb[o1[r]];
                if (REAC) pq[r] = qb[o1[r]];
            }
        }
    };
    auto lstore = [&](int kp) {
        const int s4 = kp & 3, s5 = (kp + 5) % DR;
#pragma unroll
        for (int r = 0; r < N2; ++r) {
            const int q = tid + r * NT;
            if (q < P2) {
                su[s4][q] = pu[r];
                sD[s5][q] = pD[r];
            }
        }
#pragma unroll
        for (int r = 0; r < N1; ++r) {
            const int q = tid + r * NT;
            if (q < P1) {
                sb[s4][q] = pb[r];
                if (REAC) sq[s4][q] = pq[r];
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

    gload(kb - 2);
    lstore(kb - 2);
    gload(kb - 1);
    lstore(kb - 1);
    gload(kb);
    for (int s = kb - 1; s <= ke + 1; ++s) {
        lstore(s + 1);
        __syncthreads();
        if (s <= ke) gload(s + 2);
        if (s >= 1 && s <= nz) {
            const int p = (i0 + j0 + jr + s + 1) & 1;
            const bool in = p ? rin[1] : rin[0];
            if (in) {
                T fzm, fzp, gz;
                z_factors(s, fzm, fzp, gz);
                const int q = p ? rq[1] : rq[0];
                const int e = p ? re[1] : re[0];
                const T *um = su[(s - 1) & 3], *u0 = su[s & 3], *up = su[(s + 1) & 3];
                const T *Dm = sD[(s - 1 + DR) % DR], *D0 = sD[s % DR], *Dp = sD[(s + 1) % DR];
                const T rc = REAC ? sq[REAC ? (s & 3) : 0][q] : T(0);
                sr[s & 3][q] = fast_row(D0[e], D0[e - 1], D0[e + 1], D0[e - W2], D0[e + W2], Dm[e], Dp[e], u0[e - 1],
                                        u0[e + 1], u0[e - W2], u0[e + W2], um[e], up[e], p ? rfxm[1] : rfxm[0],
                                        p ? rfxp[1] : rfxp[0], rfym, rfyp, fzm, fzp, (p ? rgxy[1] : rgxy[0]) + gz, rc,
                                        sb[s & 3][q]);
            }
        }
        __syncthreads();
        const int k = s - 1;
        if (k >= kb && k <= ke && bin) {
            T val;
            if ((bi + bj + k) & 1) {
                val = sr[k & 3][q1c];
            } else {
                T fzm, fzp, gz;
                z_factors(k, fzm, fzp, gz);
                const T *rm = sr[(k - 1) & 3], *r0 = sr[k & 3], *rp = sr[(k + 1) & 3];
                const T *Dm = sD[(k - 1 + DR) % DR], *D0 = sD[k % DR], *Dp = sD[(k + 1) % DR];
                const T rc = REAC ? sq[REAC ? (k & 3) : 0][q1c] : T(0);
                // red neighbours outside the interior hold stale values:
                // their face factors are zero, so select them away
                const T rzm = k > 1 ? rm[q1c] : T(0), rzp = k < nz ? rp[q1c] : T(0);
                const T rxm = bi > 1 ? r0[q1c - 1] : T(0), rxp = bi < nx ? r0[q1c + 1] : T(0);
                const T rym = bj > 1 ? r0[q1c - W1] : T(0), ryp = bj < ny ? r0[q1c + W1] : T(0);
                val = fast_row(D0[e2c], D0[e2c - 1], D0[e2c + 1], D0[e2c - W2], D0[e2c + W2], Dm[e2c], Dp[e2c], rxm,
                               rxp, rym, ryp, rzm, rzp, bfxm, bfxp, bfym, bfyp, fzm, fzp, bgxy + gz, rc,
                               sb[k & 3][q1c]);
            }
            uout[cb + sz * k] = val;
        }
    }
}

}  // namespace mps
