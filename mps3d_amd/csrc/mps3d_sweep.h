// Fused red-black Gauss-Seidel sweep (gfx950).
//
// One launch performs a whole red-then-black sweep: the red update reads the
// old black values from `uin`, the black update reads the new red values,
// and both colours land in `uout` (ping-pong buffers, so no block ever reads
// a value another block is rewriting). The result is bit-identical to the
// two in-place colour passes of perform_gseidel_smoothing
// (mgridlib/discretepde.f90:202-228): every cell is evaluated by row7 with
// the same neighbour values.
//
// Blocking: a workgroup owns a TX x TY column of cells and marches up a
// chunk of KZ planes (2.5D). Plane s+1 of u and D (with a 2-cell xy halo) is
// prefetched into registers while plane s is computed, then staged in a
// 4-plane LDS ring. Red cells are computed on the tile plus a 1-cell ring
// (the ring is recomputed by the neighbouring tiles too), black cells on the
// tile one plane behind, so each sweep streams u, D and b once from HBM and
// writes u once: 4 arrays per sweep instead of 8 for two colour passes.
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
    constexpr int NL = (P2 + NT - 1) / NT;
    __shared__ T su[4][P2];
    __shared__ T sD[4][P2];
    __shared__ T sr[4][P1];
    const int tid = threadIdx.x;
    const int i0 = blockIdx.x * TX + 1, j0 = blockIdx.y * TY + 1;
    const int kb = blockIdx.z * kz + 1;
    const int ke = min(kb + kz - 1, L.nz);
    const long long sy = L.sy, sz = L.sz;
    const T *__restrict__ Dg = L.D;

    T pu[NL], pD[NL];
    auto gload = [&](int kp) {
#pragma unroll
        for (int r = 0; r < NL; ++r) {
            const int q = tid + r * NT;
            pu[r] = T(0);
            pD[r] = T(0);
            if (q < P2 && kp >= 0 && kp <= L.nz + 1) {
                const int ii = q % W2, jj = q / W2;
                const int gi = i0 - 2 + ii, gj = j0 - 2 + jj;
                if (gi >= 0 && gi <= L.nx + 1 && gj >= 0 && gj <= L.ny + 1) {
                    const long long c = (long long)gi + sy * gj + sz * kp;
                    pu[r] = uin[c];
                    pD[r] = Dg[c];
                }
            }
        }
    };
    auto lstore = [&](int kp) {
        const int slot = kp & 3;
#pragma unroll
        for (int r = 0; r < NL; ++r) {
            const int q = tid + r * NT;
            if (q < P2) {
                su[slot][q] = pu[r];
                sD[slot][q] = pD[r];
            }
        }
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
        // red cells of plane s on the tile and its 1-cell ring
        if (s >= 1 && s <= L.nz) {
            const T *um = su[(s - 1) & 3], *u0 = su[s & 3], *up = su[(s + 1) & 3];
            const T *Dm = sD[(s - 1) & 3], *D0 = sD[s & 3], *Dp = sD[(s + 1) & 3];
            T *rs = sr[s & 3];
            for (int q = tid; q < P1; q += NT) {
                const int ii = q % W1, jj = q / W1;
                const int i = i0 - 1 + ii, j = j0 - 1 + jj;
                if (i < 1 || i > L.nx || j < 1 || j > L.ny) continue;
                if (((i + j + s) & 1) == 0) continue;
                const int e = (ii + 1) + W2 * (jj + 1);
                const long long c = (long long)i + sy * j + sz * s;
                T dg, of;
                row7<T, OP, REAC>(L, i, j, s, u0[e - 1], u0[e + 1], u0[e - W2], u0[e + W2], um[e], up[e], D0[e],
                                  D0[e - 1], D0[e + 1], D0[e - W2], D0[e + W2], Dm[e], Dp[e],
                                  REAC ? L.reac[c] : T(0), dg, of);
                rs[q] = (b[c] - of) / dg;
            }
        }
        __syncthreads();
        // black cells of plane s-1 on the tile; plane s-1 is complete
        const int k = s - 1;
        if (k >= kb && k <= ke) {
            const int tx = tid % TX, ty = tid / TX;
            const int i = i0 + tx, j = j0 + ty;
            if (i <= L.nx && j <= L.ny) {
                const long long c = (long long)i + sy * j + sz * k;
                const int q1 = (tx + 1) + W1 * (ty + 1);
                T val;
                if ((i + j + k) & 1) {
                    val = sr[k & 3][q1];
                } else {
                    const int e = (tx + 2) + W2 * (ty + 2);
                    const T *rm = sr[(k - 1) & 3], *r0 = sr[k & 3], *rp = sr[(k + 1) & 3];
                    const T *Dm = sD[(k - 1) & 3], *D0 = sD[k & 3], *Dp = sD[(k + 1) & 3];
                    T dg, of;
                    row7<T, OP, REAC>(L, i, j, k, r0[q1 - 1], r0[q1 + 1], r0[q1 - W1], r0[q1 + W1], rm[q1],
                                      rp[q1], D0[e], D0[e - 1], D0[e + 1], D0[e - W2], D0[e + W2], Dm[e], Dp[e],
                                      REAC ? L.reac[c] : T(0), dg, of);
                    val = (b[c] - of) / dg;
                }
                uout[c] = val;
            }
        }
        __syncthreads();
    }
}

// Dense coarsest-level solve u = A^-1 b with a precomputed inverse (row
// major, n x n, FP64), one workgroup.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_bottom_dense(Lev<T> L, T *u, const T *__restrict__ b,
                                                        const double *__restrict__ Ainv)
{
    __shared__ double sb[kBottomMaxDense];
    const int nx = L.nx, ny = L.ny, n = L.nx * L.ny * L.nz;
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        const int i = q % nx + 1, j = (q / nx) % ny + 1, k = q / (nx * ny) + 1;
        sb[q] = (double)b[(long long)i + (long long)L.sy * j + L.sz * k];
    }
    __syncthreads();
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        const double *row_ = Ainv + (long long)q * n;
        double s = 0.0;
        for (int p = 0; p < n; ++p) s += row_[p] * sb[p];
        const int i = q % nx + 1, j = (q / nx) % ny + 1, k = q / (nx * ny) + 1;
        u[(long long)i + (long long)L.sy * j + L.sz * k] = (T)s;
    }
}

}  // namespace mps
