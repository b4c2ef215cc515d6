// The smallest levels of an aligned-hierarchy V-cycle, resident in LDS.
//
// Below ~9^3 cells a level's kernels each cost a few microseconds of launch
// and dependency latency for almost no work, and every V-cycle visits each
// small level with about a dozen of them. k_lds_cycle takes the V-cycle from
// one such level down to the dense coarsest solve and back in a single
// workgroup: it loads that level's right-hand side and the 7-point rows of
// every level below once, keeps all solutions, right-hand sides and
// residuals in LDS, and writes the level's correction back. The phases --
// red-black smoothing, residual, full weighting, trilinear
// prolongation-correction -- use the same arithmetic as the per-level
// kernels for explicit-stencil rows (row() with OP_STENCIL, gs_update), so
// the V-cycle is the same; only where the data lives changes.
#pragma once

#include "mps3d_kernels.h"

namespace mps {

#ifndef MPS_LDS_BLOCK
#define MPS_LDS_BLOCK 1024
#endif
constexpr int kLdsMaxLev = 4;              // levels one launch can hold
constexpr int kLdsBlock = MPS_LDS_BLOCK;   // lanes of the workgroup

template <typename T>
struct LdsLevels {
    int nx[kLdsMaxLev], ny[kLdsMaxLev], nz[kLdsMaxLev];
    const T *st[kLdsMaxLev][ST_N];  // the levels' rows (ghosted layout, global)
    AlignedW W[kLdsMaxLev];         // W[q]: full-weighting weights onto level q
    const T *b0;                    // right-hand side of level 0 (ghosted, global)
    T *u0;                          // its correction (ghosted, global; written)
    long long sy0, sz0;             // strides of the level-0 global arrays
    const double *inv;              // dense inverse of the coarsest level
    int nlev, nu1, nu2, fast;
};

// LDS words (of T) the levels need: per level a ghosted u, an interior b and
// r, and 7 interior row arrays
__host__ __device__ inline long long lds_level_words(int nx, int ny, int nz)
{
    const long long n = (long long)nx * ny * nz;
    return (long long)(nx + 2) * (ny + 2) * (nz + 2) + 2 * n + ST_N * n;
}

template <typename T>
struct LdsLevelView {
    int nx, ny, nz, gy, gz;  // gy, gz: ghosted strides of u
    T *u, *b, *r, *st[ST_N];
};

template <typename T>
__device__ __forceinline__ T lds_gs(int fast, T b, T of, T dg)
{
    if constexpr (sizeof(T) == 4) {
        if (fast) return (b - of) * __builtin_amdgcn_rcpf(dg);
    }
    return (b - of) / dg;
}

template <typename T>
__device__ __forceinline__ void lds_rowsum(const LdsLevelView<T> &V, int n, int i, int j, int k, T &dg, T &of)
{
    const int g = i + V.gy * j + V.gz * k;
    T o = T(0);
    o = o + V.st[ST_XM][n] * V.u[g - 1];
    o = o + V.st[ST_XP][n] * V.u[g + 1];
    o = o + V.st[ST_YM][n] * V.u[g - V.gy];
    o = o + V.st[ST_YP][n] * V.u[g + V.gy];
    o = o + V.st[ST_ZM][n] * V.u[g - V.gz];
    o = o + V.st[ST_ZP][n] * V.u[g + V.gz];
    dg = V.st[ST_C][n];
    of = o;
}

template <typename T>
__device__ __forceinline__ void lds_ijk(const LdsLevelView<T> &V, int t, int &i, int &j, int &k)
{
    const int nxy = V.nx * V.ny;
    k = t / nxy;
    const int rem = t - k * nxy;
    j = rem / V.nx;
    i = rem - j * V.nx + 1;
    ++j;
    ++k;
}

// level q's arrays in the LDS block (levels packed one after the other);
// computed on demand from the kernel arguments, so no per-lane table of
// views is indexed dynamically
template <typename T>
__device__ __forceinline__ LdsLevelView<T> lds_view(const LdsLevels<T> &S, T *mem, int q)
{
    long long off = 0;
    for (int p = 0; p < q; ++p) off += lds_level_words(S.nx[p], S.ny[p], S.nz[p]);
    LdsLevelView<T> v;
    v.nx = S.nx[q], v.ny = S.ny[q], v.nz = S.nz[q];
    v.gy = v.nx + 2, v.gz = (v.nx + 2) * (v.ny + 2);
    const int n = v.nx * v.ny * v.nz;
    v.u = mem + off;
    off += (long long)v.gz * (v.nz + 2);
    v.b = mem + off;
    off += n;
    v.r = mem + off;
    off += n;
    for (int s = 0; s < ST_N; ++s) {
        v.st[s] = mem + off;
        off += n;
    }
    return v;
}

template <typename T>
__global__ void __launch_bounds__(kLdsBlock) k_lds_cycle(LdsLevels<T> S)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    T *mem = reinterpret_cast<T *>(lds_raw);
    const int tid = threadIdx.x, nt = blockDim.x;
    // load: zero solutions, rows of every level, level 0's right-hand side
    for (int q = 0; q < S.nlev; ++q) {
        const LdsLevelView<T> v = lds_view(S, mem, q);
        const int ng = v.gz * (v.nz + 2), n = v.nx * v.ny * v.nz;
        for (int t = tid; t < ng; t += nt) v.u[t] = T(0);
        const long long gsy = v.nx + 2, gsz = (long long)(v.nx + 2) * (v.ny + 2);
        for (int t = tid; t < n; t += nt) {
            int i, j, k;
            lds_ijk(v, t, i, j, k);
            const long long c = i + gsy * j + gsz * k;
            for (int s = 0; s < ST_N; ++s) v.st[s][t] = S.st[q][s][c];
            if (q == 0) v.b[t] = S.b0[i + S.sy0 * j + S.sz0 * k];
        }
    }
    __syncthreads();

    auto colour = [&](const LdsLevelView<T> &v, int red) {
        const int n = v.nx * v.ny * v.nz;
        for (int t = tid; t < n; t += nt) {
            int i, j, k;
            lds_ijk(v, t, i, j, k);
            if (((i + j + k) & 1) != red) continue;
            T dg, of;
            lds_rowsum(v, t, i, j, k, dg, of);
            v.u[i + v.gy * j + v.gz * k] = lds_gs<T>(S.fast, v.b[t], of, dg);
        }
    };
    auto smooth = [&](const LdsLevelView<T> &v, int nu) {
        for (int s = 0; s < nu; ++s) {
            colour(v, 1);
            __syncthreads();
            colour(v, 0);
            __syncthreads();
        }
    };

    const int last = S.nlev - 1;
    for (int q = 0; q < last; ++q) {
        const LdsLevelView<T> v = lds_view(S, mem, q), c = lds_view(S, mem, q + 1);
        smooth(v, S.nu1);
        const int n = v.nx * v.ny * v.nz;
        for (int t = tid; t < n; t += nt) {
            int i, j, k;
            lds_ijk(v, t, i, j, k);
            T dg, of;
            lds_rowsum(v, t, i, j, k, dg, of);
            v.r[t] = v.b[t] - (dg * v.u[i + v.gy * j + v.gz * k] + of);
        }
        __syncthreads();
        // full weighting onto level q + 1 (as k_restrict_fw; fine residual
        // values outside the interior carry zero weight)
        const AlignedW W = S.W[q + 1];
        const int nc = c.nx * c.ny * c.nz;
        for (int t = tid; t < nc; t += nt) {
            int I, J, K;
            lds_ijk(c, t, I, J, K);
            // the cell's nine weights first, all loads in flight together
            // (a zero weight marks a fine neighbour outside the interior)
            double wx[3], wy[3], wz[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                wx[d] = W.wx[3 * (I - 1) + d];
                wy[d] = W.wy[3 * (J - 1) + d];
                wz[d] = W.wz[3 * (K - 1) + d];
            }
            double acc = 0.0;
#pragma unroll
            for (int dk = 0; dk < 3; ++dk) {
                const int fk = 2 * K - 2 + dk;
#pragma unroll
                for (int dj = 0; dj < 3; ++dj) {
                    const int fj = 2 * J - 2 + dj;
                    double row_ = 0.0;
#pragma unroll
                    for (int di = 0; di < 3; ++di) {
                        const int fi = 2 * I - 2 + di;
                        if (wx[di] != 0.0 && wy[dj] != 0.0 && wz[dk] != 0.0)
                            row_ += wx[di] * (double)v.r[(fi - 1) + v.nx * ((fj - 1) + v.ny * (fk - 1))];
                    }
                    acc += (wz[dk] * wy[dj]) * row_;
                }
            }
            c.b[t] = (T)acc;
        }
        __syncthreads();
    }
    // coarsest level: u = A^-1 b with the dense inverse
    {
        const LdsLevelView<T> v = lds_view(S, mem, last);
        const int n = v.nx * v.ny * v.nz;
        for (int q = tid; q < n; q += nt) {
            const double *row_ = S.inv + (long long)q * n;
            double s = 0.0;
            for (int p = 0; p < n; ++p) s += row_[p] * (double)v.b[p];
            int i, j, k;
            lds_ijk(v, q, i, j, k);
            v.u[i + v.gy * j + v.gz * k] = (T)s;
        }
        __syncthreads();
    }
    for (int q = last - 1; q >= 0; --q) {
        const LdsLevelView<T> v = lds_view(S, mem, q), c = lds_view(S, mem, q + 1);
        // trilinear prolongation-correction (as k_prolong_tri_correct)
        const int n = v.nx * v.ny * v.nz;
        for (int t = tid; t < n; t += nt) {
            int i, j, k;
            lds_ijk(v, t, i, j, k);
            const int ia = (i + 1) >> 1, ja = (j + 1) >> 1, ka = (k + 1) >> 1;
            const int ib = (i & 1) ? ia : ia + 1, jb = (j & 1) ? ja : ja + 1, kb = (k & 1) ? ka : ka + 1;
            auto at = [&](int a, int b_, int d) { return c.u[a + c.gy * b_ + c.gz * d]; };
            auto xy = [&](int K) {
                if (i & 1) return (j & 1) ? at(ia, ja, K) : T(0.5) * (at(ia, ja, K) + at(ia, jb, K));
                const T r0 = T(0.5) * (at(ia, ja, K) + at(ib, ja, K));
                if (j & 1) return r0;
                return T(0.5) * (r0 + T(0.5) * (at(ia, jb, K) + at(ib, jb, K)));
            };
            const T e0 = xy(ka);
            const T e = (k & 1) ? e0 : T(0.5) * (e0 + xy(kb));
            v.u[i + v.gy * j + v.gz * k] += e;
        }
        __syncthreads();
        smooth(v, S.nu2);
    }
    // level 0's correction back to global memory
    {
        const LdsLevelView<T> v = lds_view(S, mem, 0);
        const int n = v.nx * v.ny * v.nz;
        for (int t = tid; t < n; t += nt) {
            int i, j, k;
            lds_ijk(v, t, i, j, k);
            S.u0[i + S.sy0 * j + S.sz0 * k] = v.u[i + v.gy * j + v.gz * k];
        }
    }
}

}  // namespace mps
