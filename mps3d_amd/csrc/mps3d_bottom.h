// Coarsest-level solve and Scharfetter-Gummel stencil assembly (gfx950).
#pragma once

#include "mps3d_kernels.h"

namespace mps {

// ------------------------------------------------------------------------
// Bottom solve: performgmres (mgridlib/gmres.f90:197-434) with find_AX and
// gseidel_precond (discretepde.f90:238-337, niter = 1) on the coarsest grid,
// as bottomsolve calls it (mgridsteps.f90:342-355): m = n, one restart,
// classical Gram-Schmidt, Givens rotations, lucky breakdown at 1e-25.
//
// One workgroup runs the whole solve. Vector updates are spread over the
// workgroup; every inner product is summed by one lane in the reference's
// memory order (its products formed in parallel), so the coarse solve
// reproduces the reference's rounding. The kd Gram-Schmidt products of
// Arnoldi step kd all take the unmodified w (classical Gram-Schmidt), so
// they run at once, one lane each, against w held in LDS. The coarsest grid
// holds at most kBottomMax cells.
// ------------------------------------------------------------------------
constexpr int kBottomMax = 512;
// Krylov dimensions up to this keep the Hessenberg matrix and the Givens
// arrays in LDS (the reference's coarse solves use m = n, which is 8-64 on
// the usual process grids); larger m keep them in the work buffer
constexpr int kBottomLdsM = 64;
// Coarse grids of up to this many cells with m <= kBottomLdsM also keep an
// interior-only copy of the Krylov basis in LDS (column stride n + 1), so
// the Gram-Schmidt products and the orthogonalised vector read no global
// memory; the ghosted basis stays in the work buffer for the operator.
constexpr int kBottomLdsN = 64;

template <typename T>
__host__ __device__ inline long long bottom_work_elems(int m, long long ng)
{
    return (long long)(m + 1) * ng + 8 * ng + (long long)(m + 1) * m + 4LL * (m + 1);
}

// performgmres with nrestarts restarts of GMRES(m); the coarse solve uses
// m = n, nrestarts = 1 and tol = ERRTOLR*0.0001 (mgridsteps.f90:305, :342-343).
// LM: m <= kBottomLdsM, the Hessenberg and Givens arrays in LDS. A compile-
// time choice, so they are addressed as LDS (a pointer chosen at run time
// between LDS and global memory is a generic one, and every access through it
// a flat instruction).
template <typename T, int OP, bool VEL, bool REAC, bool LM>
__global__ void __launch_bounds__(kBlock) k_bottom(Lev<T> L, T *u, const T *__restrict__ b, T *work, int m,
                                                  int nrestarts, T tol, int *status)
{
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int n = nx * ny * nz;
    const long long ng = (long long)(nx + 2) * (ny + 2) * (nz + 2);
    T *V = work;
    T *x = V + (long long)(m + 1) * ng;
    T *x0 = x + ng;
    T *r0 = x0 + ng;
    T *r = r0 + ng;
    T *av = r + ng;
    T *w = av + ng;
    T *wj = w + ng;
    T *mb = wj + ng;
    __shared__ T sH[(kBottomLdsM + 1) * kBottomLdsM];
    __shared__ T sg[4 * (kBottomLdsM + 1)];
    const bool lds = LM;
    T *H = LM ? sH : mb + ng;
    T *cs = LM ? sg : H + (long long)(m + 1) * m;
    T *sn = cs + (m + 1);
    T *be = sn + (m + 1);
    T *y = be + (m + 1);
    __shared__ T sh[2];
    __shared__ int cmap[kBottomMax];   // interior cell q (i fastest) -> ghosted index
    __shared__ T pr[kBottomMax];       // products of one inner product
    __shared__ T wsh[kBottomMax];      // w of the current Arnoldi step, by cell
    __shared__ T sV[(kBottomLdsM + 1) * (kBottomLdsN + 1)];
    const bool vlds = lds && n <= kBottomLdsN;
    const int ldv = n + 1;
    const int tid = threadIdx.x, nt = blockDim.x;

    auto cidx = [&](int q, int &i, int &j, int &k) {
        i = q % nx + 1;
        j = (q / nx) % ny + 1;
        k = q / (nx * ny) + 1;
        return (long long)i + (long long)L.sy * j + L.sz * k;
    };
    auto zero = [&](T *v) {
        for (long long q = tid; q < ng; q += nt) v[q] = T(0);
        __syncthreads();
    };
    auto apply = [&](T *out, const T *in) {  // out interior = A in
        for (int q = tid; q < n; q += nt) {
            int i, j, k;
            long long c = cidx(q, i, j, k);
            T dg, of;
            row<T, OP, VEL, REAC>(L, in, i, j, k, c, dg, of);
            out[c] = dg * in[c] + of;
        }
        __syncthreads();
    };
    // gseidel_precond, niter = 1 (discretepde.f90:238-337), from out = 0:
    // the interior of out is cleared, then the red cells take
    // (in - of) / dg with of formed from the cleared black neighbours (as in
    // the reference: a non-finite coefficient makes it NaN there too), then
    // the black cells from the new red ones. Invariant: every buffer that
    // precond writes (mb, r, w) has zero ghost entries, set once at the
    // kernel's start and never written afterwards; explicit (OP_STENCIL)
    // rows of boundary cells multiply those ghosts by their zero
    // coefficients.
    auto precond = [&](T *out, const T *in) {
        if (n == 1) {
            if (tid == 0) {
                int i, j, k;
                long long c = cidx(0, i, j, k);
                T dg, of;
                row<T, OP, VEL, REAC>(L, in, i, j, k, c, dg, of);
                out[c] = in[c] / dg;
            }
            __syncthreads();
            return;
        }
        for (int q = tid; q < n; q += nt) out[cmap[q]] = T(0);
        __syncthreads();
        for (int q = tid; q < n; q += nt) {   // red cells from the cleared black ones
            int i, j, k;
            long long c = cidx(q, i, j, k);
            if (((i + j + k) & 1) != 1) continue;
            T dg, of;
            row<T, OP, VEL, REAC>(L, out, i, j, k, c, dg, of);
            out[c] = (in[c] - of) / dg;
        }
        __syncthreads();
        for (int q = tid; q < n; q += nt) {   // black cells from the new red ones
            int i, j, k;
            long long c = cidx(q, i, j, k);
            if (((i + j + k) & 1) != 0) continue;
            T dg, of;
            row<T, OP, VEL, REAC>(L, out, i, j, k, c, dg, of);
            out[c] = (in[c] - of) / dg;
        }
        __syncthreads();
    };
    for (int q = tid; q < n; q += nt) {
        int i, j, k;
        cmap[q] = (int)cidx(q, i, j, k);
    }
    __syncthreads();
    auto dot = [&](const T *a, const T *c2) {  // products in parallel, one lane sums in order
        for (int q = tid; q < n; q += nt) pr[q] = a[cmap[q]] * c2[cmap[q]];
        __syncthreads();
        if (tid == 0) {
            T s = T(0);
#pragma unroll 8
            for (int q = 0; q < n; ++q) s = s + pr[q];
            sh[0] = s;
        }
        __syncthreads();
        T s = sh[0];
        __syncthreads();
        return s;
    };
#define HM(i_, j_) H[(long long)((i_) - 1) + (long long)(m + 1) * ((j_) - 1)]

    if (tid == 0) *status = 0;
    zero(x);
    zero(r0);
    zero(av);   // apply and precond write interiors only: the ghost entries
    zero(w);    // of these stay zero for the whole solve
    zero(r);
    zero(mb);
    for (int q = tid; q < n; q += nt) {
        int i, j, k;
        long long c = cidx(q, i, j, k);
        x[c] = u[c];
        r0[c] = b[c];
    }
    __syncthreads();
    for (int rs = 1; rs <= nrestarts; ++rs) {
        for (long long q = tid; q < (long long)(m + 1) * ng; q += nt) V[q] = T(0);
        for (long long q = tid; q < (long long)(m + 1) * m; q += nt) H[q] = T(0);
        for (int q = tid; q < m + 1; q += nt) {
            cs[q] = T(0);
            sn[q] = T(0);
            be[q] = T(0);
            y[q] = T(0);
        }
        __syncthreads();
        // Minvb from b with zero ghosts (r0 is reused as the ghosted b)
        zero(r0);
        for (int q = tid; q < n; q += nt) {
            int i, j, k;
            long long c = cidx(q, i, j, k);
            r0[c] = b[c];
        }
        __syncthreads();
        precond(mb, r0);
        const T bnorm = sqrt(dot(mb, mb));
        if (bnorm == T(0)) {
            zero(x);
            break;
        }
        for (long long q = tid; q < ng; q += nt) x0[q] = x[q];
        __syncthreads();
        apply(av, x0);
        zero(r0);
        for (int q = tid; q < n; q += nt) {
            int i, j, k;
            long long c = cidx(q, i, j, k);
            r0[c] = b[c] - av[c];
        }
        __syncthreads();
        precond(r, r0);
        const T beta = sqrt(dot(r, r));
        if (tid == 0) be[0] = beta;
        __syncthreads();
        if (fabs((double)beta) < (double)tol) break;
        for (long long q = tid; q < ng; q += nt) V[q] = r[q] / beta;
        if (vlds)
            for (int q = tid; q < n; q += nt) sV[q] = r[cmap[q]] / beta;
        __syncthreads();
        int optk = m;
        int lucky = 0;
        T residnorm = T(0);
        for (int kd = 1; kd <= m; ++kd) {
            const T *vj = V + ng * (long long)(kd - 1);
            apply(av, vj);
            precond(w, av);
            // The serial loops below are unrolled so that their loads issue
            // ahead of the dependent chain; the summation order is unchanged.
            // classical Gram-Schmidt: h_ii = <w, v_ii> for ii = 1..kd, one
            // lane each, then wj = w - h_1 v_1 - h_2 v_2 - ... in that order
            for (int q = tid; q < n; q += nt) wsh[q] = w[cmap[q]];
            __syncthreads();
            for (int ii = 1 + tid; ii <= kd; ii += nt) {
                T h = T(0);
                if (vlds) {
                    const T *vi = sV + ldv * (ii - 1);
#pragma unroll 8
                    for (int q = 0; q < n; ++q) h = h + wsh[q] * vi[q];
                } else {
                    const T *vi = V + ng * (long long)(ii - 1);
#pragma unroll 8
                    for (int q = 0; q < n; ++q) h = h + wsh[q] * vi[cmap[q]];
                }
                HM(ii, kd) = h;
            }
            __syncthreads();
            T hn;
            if (vlds) {
                // wj by cell over the interior (the ghost entries of every
                // basis vector are zero), kept in wsh, with the products of
                // <wj, wj> formed beside it and summed in order by one lane
                for (int q = tid; q < n; q += nt) {
                    T t = wsh[q];
#pragma unroll 8
                    for (int ii = 1; ii <= kd; ++ii) t = t - HM(ii, kd) * sV[ldv * (ii - 1) + q];
                    wsh[q] = t;
                    pr[q] = t * t;
                }
                __syncthreads();
                if (tid == 0) {
                    T s2 = T(0);
#pragma unroll 8
                    for (int q = 0; q < n; ++q) s2 = s2 + pr[q];
                    sh[0] = s2;
                }
                __syncthreads();
                hn = sqrt(sh[0]);
            } else {
                for (long long q = tid; q < ng; q += nt) {
                    T t = w[q];
                    for (int ii = 1; ii <= kd; ++ii) t = t - HM(ii, kd) * V[ng * (long long)(ii - 1) + q];
                    wj[q] = t;
                }
                __syncthreads();
                hn = sqrt(dot(wj, wj));
            }
            int nanflag = 0;
            if (hn > T(1e-25)) {
                T *vn = V + ng * (long long)kd;
                if (vlds) {
                    for (int q = tid; q < n; q += nt) {
                        const T v = wsh[q] / hn;
                        vn[cmap[q]] = v;
                        sV[ldv * kd + q] = v;
                    }
                } else {
                    for (long long q = tid; q < ng; q += nt) vn[q] = wj[q] / hn;
                }
            } else if (!(hn != hn) && !((hn / hn != T(1)) && hn != T(0))) {
                lucky = 1;
            } else {
                nanflag = 1;
            }
            __syncthreads();
            if (nanflag) {
                if (tid == 0) *status = -1;
                return;
            }
            if (tid == 0) {  // triangularize (gmres.f90:159-195)
                HM(kd + 1, kd) = hn;
                // the rotated entry ii+1 becomes the next step's entry ii:
                // carried in a register, the chain touches memory only to
                // load the untouched entries and store the finished ones
                T hc = HM(1, kd);
#pragma unroll 8
                for (int ii = 1; ii <= kd - 1; ++ii) {
                    const T hn1 = HM(ii + 1, kd), c = cs[ii - 1], s = sn[ii - 1];
                    const T t = c * hc + s * hn1;
                    hc = -s * hc + c * hn1;
                    HM(ii, kd) = t;
                }
                if (kd > 1) HM(kd, kd) = hc;
                T rr = HM(kd, kd), ss = HM(kd + 1, kd), c, s;
                if (rr == T(0)) {
                    c = T(0);
                    s = T(1);
                } else {
                    c = rr / sqrt(rr * rr + ss * ss);
                    s = ss / sqrt(rr * rr + ss * ss);
                }
                HM(kd, kd) = c * HM(kd, kd) + s * HM(kd + 1, kd);
                HM(kd + 1, kd) = T(0);
                cs[kd - 1] = c;
                sn[kd - 1] = s;
                be[kd] = -s * be[kd - 1];
                be[kd - 1] = c * be[kd - 1];
                sh[1] = fabs(be[kd]);
            }
            __syncthreads();
            residnorm = sh[1];
            __syncthreads();
            if (lucky) {
                optk = kd;
                break;
            }
        }
        if (tid == 0) {  // uppertrisolve (gmres.f90:124-142)
            for (int ii = optk; ii >= 1; --ii) {
                y[ii - 1] = be[ii - 1];
                for (int jj = ii + 1; jj <= optk; ++jj) y[ii - 1] = y[ii - 1] - HM(ii, jj) * y[jj - 1];
                y[ii - 1] = y[ii - 1] / HM(ii, ii);
            }
        }
        __syncthreads();
        for (long long q = tid; q < ng; q += nt) {
            T xv = x0[q];
            for (int jj = 1; jj <= optk; ++jj) xv = xv + y[jj - 1] * V[ng * (long long)(jj - 1) + q];
            x[q] = xv;
        }
        __syncthreads();
        if (residnorm / bnorm <= tol || lucky) break;
    }
    for (int q = tid; q < n; q += nt) {
        int i, j, k;
        long long c = cidx(q, i, j, k);
        u[c] = x[c];
    }
#undef HM
}

// ------------------------------------------------------------------------
// Scharfetter-Gummel operator rows
// ------------------------------------------------------------------------

// Bernoulli function B(x) = x/(e^x - 1), evaluated in FP64 with expm1 and a
// Taylor series near 0 (B = 1 - x/2 + x^2/12 - x^4/720 + ...) so that the
// small-drift limit stays exact.
__device__ __forceinline__ double bernoulli(double x)
{
    const double ax = fabs(x);
    if (ax < 1e-3) {
        const double x2 = x * x;
        return 1.0 - 0.5 * x + x2 * (1.0 / 12.0 - x2 * (1.0 / 720.0));
    }
    if (x > 700.0) return x * exp(-x);
    return x / expm1(x);
}

// SG weights of one face (wL multiplies the left value, wR the right) in
// the form of find_sg_exponential_flux (mps1d/source/convdiff.f90:10-64):
// c = mean velocity, D = harmonic mean with the reference's 1e-10 guard,
// pure upwind for vanishing D, central diffusion for vanishing Peclet.
__device__ __forceinline__ void sg_weights(double cL, double cR, double DL, double DR, double d, double &wL,
                                           double &wR)
{
    const double eps = 1e-10;
    const double ch = 0.5 * (cL + cR);
    const double Dh = 2.0 * DL * DR / (DL + DR + eps);
    if (fabs(Dh) < eps) {
        const double sgn = ch <= 0.0 ? -1.0 : 1.0;
        wL = cL * 0.5 * (1.0 + sgn);
        wR = cR * 0.5 * (1.0 - sgn);
        return;
    }
    const double Pe = ch * d / Dh;
    const double g = Dh / d;
    if (fabs(Pe) < eps) {
        wL = g;
        wR = -g;
        return;
    }
    wL = g * bernoulli(-Pe);
    wR = -g * bernoulli(Pe);
}

// Assembles the 7-point SG rows of one level: each interior face carries
// the flux F = wL u_L + wR u_R (positive along +axis); row L gains F/w_L,
// row R loses F/w_R. Physical faces keep the reference's boundary rows
// (DIRC diffusion to the wall plus upwind outflow, ZGRD outflow). Reaction
// and 1/dt sit on the diagonal. transpose != 0 stores A^T.
template <int OP>
__global__ void __launch_bounds__(kBlock) k_sg_stencil(Lev<double> L, double *st0, double *st1, double *st2,
                                                      double *st3, double *st4, double *st5, double *st6,
                                                      int transpose)
{
    int i, j, k;
    long long c;
    if (!cell_of(L, i, j, k, c)) return;
    double *st[ST_N] = {st0, st1, st2, st3, st4, st5, st6};
    double a[ST_N];
    for (int q = 0; q < ST_N; ++q) a[q] = 0.0;
    double d = 0.0;
    const double *D = L.D;
    const double Dc = D[c];
    const double *vel[3] = {L.vx, L.vy, L.vz};
    const long long s[3] = {1, L.sy, L.sz};
    const int idx[3] = {i, j, k};
    const int nn[3] = {L.nx, L.ny, L.nz};
    const double h[3] = {L.dx, L.dy, L.dz};
    const double ai[3] = {L.ax, L.ay, L.az};
    const double ab[3] = {L.axb, L.ayb, L.azb};
    const double i0[3] = {L.ix0, L.iy0, L.iz0};
    for (int dim = 0; dim < 3; ++dim) {
        const bool edge = idx[dim] == 1 || idx[dim] == nn[dim];
        const double abv = (OP == OP_ALIGNED && edge) ? ab[dim] : ai[dim];
        const double vc = vel[dim] ? vel[dim][c] : 0.0;
        if (idx[dim] > 1) {  // face (c-s | c): this cell is R
            const long long m = c - s[dim];
            double wL, wR;
            sg_weights(vel[dim] ? vel[dim][m] : 0.0, vc, D[m], Dc, h[dim], wL, wR);
            d = d - wR * abv;
            a[1 + 2 * dim] = -wL * abv;
        }
        if (idx[dim] < nn[dim]) {  // face (c | c+s): this cell is L
            const long long p = c + s[dim];
            double wL, wR;
            sg_weights(vc, vel[dim] ? vel[dim][p] : 0.0, Dc, D[p], h[dim], wL, wR);
            d = d + wL * abv;
            a[2 + 2 * dim] = wR * abv;
        }
        if (idx[dim] == 1) bc_row<double, OP, true>(L.bc[2 * dim], false, Dc, vc, h[dim], abv, i0[dim], d);
        if (idx[dim] == nn[dim]) bc_row<double, OP, true>(L.bc[2 * dim + 1], true, Dc, vc, h[dim], abv, i0[dim], d);
    }
    if (L.reac) d = d - L.reac[c];
    if (L.timederiv) d = d + 1.0 / L.dt;
    a[ST_C] = d;
    if (!transpose) {
        for (int q = 0; q < ST_N; ++q) st[q][c] = a[q];
        return;
    }
    st[ST_C][c] = a[ST_C];
    if (i > 1) st[ST_XP][c - 1] = a[ST_XM];
    if (i < L.nx) st[ST_XM][c + 1] = a[ST_XP];
    if (j > 1) st[ST_YP][c - L.sy] = a[ST_YM];
    if (j < L.ny) st[ST_YM][c + L.sy] = a[ST_YP];
    if (k > 1) st[ST_ZP][c - L.sz] = a[ST_ZM];
    if (k < L.nz) st[ST_ZM][c + L.sz] = a[ST_ZP];
}

// transposes an explicit stencil (st -> stT); stT must be zero-filled
template <typename T>
__global__ void __launch_bounds__(kBlock) k_transpose_stencil(Lev<T> L, const T *s0, const T *s1, const T *s2,
                                                             const T *s3, const T *s4, const T *s5, const T *s6,
                                                             T *t0, T *t1, T *t2, T *t3, T *t4, T *t5, T *t6)
{
    int i, j, k;
    long long c;
    if (!cell_of(L, i, j, k, c)) return;
    t0[c] = s0[c];
    if (i > 1) t2[c - 1] = s1[c];
    if (i < L.nx) t1[c + 1] = s2[c];
    if (j > 1) t4[c - L.sy] = s3[c];
    if (j < L.ny) t3[c + L.sy] = s4[c];
    if (k > 1) t6[c - L.sz] = s5[c];
    if (k < L.nz) t5[c + L.sz] = s6[c];
}

// ------------------------------------------------------------------------
// Dense inverse of the coarsest operator on the device, for the reference
// hierarchy in fast arithmetic: the reference's coarse GMRES (m = n, one
// restart, tol 1e-15, mgridsteps.f90:342-355) is an exact solve of the
// coarsest system up to rounding, which x = A^-1 b gives directly
// (k_bottom_dense, one small matrix-vector product per V-cycle).
//
// One workgroup assembles A (row-major, n x n) from the level's 7-point rows
// (k_build_stencil output) and inverts it in place by Gauss-Jordan
// elimination with partial pivoting (rows swapped during the sweep, the
// swaps undone as column swaps at the end); the matrix sits in LDS for
// n <= kDenseLdsN, else in the global work array. Ties in the pivot search
// go to the lowest row, so the result is deterministic. A zero pivot sets
// *status = -1 (the inverse is then meaningless).
// ------------------------------------------------------------------------
constexpr int kDenseLdsN = 64;
constexpr int kDenseThreads = 1024;

__global__ void __launch_bounds__(kDenseThreads) k_dense_inverse(Lev<double> L, const double *__restrict__ st0,
                                                                 const double *__restrict__ st1,
                                                                 const double *__restrict__ st2,
                                                                 const double *__restrict__ st3,
                                                                 const double *__restrict__ st4,
                                                                 const double *__restrict__ st5,
                                                                 const double *__restrict__ st6, double *gwork,
                                                                 double *Ainv, int *status)
{
    __shared__ double sM[kDenseLdsN * kDenseLdsN];
    __shared__ double col[kBottomMaxDense];
    __shared__ int perm[kBottomMaxDense];
    __shared__ double rv[kDenseThreads / 64];
    __shared__ int ri[kDenseThreads / 64];
    __shared__ int piv;
    const int nx = L.nx, ny = L.ny, n = L.nx * L.ny * L.nz;
    const int tid = threadIdx.x, nt = blockDim.x;
    const long long nn = (long long)n * n;
    // A into the matrix: called with the LDS array or the global work array
    // itself, so each call addresses one memory kind
    auto assemble = [&](double *M) {
        for (long long q = tid; q < nn; q += nt) M[q] = 0.0;
        __syncthreads();
        for (int r = tid; r < n; r += nt) {
            const int i = r % nx + 1, j = (r / nx) % ny + 1, k = r / (nx * ny) + 1;
            const long long c = (long long)i + (long long)L.sy * j + L.sz * k;
            double *row = M + (long long)r * n;
            row[r] = st0[c];
            if (i > 1) row[r - 1] = st1[c];
            if (i < nx) row[r + 1] = st2[c];
            if (j > 1) row[r - nx] = st3[c];
            if (j < ny) row[r + nx] = st4[c];
            if (k > 1) row[r - nx * ny] = st5[c];
            if (k < L.nz) row[r + nx * ny] = st6[c];
        }
        if (tid == 0 && status) *status = 0;
        __syncthreads();
    };
    if (n <= kDenseLdsN) {
        assemble(sM);
        // LDS-resident matrix, one column per lane. Wave 0 alone finds the
        // pivot (a butterfly over the candidate rows, one per lane), swaps
        // the rows and scales the pivot row, with no barrier between those
        // steps; then every wave updates its rows (wave w of nw: rows
        // w + nw m), each lane loading a batch of rows before it stores them
        // (any whole number of waves: the host picks it). Two barriers
        // per pivot. The row swaps are recorded as the original row index
        // at each position (pos, one per lane of wave 0) and undone at the
        // end by scattering each row's columns to their original places,
        // which moves the same values as swapping the columns back, last
        // swap first.
        const int lane = tid & 63, wv = tid >> 6, nw = nt >> 6;
        constexpr int kRows = 4;   // rows per wave and batch: all loads of a batch before its stores
        int pos = lane;
        for (int kk = 0; kk < n; ++kk) {
            if (wv == 0) {
                const int r = kk + lane;
                double v = r < n ? fabs(sM[r * n + kk]) : -1.0;
                if (!(v >= 0.0)) v = -1.0;   // NaN never wins
                int bi = r;
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) {
                    const double ov = __shfl_xor(v, off, 64);
                    const int oi = __shfl_xor(bi, off, 64);
                    if (ov > v || (ov == v && oi < bi)) v = ov, bi = oi;
                }
                const int p = bi;   // the same in every lane
                if (lane == 0 && !(v > 0.0) && status) *status = -1;
                const int pk = __shfl(pos, kk, 64), pp = __shfl(pos, p, 64);
                if (lane == kk)
                    pos = pp;
                else if (lane == p)
                    pos = pk;
                double a = 0.0;
                if (lane < n) {
                    a = sM[kk * n + lane];
                    if (p != kk) {
                        const double b = sM[p * n + lane];
                        sM[p * n + lane] = a;
                        a = b;
                    }
                }
                const double inv = 1.0 / __shfl(a, kk, 64);
                if (lane < n) sM[kk * n + lane] = (lane == kk ? 1.0 : a) * inv;
            }
            __syncthreads();
            const double rkc = lane < n ? sM[kk * n + lane] : 0.0;
            for (int r0 = wv; r0 < n; r0 += kRows * nw) {
                double f[kRows], x[kRows];
#pragma unroll
                for (int m = 0; m < kRows; ++m) {
                    const int r = r0 + m * nw;
                    f[m] = r < n ? sM[r * n + kk] : 0.0;
                    x[m] = (r < n && lane < n) ? sM[r * n + lane] : 0.0;
                }
#pragma unroll
                for (int m = 0; m < kRows; ++m) {
                    const int r = r0 + m * nw;
                    if (r < n && r != kk && lane < n) sM[r * n + lane] = (lane == kk ? 0.0 : x[m]) - f[m] * rkc;
                }
            }
            __syncthreads();
        }
        if (wv == 0 && lane < n) perm[lane] = pos;
        __syncthreads();
        for (int r = wv; r < n; r += nw)
            if (lane < n) Ainv[r * n + perm[lane]] = sM[r * n + lane];
        return;
    }
    double *const M = gwork;
    assemble(M);
    for (int kk = 0; kk < n; ++kk) {
        // pivot: the largest |M[r][kk]|, r >= kk, lowest r on ties
        double best = -1.0;
        int bi = kk;
        for (int r = kk + tid; r < n; r += nt) {
            const double v = fabs(M[(long long)r * n + kk]);
            if (v > best) best = v, bi = r;
        }
        for (int off = 32; off > 0; off >>= 1) {
            const double ov = __shfl_down(best, off, 64);
            const int oi = __shfl_down(bi, off, 64);
            if (ov > best || (ov == best && oi < bi)) best = ov, bi = oi;
        }
        if ((tid & 63) == 0) rv[tid >> 6] = best, ri[tid >> 6] = bi;
        __syncthreads();
        if (tid == 0) {
            double b = rv[0];
            int p = ri[0];
            for (int w = 1; w < nt / 64; ++w)
                if (rv[w] > b || (rv[w] == b && ri[w] < p)) b = rv[w], p = ri[w];
            piv = p;
            perm[kk] = p;
            if (!(b > 0.0) && status) *status = -1;
        }
        __syncthreads();
        const int p = piv;
        double *rk = M + (long long)kk * n;
        if (p != kk) {
            double *rp = M + (long long)p * n;
            for (int q = tid; q < n; q += nt) {
                const double t = rk[q];
                rk[q] = rp[q];
                rp[q] = t;
            }
            __syncthreads();
        }
        const double inv = 1.0 / rk[kk];
        __syncthreads();
        // pivot row: M[kk][kk] = 1, then the row times 1 / pivot
        for (int q = tid; q < n; q += nt) rk[q] = (q == kk ? 1.0 : rk[q]) * inv;
        // the pivot column of the other rows, before the update
        for (int r = tid; r < n; r += nt) col[r] = M[(long long)r * n + kk];
        __syncthreads();
        // rows over the waves, columns over the lanes (no index division)
        for (int r = tid >> 6; r < n; r += nt >> 6) {
            if (r == kk) continue;
            const double f = col[r];
            double *row = M + (long long)r * n;
            for (int cidx = tid & 63; cidx < n; cidx += 64) row[cidx] = (cidx == kk ? 0.0 : row[cidx]) - f * rk[cidx];
        }
        __syncthreads();
    }
    // undo the row swaps as column swaps, last first
    for (int kk = n - 1; kk >= 0; --kk) {
        const int p = perm[kk];
        if (p != kk) {
            for (int r = tid; r < n; r += nt) {
                double *row = M + (long long)r * n;
                const double t = row[kk];
                row[kk] = row[p];
                row[p] = t;
            }
        }
        __syncthreads();
    }
    for (long long q = tid; q < nn; q += nt) Ainv[q] = M[q];
}

}  // namespace mps

namespace mps {

// dB/dx of the Bernoulli function, B'(x) = B(x) (1 - B(-x)) / x, with its
// Taylor series near 0 (B' = -1/2 + x/6 - x^3/180).
__device__ __forceinline__ double bernoulli_prime(double x)
{
    if (fabs(x) < 1e-3) {
        const double x2 = x * x;
        return -0.5 + x * (1.0 / 6.0 - x2 * (1.0 / 180.0));
    }
    return bernoulli(x) * (1.0 - bernoulli(-x)) / x;
}

// d(wL)/d(c) and d(wR)/d(c) of a face's flux weights with respect to the
// face velocity c = (cL + cR)/2 (SG: dwL/dc = -B'(-Pe), dwR/dc = -B'(Pe);
// upwind: the flux is c u_upwind); for vanishing D the SG weights are the
// upwind ones with the cell velocities, returned through dL/dR per cell.
template <int FLUX>
__device__ __forceinline__ void flux_dc(double cL, double cR, double DL, double DR, double d, double uL, double uR,
                                        double &g_face, double &g_L, double &g_R)
{
    const double ch = 0.5 * (cL + cR);
    g_face = 0.0;
    g_L = 0.0;
    g_R = 0.0;
    if (FLUX == MPS_FLUX_UPWIND_K) {
        g_face = ch > 0.0 ? uL : uR;
        return;
    }
    const double eps = 1e-10;
    const double Dh = 2.0 * DL * DR / (DL + DR + eps);
    if (fabs(Dh) < eps) {
        const double sgn = ch <= 0.0 ? -1.0 : 1.0;
        g_L = 0.5 * (1.0 + sgn) * uL;
        g_R = 0.5 * (1.0 - sgn) * uR;
        return;
    }
    const double Pe = ch * d / Dh;
    g_face = -bernoulli_prime(-Pe) * uL - bernoulli_prime(Pe) * uR;
}

// Velocity sensitivity of the adjoint pairing:  g_v(c) = -d(lam^T A(v) x)/dv(c)
// for the three cell velocity components, i.e. the gradient of a scalar J
// through x = A(v)^-1 b once lam = A^-T dJ/dx. Each face couples its two
// cells through the face velocity (v_L + v_R)/2 and the flux pairing
// (lam_L/w_L - lam_R/w_R) F; the boundary rows add their convective terms.
// Every cell evaluates both of its faces, so no atomics are needed.
template <int OP, int FLUX>
__global__ void __launch_bounds__(kBlock) k_velocity_grad(Lev<double> L, const double *__restrict__ lam,
                                                         const double *__restrict__ x, double *gx, double *gy,
                                                         double *gz)
{
    int i, j, k;
    long long c;
    if (!cell_of(L, i, j, k, c)) return;
    const double *vel[3] = {L.vx, L.vy, L.vz};
    const long long s[3] = {1, L.sy, L.sz};
    const int idx[3] = {i, j, k};
    const int nn[3] = {L.nx, L.ny, L.nz};
    const double h[3] = {L.dx, L.dy, L.dz};
    const double ai[3] = {L.ax, L.ay, L.az};
    const double ab[3] = {L.axb, L.ayb, L.azb};
    double *out[3] = {gx, gy, gz};
    const long long n = (long long)(i - 1) + (long long)L.nx * ((j - 1) + (long long)L.ny * (k - 1));
    for (int dim = 0; dim < 3; ++dim) {
        double g = 0.0;
        const double vc = vel[dim] ? vel[dim][c] : 0.0;
        auto abv_of = [&](int ii) {
            const bool edge = ii == 1 || ii == nn[dim];
            return (OP == OP_ALIGNED && edge) ? ab[dim] : ai[dim];
        };
        if (idx[dim] > 1) {  // face (m | c)
            const long long m = c - s[dim];
            double gf, gl, gr;
            flux_dc<FLUX>(vel[dim] ? vel[dim][m] : 0.0, vc, L.D[m], L.D[c], h[dim], x[m], x[c], gf, gl, gr);
            const double pair = lam[m] * abv_of(idx[dim] - 1) - lam[c] * abv_of(idx[dim]);
            g += pair * (0.5 * gf + gr);
        }
        if (idx[dim] < nn[dim]) {  // face (c | p)
            const long long p = c + s[dim];
            double gf, gl, gr;
            flux_dc<FLUX>(vc, vel[dim] ? vel[dim][p] : 0.0, L.D[c], L.D[p], h[dim], x[c], x[p], gf, gl, gr);
            const double pair = lam[c] * abv_of(idx[dim]) - lam[p] * abv_of(idx[dim] + 1);
            g += pair * (0.5 * gf + gl);
        }
        // boundary rows: DIRC upwind outflow and ZGRD convection
        // (discretepde.f90:810-828 and the + face copy)
        const double abv = abv_of(idx[dim]);
        const double hh = OP == OP_REF ? 1.0 / h[dim] : abv;
        if (idx[dim] == 1) {
            const int code = L.bc[2 * dim];
            if ((code == BC_DIRC && vc < 0.0) || code == BC_ZGRD) g += lam[c] * (-hh) * x[c];
        }
        if (idx[dim] == nn[dim]) {
            const int code = L.bc[2 * dim + 1];
            if ((code == BC_DIRC && vc > 0.0) || code == BC_ZGRD) g += lam[c] * hh * x[c];
        }
        out[dim][n] = -g;
    }
}

}  // namespace mps
