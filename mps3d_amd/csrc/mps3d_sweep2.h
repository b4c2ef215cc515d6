// Two red-black Gauss-Seidel sweeps in one pass over HBM (gfx950).
//
// k_sweep2_fast performs red, black, red, black -- the two sweeps of a
// pre- or post-smoothing step (perform_gseidel_smoothing with nsmoothsteps
// = 2, mgridlib/discretepde.f90:132-236, mgridsteps.f90:473) -- reading u,
// the coefficient field D and b once and writing u once (temporal blocking).
// Rows are the fast face-factor rows of k_sweep_fast (no convection, no
// reaction term, uniform face factors: OP_REF), so the result is the same
// arithmetic as two k_sweep_fast launches up to the rounding of the DIRC
// wall term (below).
//
// A workgroup owns a TX x TY column of cells and marches its chunk of kz
// planes. Per step s it stages plane s + 1 of u and D (and b) on the tile
// plus a 4-cell ring into LDS rings, then, separated by barriers, computes
//   R1: red cells of plane s     on the tile + 3-cell ring (from old black),
//   B1: black cells of plane s-1 on the tile + 2-cell ring (from R1),
//   R2: red cells of plane s-2   on the tile + 1-cell ring (from B1),
//   B2: black cells of plane s-3 on the tile               (from R2),
// all in place in one u ring (each stage writes its colour of its plane, and
// no stage of the same step reads what another writes). R2's and B2's tile
// cells are final and go straight to uout. The rings hold every plane a
// stage of this or the previous step still reads: u and D planes s-5 .. s+1
// (7), b planes s-4 .. s+1 (6).
//
// Boundary rows without per-cell factors: values outside the interior are
// staged as zeros for u and b, and the coefficient staged one cell outside a
// physical face is D_c * m_f with D_c the coefficient of the boundary cell
// it faces: m_f = -1 makes the face weight (D_c + D_ghost) h exactly zero
// (no flux), m_f = g/h - 1 makes it D_c g (the DIRC wall term). The zero
// ghost u leaves the off-diagonal of the face at zero either way.
//
// Staging loads are buffer loads with the plane offset in a scalar register
// and a per-lane 32-bit offset fixed for the whole launch (the host keeps
// ghosted fields below 2^31 bytes).
#pragma once

#include "mps3d_sweep.h"

namespace mps {

constexpr int kSweep2Threads = 512;


template <typename T, int TX, int TY, bool ZG>
__global__ void __launch_bounds__(kSweep2Threads) k_sweep2_fast(Lev<T> L, GhostD<T> G, const T *__restrict__ uin,
                                                                T *__restrict__ uout, const T *__restrict__ b, int kz)
{
#pragma clang fp contract(fast)
    constexpr int HW = 4;
    constexpr int W = TX + 2 * HW, H = TY + 2 * HW, P = W * H;
    constexpr int NT = kSweep2Threads;
    constexpr int NS = (P + NT - 1) / NT;  // staging positions per lane
    constexpr int RU = 7, RB = 6;          // ring depths: u and D, b
    static_assert(TX % 2 == 0, "tile width must be even");
    static_assert((TX + 6) / 2 * (TY + 6) <= NT, "one R1 cell per lane");
    __shared__ T sU[RU][P];
    __shared__ T sD[RU][P];
    __shared__ T sB[RB][P];

    const int tid = threadIdx.x;
    int bx, by, bz;
    xcd_tile(bx, by, bz);
    const int i0 = bx * TX + 1, j0 = by * TY + 1;
    const int kb = bz * kz + 1, ke = min(kb + kz - 1, L.nz);
    const int nx = L.nx, ny = L.ny, nz = L.nz, sy = L.sy;
    const int szb = (int)L.sz * (int)sizeof(T);   // plane stride in bytes
    const int nbytes = (int)((L.nz + 2) * L.sz * (long long)sizeof(T));
    const __amdgpu_buffer_rsrc_t Ru = __builtin_amdgcn_make_buffer_rsrc((void *)uin, (short)0, nbytes, kBufferWord3);
    const __amdgpu_buffer_rsrc_t Rd = __builtin_amdgcn_make_buffer_rsrc((void *)L.D, (short)0, nbytes, kBufferWord3);
    const __amdgpu_buffer_rsrc_t Rb = __builtin_amdgcn_make_buffer_rsrc((void *)b, (short)0, nbytes, kBufferWord3);
    const __amdgpu_buffer_rsrc_t Ro = __builtin_amdgcn_make_buffer_rsrc((void *)uout, (short)0, nbytes, kBufferWord3);

    // staging positions: xy byte offset of the position clamped into the
    // interior (a valid address on every plane), inside flag, and the
    // coefficient multiplier (1 inside, m_f one cell outside one x or y
    // face, 0 further out: never read by an interior row)
    int so[NS];
    bool sin_[NS], son[NS];
    T smul[NS];
#pragma unroll
    for (int r = 0; r < NS; ++r) {
        const int q = tid + r * NT;
        son[r] = q < P;
        const int gi = i0 - HW + q % W, gj = j0 - HW + q / W;
        const int ci = min(max(gi, 1), nx), cj = min(max(gj, 1), ny);
        so[r] = (ci + sy * cj) * (int)sizeof(T);
        const bool xin = gi >= 1 && gi <= nx, yin = gj >= 1 && gj <= ny;
        sin_[r] = son[r] && xin && yin;
        T m = T(0);
        if (xin && yin) m = T(1);
        else if (yin && gi == 0) m = G.m[LEFT];
        else if (yin && gi == nx + 1) m = G.m[RIGHT];
        else if (xin && gj == 0) m = G.m[BOTTOM];
        else if (xin && gj == ny + 1) m = G.m[TOP];
        smul[r] = m;
    }
    T pu[NS], pD[NS], pb[NS];
    // plane kp into registers (raw values; masks applied at the store)
    auto gload = [&](int kp) {
        const int soff = min(max(kp, 1), nz) * szb;
#pragma unroll
        for (int r = 0; r < NS; ++r) {
            pD[r] = buf_load<T>(Rd, so[r], soff);
            pb[r] = buf_load<T>(Rb, so[r], soff);
            pu[r] = ZG ? T(0) : buf_load<T>(Ru, so[r], soff);
        }
    };
    auto lstore = [&](int kp, int su, int sb) {
        const bool zin = kp >= 1 && kp <= nz;
        const T zm = kp == 0 ? G.m[BACK] : kp == nz + 1 ? G.m[FRONT] : T(0);
#pragma unroll
        for (int r = 0; r < NS; ++r) {
            if (!son[r]) continue;
            const int q = tid + r * NT;
            const bool in = sin_[r] && zin;
            const T m = zin ? smul[r] : (sin_[r] ? zm : T(0));
            sU[su][q] = in ? pu[r] : T(0);
            sD[su][q] = pD[r] * m;
            sB[sb][q] = in ? pb[r] : T(0);
        }
    };

    // stage lanes: stage x (halo h = 3 - x, red for x = 0, 2) owns cell
    // (2 mm + p, jj) of its region's row jj, p set by the plane parity
    struct Lane {
        bool act;     // this lane has a cell in the stage
        bool jin;     // its row is an interior row
        bool tile;    // (R2) its row lies in the tile
        int ebase;    // LDS index of column 2 mm of its row
        int ibase;    // global i of column 2 mm
        int obase;    // byte offset of (ibase, j) within a plane
        int par;      // parity of (colour + i0 + j0 + jj)
    };
    auto lane_of = [&](int h, int colour) {
        Lane a{};
        const int rw = (TX + 2 * h) / 2, n = rw * (TY + 2 * h);
        a.act = tid < n;
        const int jj = a.act ? tid / rw : 0, mm = a.act ? tid - jj * rw : 0;
        const int j = j0 - h + jj;
        a.jin = a.act && j >= 1 && j <= ny;
        a.tile = jj >= 1 && jj <= TY;
        a.ebase = (2 * mm + HW - h) + W * (jj + HW - h);
        a.ibase = i0 - h + 2 * mm;
        a.obase = (a.ibase + sy * j) * (int)sizeof(T);
        a.par = (colour + i0 + j0 + jj) & 1;
        return a;
    };
    const Lane LR1 = lane_of(3, 1), LB1 = lane_of(2, 0), LR2 = lane_of(1, 1), LB2 = lane_of(0, 0);

    const T hx = L.hx, hy = L.hy, hz = L.hz, dtinv = L.dtinv;
    // Gauss-Seidel value of the cell at LDS index e of plane slot s0 (planes
    // below and above in sm, sp), right-hand side from b slot bs
    auto update = [&](int e, int sm, int s0, int sp, int bs) {
        const T *U0 = sU[s0], *Um = sU[sm], *Up = sU[sp];
        const T *D0 = sD[s0], *Dm = sD[sm], *Dp = sD[sp];
        const T Dc = D0[e];
        const T wxm = (D0[e - 1] + Dc) * hx, wxp = (Dc + D0[e + 1]) * hx;
        const T wym = (D0[e - W] + Dc) * hy, wyp = (Dc + D0[e + W]) * hy;
        const T wzm = (Dm[e] + Dc) * hz, wzp = (Dc + Dp[e]) * hz;
        T d = ((wxm + wxp) + (wym + wyp)) + (wzm + wzp);
        d = d + dtinv;
        T s = sB[bs][e] + wxm * U0[e - 1];
        s = s + wxp * U0[e + 1];
        s = s + wym * U0[e - W];
        s = s + wyp * U0[e + W];
        s = s + wzm * Um[e];
        s = s + wzp * Up[e];
        return s * fast_rcp(d);
    };
    // one stage on plane k (slots of k-1, k, k+1 and of b(k)); returns the
    // value and whether the lane's cell is an interior cell
    auto stage = [&](const Lane &a, int k, int sm, int s0, int sp, int bs, T &val, int &voff) {
        const int p = a.par ^ (k & 1);
        const int i = a.ibase + p;
        const bool ok = a.jin && i >= 1 && i <= nx;
        voff = a.obase + p * (int)sizeof(T);
        if (ok) {
            const int e = a.ebase + p;
            val = update(e, sm, s0, sp, bs);
            sU[s0][e] = val;
        }
        return ok;
    };

    // ring slots of plane kb - 4 + t: u/D t % 7, b t % 6
    int u0 = 0, b0s = 0;   // slots of plane kb - 4
    gload(kb - 4);
    lstore(kb - 4, 0, 0);
    gload(kb - 3);
    lstore(kb - 3, 1, 1);
    gload(kb - 2);
    // slots of plane s at step s = kb - 3
    int us = 1, bs = 1;
    auto inc7 = [](int a) { return a == RU - 1 ? 0 : a + 1; };
    auto dec7 = [](int a) { return a == 0 ? RU - 1 : a - 1; };
    auto inc6 = [](int a) { return a == RB - 1 ? 0 : a + 1; };
    auto dec6 = [](int a) { return a == 0 ? RB - 1 : a - 1; };
    (void)u0;
    (void)b0s;
    const int r1lo = max(1, kb - 3), r1hi = min(nz, ke + 3);
    const int b1lo = max(1, kb - 2), b1hi = min(nz, ke + 2);
    const int r2lo = max(1, kb - 1), r2hi = min(nz, ke + 1);
    for (int s = kb - 3; s <= ke + 3; ++s) {
        // slots of planes s+1, s, s-1, s-2, s-3, s-4 (u/D) and s+1 .. s-3 (b)
        const int up1 = inc7(us), um1 = dec7(us), um2 = dec7(um1), um3 = dec7(um2), um4 = dec7(um3);
        const int bp1 = inc6(bs), bm1 = dec6(bs), bm2 = dec6(bm1), bm3 = dec6(bm2);
        lstore(s + 1, up1, bp1);
        __syncthreads();
        if (s + 2 <= ke + 4) gload(s + 2);
        T val;
        int voff;
        // R1: red, plane s, tile + 3
        if (s >= r1lo && s <= r1hi && LR1.act) stage(LR1, s, um1, us, up1, bs, val, voff);
        __syncthreads();
        // B1: black, plane s-1, tile + 2
        if (s - 1 >= b1lo && s - 1 <= b1hi && LB1.act) stage(LB1, s - 1, um2, um1, us, bm1, val, voff);
        __syncthreads();
        // R2: red, plane s-2, tile + 1; tile cells are final
        if (s - 2 >= r2lo && s - 2 <= r2hi && LR2.act) {
            const int k = s - 2;
            if (stage(LR2, k, um3, um2, um1, bm2, val, voff)) {
                const int ii = (LR2.ibase + (LR2.par ^ (k & 1))) - (i0 - 1);
                if (LR2.tile && ii >= 1 && ii <= TX && k >= kb && k <= ke) buf_store<T>(val, Ro, voff, k * szb);
            }
        }
        __syncthreads();
        // B2: black, plane s-3, tile: final
        if (s - 3 >= kb && s - 3 <= ke && LB2.act) {
            const int k = s - 3;
            const int p = LB2.par ^ (k & 1);
            const int i = LB2.ibase + p;
            if (LB2.jin && i <= nx) {
                const T v = update(LB2.ebase + p, um4, um3, um2, bm3);
                buf_store<T>(v, Ro, LB2.obase + p * (int)sizeof(T), k * szb);
            }
        }
        us = up1;
        bs = bp1;
    }
}

// One fused red-black sweep (red, then black; the schedule of k_sweep_fast)
// for rows with uniform face factors and no reaction term (OP_REF: the
// finest level of either hierarchy), fast arithmetic, with the boundary
// rows carried by the staged ghost coefficients as in k_sweep2_fast and
// scalar-register plane offsets for every global access.
//
// A workgroup owns a TX x TY column of cells and marches its chunk of kz
// planes. Step s stages plane s + 1 of u and D on the tile plus a 2-cell
// ring, then computes the red cells of plane s on the tile plus a 1-cell
// ring (from old black values, written in place into the u ring) and,
// after a barrier, the black cells of plane s - 1 on the tile (from the new
// red values). Red and black tile cells go to uout. The rings hold planes
// s - 3 .. s + 1 (the black pass of step s - 1 still reads s - 3 when step
// s stages s + 1). Each lane keeps its red and black cells' b in
// registers, loaded one step ahead; the next plane's u and D loads are in
// flight during the step.
template <typename T, int TX, int TY, bool ZG>
__global__ void __launch_bounds__(TX *TY) k_sweep_ref(Lev<T> L, GhostD<T> G, const T *__restrict__ uin,
                                                     T *__restrict__ uout, const T *__restrict__ b, int kz)
{
#pragma clang fp contract(fast)
    constexpr int HW = 2;
    constexpr int W = TX + 2 * HW, H = TY + 2 * HW, P = W * H;
    constexpr int NT = TX * TY;
    constexpr int NS = (P + NT - 1) / NT;
    constexpr int RU = 5;
    static_assert(TX % 2 == 0, "tile width must be even");
    static_assert((TX + 2) / 2 * (TY + 2) <= NT, "one red cell per lane");
    __shared__ T sU[RU][P];
    __shared__ T sD[RU][P];

    const int tid = threadIdx.x;
    int bx, by, bz;
    xcd_tile(bx, by, bz);
    const int i0 = bx * TX + 1, j0 = by * TY + 1;
    const int kb = bz * kz + 1, ke = min(kb + kz - 1, L.nz);
    const int nx = L.nx, ny = L.ny, nz = L.nz, sy = L.sy;
    const int szb = (int)L.sz * (int)sizeof(T);
    const int nbytes = (int)((L.nz + 2) * L.sz * (long long)sizeof(T));
    const __amdgpu_buffer_rsrc_t Ru = __builtin_amdgcn_make_buffer_rsrc((void *)uin, (short)0, nbytes, kBufferWord3);
    const __amdgpu_buffer_rsrc_t Rd = __builtin_amdgcn_make_buffer_rsrc((void *)L.D, (short)0, nbytes, kBufferWord3);
    const __amdgpu_buffer_rsrc_t Rb = __builtin_amdgcn_make_buffer_rsrc((void *)b, (short)0, nbytes, kBufferWord3);
    const __amdgpu_buffer_rsrc_t Ro = __builtin_amdgcn_make_buffer_rsrc((void *)uout, (short)0, nbytes, kBufferWord3);

    int so[NS];
    bool sin_[NS], son[NS];
    T smul[NS];
#pragma unroll
    for (int r = 0; r < NS; ++r) {
        const int q = tid + r * NT;
        son[r] = q < P;
        const int gi = i0 - HW + q % W, gj = j0 - HW + q / W;
        const int ci = min(max(gi, 1), nx), cj = min(max(gj, 1), ny);
        so[r] = (ci + sy * cj) * (int)sizeof(T);
        const bool xin = gi >= 1 && gi <= nx, yin = gj >= 1 && gj <= ny;
        sin_[r] = son[r] && xin && yin;
        T m = T(0);
        if (xin && yin) m = T(1);
        else if (yin && gi == 0) m = G.m[LEFT];
        else if (yin && gi == nx + 1) m = G.m[RIGHT];
        else if (xin && gj == 0) m = G.m[BOTTOM];
        else if (xin && gj == ny + 1) m = G.m[TOP];
        smul[r] = m;
    }
    T pu[NS], pD[NS];
    auto gload = [&](int kp) {
        const int soff = min(max(kp, 1), nz) * szb;
#pragma unroll
        for (int r = 0; r < NS; ++r) {
            pD[r] = buf_load<T>(Rd, so[r], soff);
            pu[r] = ZG ? T(0) : buf_load<T>(Ru, so[r], soff);
        }
    };
    auto lstore = [&](int kp, int su) {
        const bool zin = kp >= 1 && kp <= nz;
        const T zm = kp == 0 ? G.m[BACK] : kp == nz + 1 ? G.m[FRONT] : T(0);
#pragma unroll
        for (int r = 0; r < NS; ++r) {
            if (!son[r]) continue;
            const int q = tid + r * NT;
            const T m = zin ? smul[r] : (sin_[r] ? zm : T(0));
            sU[su][q] = (sin_[r] && zin) ? pu[r] : T(0);
            sD[su][q] = pD[r] * m;
        }
    };

    // red lane: cell (2 mm + p, jj) of the 1-cell-ring region; black lane:
    // cell (2 mm + p, jj) of the tile; p from the plane parity
    const int rrw = (TX + 2) / 2, brw = TX / 2;
    const bool ract = tid < rrw * (TY + 2), bact = tid < brw * TY;
    const int rjj = ract ? tid / rrw : 0, rmm = ract ? tid - rjj * rrw : 0;
    const int bjj = bact ? tid / brw : 0, bmm = bact ? tid - bjj * brw : 0;
    const int rj = j0 - 1 + rjj, bj = j0 + bjj;
    const bool rjin = ract && rj >= 1 && rj <= ny, bjin = bact && bj <= ny;
    const bool rtile = rjj >= 1 && rjj <= TY;
    const int re = (2 * rmm + HW - 1) + W * (rjj + HW - 1), be = (2 * bmm + HW) + W * (bjj + HW);
    const int ri = i0 - 1 + 2 * rmm, bi = i0 + 2 * bmm;
    const int ro = (ri + sy * rj) * (int)sizeof(T), bo = (bi + sy * bj) * (int)sizeof(T);
    const int rpar = (1 + i0 + j0 + rjj) & 1, bpar = (i0 + j0 + bjj) & 1;

    const T hx = L.hx, hy = L.hy, hz = L.hz, dtinv = L.dtinv;
    // Gauss-Seidel value of the cell at LDS index e: in-plane neighbours
    // through one base at e - W (non-negative immediate offsets)
    auto update = [&](int e, int sm, int s0, int sp, T bv) {
        const T *U0 = &sU[s0][e - W], *D0 = &sD[s0][e - W];
        const T Dc = D0[W];
        const T wxm = (D0[W - 1] + Dc) * hx, wxp = (Dc + D0[W + 1]) * hx;
        const T wym = (D0[0] + Dc) * hy, wyp = (Dc + D0[2 * W]) * hy;
        const T wzm = (sD[sm][e] + Dc) * hz, wzp = (Dc + sD[sp][e]) * hz;
        T d = ((wxm + wxp) + (wym + wyp)) + (wzm + wzp);
        d = d + dtinv;
        T s = bv + wxm * U0[W - 1];
        s = s + wxp * U0[W + 1];
        s = s + wym * U0[0];
        s = s + wyp * U0[2 * W];
        s = s + wzm * sU[sm][e];
        s = s + wzp * sU[sp][e];
        return s * fast_rcp(d);
    };

    // b of this lane's red cell on plane k and black cell on plane k - 1
    T rb = T(0), bb = T(0);
    auto bload = [&](int k, T &r_, T &b_) {
        const int pr = rpar ^ (k & 1), pb = bpar ^ ((k - 1) & 1);
        r_ = buf_load<T>(Rb, ract ? ro + pr * (int)sizeof(T) : 0, min(max(k, 0), nz + 1) * szb);
        b_ = buf_load<T>(Rb, bact ? bo + pb * (int)sizeof(T) : 0, min(max(k - 1, 0), nz + 1) * szb);
    };

    gload(kb - 2);
    lstore(kb - 2, 0);
    gload(kb - 1);
    lstore(kb - 1, 1);
    gload(kb);
    bload(kb - 1, rb, bb);
    int us = 1;   // slot of plane s
    const int rlo = max(1, kb - 1), rhi = min(nz, ke + 1);
    for (int s = kb - 1; s <= ke + 1; ++s) {
        const int up1 = us == RU - 1 ? 0 : us + 1, um1 = us == 0 ? RU - 1 : us - 1, um2 = um1 == 0 ? RU - 1 : um1 - 1;
        lstore(s + 1, up1);
        const T rbv = rb, bbv = bb;
        __syncthreads();
        if (s + 2 <= ke + 2) gload(s + 2);
        bload(s + 1, rb, bb);
        // red cells of plane s on the tile + 1 ring, in place
        if (s >= rlo && s <= rhi && ract) {
            const int p = rpar ^ (s & 1);
            const int i = ri + p;
            if (rjin && i >= 1 && i <= nx) {
                const int e = re + p;
                const T v = update(e, um1, us, up1, rbv);
                sU[us][e] = v;
                const int ii = i - (i0 - 1);
                if (rtile && ii >= 1 && ii <= TX && s >= kb && s <= ke)
                    buf_store<T>(v, Ro, ro + p * (int)sizeof(T), s * szb);
            }
        }
        __syncthreads();
        // black cells of plane s - 1 on the tile: final
        const int k = s - 1;
        if (k >= kb && k <= ke && bact) {
            const int p = bpar ^ (k & 1);
            if (bjin && bi + p <= nx) {
                const T v = update(be + p, um2, um1, us, bbv);
                buf_store<T>(v, Ro, bo + p * (int)sizeof(T), k * szb);
            }
        }
        us = up1;
    }
}

}  // namespace mps
