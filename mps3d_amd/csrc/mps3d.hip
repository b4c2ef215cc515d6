// mps3d_amd host driver and C ABI (include/mps3d.h).
//
// One context = one grid on one GPU. All fields stay resident in HBM across
// the V-cycles of a solve; the host only sequences kernels on the context's
// stream and reads back one residual norm per cycle.
#include "../../include/mps3d.h"
#include "mps3d_bottom.h"
#include "mps3d_kernels.h"
#include "mps3d_sweep.h"
#include "mps3d_sweep2.h"
#include "mps3d_lds_cycle.h"
#include "mps3d_transport.h"
#include "mps3d_stream.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

using namespace mps;

namespace {

thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

struct HipError {
    hipError_t e;
    const char *what;
};

#define HIPCHK(expr)                                              \
    do {                                                          \
        hipError_t e_ = (expr);                                   \
        if (e_ != hipSuccess) throw HipError{e_, #expr};          \
    } while (0)

template <bool B>
using BoolC = std::integral_constant<bool, B>;
template <int I>
using IntC = std::integral_constant<int, I>;

// runtime (op, vel, reac) -> compile-time template arguments
template <typename F>
void with_op(int op, bool vel, bool reac, F &&f)
{
    if (op == OP_STENCIL) {
        f(IntC<OP_STENCIL>{}, BoolC<false>{}, BoolC<false>{});
        return;
    }
    auto inner = [&](auto O) {
        if (vel) {
            if (reac) f(O, BoolC<true>{}, BoolC<true>{});
            else f(O, BoolC<true>{}, BoolC<false>{});
        } else {
            if (reac) f(O, BoolC<false>{}, BoolC<true>{});
            else f(O, BoolC<false>{}, BoolC<false>{});
        }
    };
    if (op == OP_REF) inner(IntC<OP_REF>{});
    else inner(IntC<OP_ALIGNED>{});
}

bool is_device_ptr(const void *p)
{
    if (!p) return false;
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice;
}

template <typename T>
T *dalloc(size_t n)
{
    void *p = nullptr;
    HIPCHK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
    return static_cast<T *>(p);
}

void dfree(void *p)
{
    if (p) (void)hipFree(p);
}

// geometry of one level (host side)
struct Geom {
    int nx, ny, nz;        // global cells
    int lx, ly, lz;        // per-rank cells (reference hierarchy)
    double dx, dy, dz;     // centre spacing
    double ax, ay, az;     // 1/width interior
    double axb, ayb, azb;  // 1/width boundary (aligned)
    double ix0, iy0, iz0;  // 1/wall distance (aligned)
    bool bottom;
    int op;                // operator form of this level
    std::vector<double> w[3];  // aligned restriction weights into this level
    long long ng() const { return (long long)(nx + 2) * (ny + 2) * (nz + 2); }
    long long n() const { return (long long)nx * ny * nz; }
};

// device buffers of one level in precision T
template <typename T>
struct LevelBuf {
    T *u = nullptr, *b = nullptr, *r = nullptr, *uold = nullptr;
    T *D = nullptr, *vx = nullptr, *vy = nullptr, *vz = nullptr, *reac = nullptr;
    T *st[ST_N] = {};
    T *fc[4] = {};                // face form of st (fc[0] aliases st[ST_C]; see Lev::fc)
    T *faces[6] = {};
    double *w[3] = {};
    T *u2 = nullptr;              // ping-pong partner of u for fused sweeps
    const double *inv = nullptr;  // dense inverse of the coarsest operator
    bool own_u = true;
    int op = -1;  // operator form override (>= 0), e.g. transposed stencils
};

// one outer solve step captured as a HIP graph, with what it was captured
// for (rebuilt when the pointers, the kind of step or the problem change)
struct StepGraph {
    hipGraphExec_t exec = nullptr;
    const void *u = nullptr, *b = nullptr;
    int kind = -1;
    unsigned gen = ~0u;
};

// FP32 V-cycles per FP64 defect-correction step: opts.dc_inner when set
// (1..kDcInnerMax), else chosen per step from the observed contraction
constexpr int kDcInnerMax = 4;

// FGMRES(m) work vectors of the finest level, kept by the context between
// calls (allocated on the first call with a given m or more)
struct GmresBufs {
    int m = 0;
    std::vector<double *> V, Z;
    double *w = nullptr, *bg = nullptr, *dh = nullptr, *dpart = nullptr;
    double *hh = nullptr;  // pinned host mirror of dh
    int nblocks = 0;
    // device Arnoldi state: H, cs, sn, gv, y, hn, the second Gram-Schmidt
    // pass's products h2, the subtracted coefficients ny and |w|^2
    double *darn = nullptr;
    int *dstate = nullptr;
    int *hflag = nullptr;                // pinned: per step the done flag, then kk and the NaN flag
    std::vector<hipEvent_t> ev;          // one per Arnoldi step
};

}  // namespace

struct mps_ctx {
    mps_grid grid{};
    mps_opts opts{};
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<Geom> geom;
    std::vector<LevelBuf<double>> L64;
    std::vector<LevelBuf<float>> L32;
    std::vector<LevelBuf<double>> LT;   // transposed (adjoint) operators
    int bc[6] = {0, 0, 0, 0, 0, 0};
    std::vector<std::vector<double>> faces_host;  // [level*6 + face]
    double *faces0[6] = {};              // finest faces, device
    bool has_vel = false, has_reac = false, timederiv = false;
    double dt = 1.0;
    bool coeffs_set = false, bc_set = false, transpose_ready = false;
    double *source = nullptr;            // finest source, ghosted
    double *b0 = nullptr;                // finest right-hand side
    double *partials = nullptr;
    int max_partials = 0;
    double *d_scalar = nullptr;
    double *d_dcscale = nullptr;         // FP32 defect scale and its inverse
    double *h_scalar = nullptr;          // pinned
    int *d_status = nullptr;
    double *bottom64 = nullptr;
    float *bottom32 = nullptr;
    double *bottomT = nullptr;
    double *user_tmp = nullptr;          // staging for host arrays (ghosted)
    double *user_tmp2 = nullptr;         // staging for interior arrays
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t ev_in = nullptr;          // orders calls after the null stream
    double last_ms = 0.0;
    int last_cycles = 0;
    double *dinv = nullptr;              // dense inverse of the coarsest operator
    double *dinvT = nullptr;             // and of its transpose
    double *bst[ST_N] = {};              // coarsest-level rows for a device-built inverse
    double *dwork = nullptr;             // its work matrix when it does not fit LDS
    double *gbuf = nullptr;              // ghosted right-hand side of adjoint solves
    double *fmg_r = nullptr, *fmg_e = nullptr;  // FP64 defect and correction of a full-multigrid start
    bool exact_now = false;              // reference-mirroring call in progress
    // outer solve steps captured as HIP graphs and replayed (see
    // outer_step): [0] the FP64 V-cycle step, [1] the full-multigrid first
    // step, [1 + k] the defect-correction step of k FP32 V-cycles
    bool graphs = true;                  // MPS3D_NO_GRAPHS=1 turns replay off
    bool fused_rr = true;                // MPS3D_NO_FUSED_RESTRICT=1: residual and restriction apart
    bool sweep_ref = false;              // k_sweep_ref for OP_REF rows (MPS3D_SWEEP_REF=1: on)
    double omega = 1.0;                  // fast-sweep over-relaxation of solves (opts.omega;
                                         // MPS3D_OMEGA overrides)
    bool relax_now = false;              // a solve (solve_core) is running: omega applies
    int sweep2 = 1;                      // precisions whose smoothing pairs run in one pass
                                         // (bit 1 FP32, bit 2 FP64; MPS3D_SWEEP2 overrides)
    StepGraph step[2 + kDcInnerMax];
    int last_vcycles = 0;                // V-cycles of the last solve (mps_solve_vcycles)
    // in-solve kernel timing (mps_time_in_solve): the kernel and precision
    // whose launches get an event pair, the events, pairs recorded
    int kt_kernel = -1;
    bool kt_fp32 = false;
    std::vector<hipEvent_t> kt_ev;
    int kt_n = 0;
    unsigned gen = 0;                    // bumped when set_bc / set_coeffs change what
                                         // captured step graphs bake in (see touch_gen)
    int sig_bc[6] = {-1, -1, -1, -1, -1, -1};
    int sig_flags = -1;                  // has_vel | has_reac << 1 | timederiv << 2
    double sig_dt = 0.0;
    ChemDev *chem = nullptr;             // chemistry tables (mps_set_chemistry)
    ChemDev chem_host{};
    bool mg_ok = true;                   // a multigrid hierarchy exists
    std::string mg_error;                // why not, when it does not
    GmresBufs gm;                        // FGMRES work vectors (mps_gmres)

    int nlev() const { return (int)geom.size(); }
};

namespace {

dim3 cells_grid(const Geom &g) { return dim3((unsigned)((g.nx * g.ny + kBlock - 1) / kBlock), (unsigned)g.nz); }
// one colour of a red-black sweep: (nx + 1) / 2 threads per row
dim3 colour_grid(const Geom &g)
{
    return dim3((unsigned)(((g.nx + 1) / 2 * g.ny + kBlock - 1) / kBlock), (unsigned)g.nz);
}
// reductions: blocks loop over k so a grid has few partial sums
dim3 red_grid(const Geom &g)
{
    return dim3((unsigned)((g.nx * g.ny + kBlock - 1) / kBlock), (unsigned)std::min(g.nz, 16));
}
dim3 block() { return dim3(kBlock); }

template <typename T>
Lev<T> view(const mps_ctx *c, const std::vector<LevelBuf<T>> &L, int l, int op_override = -1)
{
    const Geom &g = c->geom[l];
    const LevelBuf<T> &B = L[l];
    Lev<T> v{};
    v.nx = g.nx;
    v.ny = g.ny;
    v.nz = g.nz;
    v.sy = g.nx + 2;
    v.sz = (long long)(g.nx + 2) * (g.ny + 2);
    v.D = B.D;
    v.vx = B.vx;
    v.vy = B.vy;
    v.vz = B.vz;
    v.reac = B.reac;
    for (int q = 0; q < ST_N; ++q) v.st[q] = B.st[q];
    for (int q = 0; q < 4; ++q) v.fc[q] = B.fc[q];
    v.dx = (T)g.dx;
    v.dy = (T)g.dy;
    v.dz = (T)g.dz;
    v.ax = (T)g.ax;
    v.ay = (T)g.ay;
    v.az = (T)g.az;
    v.axb = (T)g.axb;
    v.ayb = (T)g.ayb;
    v.azb = (T)g.azb;
    v.ix0 = (T)g.ix0;
    v.iy0 = (T)g.iy0;
    v.iz0 = (T)g.iz0;
    for (int f = 0; f < 6; ++f) v.bc[f] = c->bc[f];
    v.timederiv = c->timederiv ? 1 : 0;
    v.dt = (T)c->dt;
    // fast arithmetic: same operator with precomputed factors
    v.fast = (c->opts.arith == MPS_ARITH_FAST && !c->has_vel && !c->exact_now) ? 1 : 0;
    v.hx = (T)(0.5 * g.ax / g.dx);
    v.hy = (T)(0.5 * g.ay / g.dy);
    v.hz = (T)(0.5 * g.az / g.dz);
    v.hxb = (T)(0.5 * g.axb / g.dx);
    v.hyb = (T)(0.5 * g.ayb / g.dy);
    v.hzb = (T)(0.5 * g.azb / g.dz);
    const bool ref_geom = c->geom.size() == 1 || c->opts.hierarchy == MPS_HIER_REF || l == 0;
    v.gx = (T)(ref_geom ? 2.0 / (g.dx * g.dx) : g.ix0 * g.ax);
    v.gy = (T)(ref_geom ? 2.0 / (g.dy * g.dy) : g.iy0 * g.ay);
    v.gz = (T)(ref_geom ? 2.0 / (g.dz * g.dz) : g.iz0 * g.az);
    v.gxb = (T)(g.ix0 * g.axb);
    v.gyb = (T)(g.iy0 * g.ayb);
    v.gzb = (T)(g.iz0 * g.azb);
    v.dtinv = c->timederiv ? (T)(1.0 / c->dt) : T(0);
    v.omega = (c->relax_now && !c->exact_now && c->opts.arith == MPS_ARITH_FAST) ? (T)c->omega : T(1);
    return v;
}

template <typename T>
int level_op(const mps_ctx *c, const std::vector<LevelBuf<T>> &L, int l)
{
    return (l < (int)L.size() && L[l].op >= 0) ? L[l].op : c->geom[l].op;
}
int level_op(const mps_ctx *c, int l) { return c->geom[l].op; }

// ---------------------------------------------------------------- hierarchy
// Captured step graphs bake in the kernels' arguments: buffer pointers
// (fixed once allocated) and the level views' BC codes, coefficient flags and
// time step. New boundary values or coefficient values leave them valid;
// a change of those settings invalidates them.
void touch_gen(mps_ctx *c)
{
    const int flags = (c->has_vel ? 1 : 0) | (c->has_reac ? 2 : 0) | (c->timederiv ? 4 : 0);
    bool same = flags == c->sig_flags && c->dt == c->sig_dt;
    for (int f = 0; f < 6; ++f) same = same && c->bc[f] == c->sig_bc[f];
    if (same) return;
    for (int f = 0; f < 6; ++f) c->sig_bc[f] = c->bc[f];
    c->sig_flags = flags;
    c->sig_dt = c->dt;
    ++c->gen;
}

// keeps only the finest level: operator-level entry points still work,
// multigrid entry points report the reason
int no_hierarchy(mps_ctx *c, const char *why)
{
    c->geom.resize(1);
    c->geom[0].bottom = false;
    for (auto &w : c->geom[0].w) w.clear();
    c->mg_ok = false;
    c->mg_error = why;
    return MPS_OK;
}

int build_geometry(mps_ctx *c)
{
    const mps_grid &G = c->grid;
    const mps_opts &O = c->opts;
    c->geom.clear();
    if (O.hierarchy == MPS_HIER_REF) {
        if (G.px < 1 || G.py < 1 || G.pz < 1 || G.nx % G.px || G.ny % G.py || G.nz % G.pz)
            return fail(MPS_ERR_ARG, "grid %dx%dx%d does not divide into process grid %dx%dx%d", G.nx, G.ny,
                        G.nz, G.px, G.py, G.pz);
        const int glx = G.nx / G.px, gly = G.ny / G.py, glz = G.nz / G.pz;
        // g_dx = g_xlen/g_Nx, g_llenx = g_lx*g_dx (par_decompose.f90:20-42)
        const double gdx = G.lenx / G.nx, gdy = G.leny / G.ny, gdz = G.lenz / G.nz;
        const double llx = glx * gdx, lly = gly * gdy, llz = glz * gdz;
        int lx = glx, ly = gly, lz = glz;
        for (;;) {
            Geom g{};
            g.lx = lx;
            g.ly = ly;
            g.lz = lz;
            g.nx = lx * G.px;
            g.ny = ly * G.py;
            g.nz = lz * G.pz;
            g.bottom = (lx <= 1 || ly <= 1 || lz <= 1);
            if (g.bottom) {  // bottomsolve's serial lengths (mgridsteps.f90:334-336)
                g.dx = (llx * G.px) / g.nx;
                g.dy = (lly * G.py) / g.ny;
                g.dz = (llz * G.pz) / g.nz;
            } else {
                g.dx = llx / lx;
                g.dy = lly / ly;
                g.dz = llz / lz;
            }
            g.ax = 1.0 / g.dx;
            g.ay = 1.0 / g.dy;
            g.az = 1.0 / g.dz;
            g.axb = g.ax;
            g.ayb = g.ay;
            g.azb = g.az;
            g.ix0 = 2.0 / g.dx;
            g.iy0 = 2.0 / g.dy;
            g.iz0 = 2.0 / g.dz;
            g.op = O.flux == MPS_FLUX_SG ? OP_STENCIL : OP_REF;
            c->geom.push_back(g);
            if (g.bottom) break;
            // colouring by global (i+j+k) parity equals the reference's
            // per-rank offsets only for even per-rank sizes, which every
            // non-bottom level needs anyway (restriction halves them)
            if ((lx & 1) || (ly & 1) || (lz & 1)) {
                char buf[256];
                snprintf(buf, sizeof buf,
                         "per-rank level %dx%dx%d is odd above the coarsest level (the reference would "
                         "truncate it); use sizes nx/px.. = m*2^k or MPS_HIER_ALIGNED",
                         lx, ly, lz);
                return no_hierarchy(c, buf);
            }
            lx /= 2;
            ly /= 2;
            lz /= 2;
        }
        if (c->geom.back().n() > kBottomMax) {
            char buf[128];
            snprintf(buf, sizeof buf, "coarsest grid %lld cells exceeds %d", c->geom.back().n(), kBottomMax);
            return no_hierarchy(c, buf);
        }
        return MPS_OK;
    }
    if (O.hierarchy == MPS_HIER_ALIGNED) {
        const int N[3] = {G.nx, G.ny, G.nz};
        const double len[3] = {G.lenx, G.leny, G.lenz};
        double h0[3];
        for (int d = 0; d < 3; ++d) h0[d] = len[d] / N[d];
        int n[3] = {N[0], N[1], N[2]};
        double H[3] = {h0[0], h0[1], h0[2]};
        int lev = 0;
        for (;;) {
            Geom g{};
            g.nx = n[0];
            g.ny = n[1];
            g.nz = n[2];
            g.lx = n[0];
            g.ly = n[1];
            g.lz = n[2];
            const int mn = std::min(n[0], std::min(n[1], n[2]));
            const bool odd = (n[0] & 1) && (n[1] & 1) && (n[2] & 1);
            g.bottom = (mn <= 3) || !odd;
            double *dd[3] = {&g.dx, &g.dy, &g.dz};
            double *aa[3] = {&g.ax, &g.ay, &g.az};
            double *bb[3] = {&g.axb, &g.ayb, &g.azb};
            double *ii[3] = {&g.ix0, &g.iy0, &g.iz0};
            for (int d = 0; d < 3; ++d) {
                *dd[d] = H[d];
                *aa[d] = 1.0 / H[d];
                // boundary cells reach from the wall (h0/2 away) to the
                // mid-point towards the next centre
                *bb[d] = 1.0 / (0.5 * H[d] + 0.5 * h0[d]);
                *ii[d] = 2.0 / h0[d];
            }
            if (O.flux == MPS_FLUX_SG) g.op = OP_STENCIL;
            else g.op = lev == 0 ? OP_REF : OP_ALIGNED;
            if (lev > 0) {
                // restriction weights: coarse I covers fine 2I-1 fully and
                // half of 2I-2 and 2I, volume weighted
                const Geom &f = c->geom[lev - 1];
                const int fn[3] = {f.nx, f.ny, f.nz};
                const double fH[3] = {f.dx, f.dy, f.dz};
                const double fwb[3] = {1.0 / f.axb, 1.0 / f.ayb, 1.0 / f.azb};
                for (int d = 0; d < 3; ++d) {
                    g.w[d].assign(3 * (size_t)n[d], 0.0);
                    auto vol = [&](int fi) {
                        if (fi < 1 || fi > fn[d]) return 0.0;
                        return (fi == 1 || fi == fn[d]) ? fwb[d] : fH[d];
                    };
                    for (int I = 1; I <= n[d]; ++I) {
                        double a = 0.5 * vol(2 * I - 2), b = vol(2 * I - 1), cc = 0.5 * vol(2 * I);
                        double s = a + b + cc;
                        g.w[d][3 * (I - 1) + 0] = a / s;
                        g.w[d][3 * (I - 1) + 1] = b / s;
                        g.w[d][3 * (I - 1) + 2] = cc / s;
                    }
                }
            }
            c->geom.push_back(g);
            if (g.bottom) break;
            for (int d = 0; d < 3; ++d) {
                n[d] = (n[d] + 1) / 2;
                H[d] *= 2.0;
            }
            ++lev;
        }
        if (c->geom.back().n() > kBottomMax) {
            char buf[256];
            snprintf(buf, sizeof buf,
                     "aligned hierarchy needs every size = m*2^k+1 with a coarsest grid of at most %d cells; "
                     "coarsest here is %dx%dx%d",
                     kBottomMax, c->geom.back().nx, c->geom.back().ny, c->geom.back().nz);
            return no_hierarchy(c, buf);
        }
        return MPS_OK;
    }
    return fail(MPS_ERR_ARG, "unknown hierarchy %d", O.hierarchy);
}

template <typename T>
void fill(mps_ctx *c, T *p, long long n, T v);

template <typename T>
void alloc_levels(mps_ctx *c, std::vector<LevelBuf<T>> &L, bool with_coeffs)
{
    L.assign(c->geom.size(), LevelBuf<T>{});
    for (size_t l = 0; l < c->geom.size(); ++l) {
        const Geom &g = c->geom[l];
        const long long ng = g.ng();
        LevelBuf<T> &B = L[l];
        B.u = dalloc<T>(ng);
        B.u2 = dalloc<T>(ng);
        // ghost layers stay zero from here on (kernels write interiors only)
        fill<T>(c, B.u, ng, T(0));
        fill<T>(c, B.u2, ng, T(0));
        B.b = dalloc<T>(ng);
        B.r = dalloc<T>(ng);
        if (l > 0) B.uold = dalloc<T>(ng);
        if (with_coeffs) {
            B.D = dalloc<T>(ng);
            fill<T>(c, B.D, ng, T(0));
        }
        for (int d = 0; d < 3; ++d)
            if (!g.w[d].empty()) {
                B.w[d] = dalloc<double>(g.w[d].size());
                HIPCHK(hipMemcpy(B.w[d], g.w[d].data(), g.w[d].size() * sizeof(double), hipMemcpyHostToDevice));
            }
    }
}

template <typename T>
void free_levels(std::vector<LevelBuf<T>> &L)
{
    for (auto &B : L) {
        if (B.own_u) dfree(B.u);
        dfree(B.u2);
        dfree(B.b);
        dfree(B.r);
        dfree(B.uold);
        dfree(B.D);
        dfree(B.vx);
        dfree(B.vy);
        dfree(B.vz);
        dfree(B.reac);
        for (auto p : B.st) dfree(p);
        for (int q = 1; q < 4; ++q) dfree(B.fc[q]);
        for (auto p : B.faces) dfree(p);
        for (auto p : B.w) dfree(p);
    }
    L.clear();
}

template <typename T>
void fill(mps_ctx *c, T *p, long long n, T v)
{
    long long blocks = std::min<long long>((n + kBlock - 1) / kBlock, 65536);
    hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)std::max<long long>(blocks, 1)), block(), 0, c->stream, p, n, v);
}

template <typename Ti, typename To>
void convert(mps_ctx *c, const Ti *in, To *out, long long n)
{
    long long blocks = std::min<long long>((n + kBlock - 1) / kBlock, 65536);
    hipLaunchKernelGGL((k_convert<Ti, To>), dim3((unsigned)std::max<long long>(blocks, 1)), block(), 0, c->stream,
                       in, out, n);
}

// copies a caller array (host or device, ghosted layout) into dst
void stage_in(mps_ctx *c, const double *src, double *dst, long long n)
{
    if (is_device_ptr(src))
        HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    else
        HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
}

// up to six same-length device-to-device copies (a NULL source: zero fill)
// in one launch, blockIdx.y picking the array
struct CopySet {
    const double *src[6];
    double *dst[6];
};
__global__ void __launch_bounds__(kBlock) k_copy_set(CopySet S, long long n)
{
    const double *s = S.src[blockIdx.y];
    double *d = S.dst[blockIdx.y];
    for (long long t = blockIdx.x * (long long)kBlock + threadIdx.x; t < n; t += (long long)gridDim.x * kBlock)
        d[t] = s ? s[t] : 0.0;
}

// collects stage_in's device copies (and zero fills) of one length into one
// k_copy_set launch; host sources are copied at once
struct StageSet {
    mps_ctx *c;
    long long n;
    CopySet S{};
    int k = 0;
    StageSet(mps_ctx *c_, long long n_) : c(c_), n(n_) {}
    void add(const double *src, double *dst)
    {
        if (src && !is_device_ptr(src)) {
            HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
            return;
        }
        if (k == 6) flush();
        S.src[k] = src;
        S.dst[k] = dst;
        ++k;
    }
    void flush()
    {
        if (k == 0) return;
        const long long blocks = std::min<long long>((n + kBlock - 1) / kBlock, 4096);
        hipLaunchKernelGGL(k_copy_set, dim3((unsigned)blocks, (unsigned)k), dim3(kBlock), 0, c->stream, S, n);
        HIPCHK(hipGetLastError());
        k = 0;
    }
};

void stage_out(mps_ctx *c, const double *src, double *dst, long long n)
{
    if (is_device_ptr(dst))
        HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    else
        HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
}

// ------------------------------------------------------------ launches
template <typename T>
void launch_colour(mps_ctx *c, std::vector<LevelBuf<T>> &L, int l, int red, int op)
{
    Lev<T> v = view(c, L, l);
    const Geom &g = c->geom[l];
    // threads for the colour's cells only once that still leaves 256
    // workgroups (128^3: 11 % faster per launch; at 32^3 it was 5 % slower)
    const int compact = (long long)((g.nx + 1) / 2) * g.ny * g.nz >= 65536 ? 1 : 0;
    with_op(op, c->has_vel, c->has_reac, [&](auto O, auto V, auto R) {
        hipLaunchKernelGGL((k_colour<T, decltype(O)::value, decltype(V)::value, decltype(R)::value>),
                           compact ? colour_grid(g) : cells_grid(g), block(), 0, c->stream, v, L[l].u,
                           (const T *)L[l].b, red, compact);
    });
}

// z-planes per block of the plane-marching kernels: enough blocks to fill
// the chip several times over on large grids, short serial z chains (each
// plane waits on one memory round trip) on small ones
// In-solve kernel timing: while mps_time_in_solve runs a solve, a KtScope
// around a finest-level launch of the timed kernel (0 sweep, 1 residual,
// 2 restriction, 3 prolongation) records an event before and after it on
// the context's stream.
constexpr int kKtMax = 256;  // launches recorded per timed solve

struct KtScope {
    mps_ctx *c;
    bool on;
    KtScope(mps_ctx *c_, int kernel, bool fp32)
        : c(c_), on(c_->kt_kernel == kernel && c_->kt_fp32 == fp32 && c_->kt_n < kKtMax)
    {
        if (on) HIPCHK(hipEventRecord(c->kt_ev[2 * c->kt_n], c->stream));
    }
    ~KtScope()
    {
        if (on && hipEventRecord(c->kt_ev[2 * c->kt_n + 1], c->stream) == hipSuccess) ++c->kt_n;
    }
};

int z_chunk(const Geom &g, int tx, int ty, int target_blocks, int kmax)
{
    const long long cols = (long long)((g.nx + tx - 1) / tx) * ((g.ny + ty - 1) / ty);
    const long long kz = cols * g.nz / target_blocks;
    return (int)std::max(2LL, std::min((long long)kmax, kz));
}

#ifndef MPS_RES_TX
#define MPS_RES_TX 32
#endif
#ifndef MPS_RES_TY
#define MPS_RES_TY 8
#endif
#ifndef MPS_RES_RB
#define MPS_RES_RB 1
#endif
#ifndef MPS_RES_KZ
#define MPS_RES_KZ 16
#endif
#ifndef MPS_RES_TARGET
#define MPS_RES_TARGET 4096
#endif
constexpr int kResTX = MPS_RES_TX, kResTY = MPS_RES_TY, kResRB = MPS_RES_RB, kResKZ = MPS_RES_KZ;

dim3 residual_planes_grid(const Geom &g, int &kz)
{
    kz = z_chunk(g, kResTX, kResTY * kResRB, MPS_RES_TARGET, kResKZ);
    return dim3((unsigned)((g.nx + kResTX - 1) / kResTX), (unsigned)((g.ny + kResTY * kResRB - 1) / (kResTY * kResRB)),
                (unsigned)((g.nz + kz - 1) / kz));
}

// aligned-hierarchy transfers (plane-marching kernels): restriction of fine
// level l's r into level l+1's b, prolongation-correction of level l+1's u
// into level l's u
#ifndef MPS_PROL_TX
#define MPS_PROL_TX 64
#endif
#ifndef MPS_PROL_TY
#define MPS_PROL_TY 4
#endif
constexpr int kRestCX = 32, kRestCY = 8, kProlTX = MPS_PROL_TX, kProlTY = MPS_PROL_TY, kProlKZ = 16;

template <typename T>
void launch_restrict_fw(mps_ctx *c, std::vector<LevelBuf<T>> &L, int l, const T *src = nullptr)
{
    const Geom &gc = c->geom[l + 1];
    Lev<T> vf = view(c, L, l), vc = view(c, L, l + 1);
    AlignedW W{L[l + 1].w[0], L[l + 1].w[1], L[l + 1].w[2]};
    const int kz = z_chunk(gc, kRestCX, kRestCY, 2048, 16);
    dim3 gr((unsigned)((gc.nx + kRestCX - 1) / kRestCX), (unsigned)((gc.ny + kRestCY - 1) / kRestCY),
            (unsigned)((gc.nz + kz - 1) / kz));
    hipLaunchKernelGGL((k_restrict_fw_planes<T, kRestCX, kRestCY>), gr, dim3(kRestCX * kRestCY), 0, c->stream, vf,
                       vc, W, src ? src : (const T *)L[l].r, L[l + 1].b, kz);
}

// Residual and restriction in one pass (k_resrestrict_planes): aligned
// hierarchy, rows without convection, fields addressed with 32-bit plane
// offsets. MPS_FUSED_RESTRICT selects where: bit 1 FP32 face-factor rows,
// bit 2 FP64 face-factor rows, bit 4 explicit rows.
#ifndef MPS_FUSED_RESTRICT
#define MPS_FUSED_RESTRICT 1
#endif
#ifndef MPS_RR_CY
#define MPS_RR_CY 8
#endif
#ifndef MPS_RR_TARGET
#define MPS_RR_TARGET 2048
#endif
#ifndef MPS_RR_KZ
#define MPS_RR_KZ 16
#endif
#ifndef MPS_RR_M
#define MPS_RR_M 1
#endif
constexpr int kRRCX = 32, kRRCY = MPS_RR_CY, kRRM = MPS_RR_M;

template <typename T>
bool resrestrict_ok(const mps_ctx *c, const std::vector<LevelBuf<T>> &L, int l, int op)
{
    if (!c->fused_rr || c->opts.hierarchy != MPS_HIER_ALIGNED || c->has_vel) return false;
    if (c->geom[l].ng() >= (1LL << 31)) return false;
    if (op == OP_STENCIL) return (MPS_FUSED_RESTRICT & 4) != 0;
    return view(c, L, l).fast && (MPS_FUSED_RESTRICT & (sizeof(T) == 4 ? 1 : 2)) != 0;
}

template <typename T>
void launch_resrestrict(mps_ctx *c, std::vector<LevelBuf<T>> &L, int l, int op)
{
    const Geom &gc = c->geom[l + 1];
    Lev<T> vf = view(c, L, l), vc = view(c, L, l + 1);
    AlignedW W{L[l + 1].w[0], L[l + 1].w[1], L[l + 1].w[2]};
    const int kz = z_chunk(gc, kRRCX, kRRCY, MPS_RR_TARGET, MPS_RR_KZ);
    dim3 gr((unsigned)((gc.nx + kRRCX - 1) / kRRCX), (unsigned)((gc.ny + kRRCY - 1) / kRRCY),
            (unsigned)((gc.nz + kz - 1) / kz));
    auto go = [&](auto O, auto R) {
        hipLaunchKernelGGL((k_resrestrict_planes<T, decltype(O)::value, decltype(R)::value, kRRCX, kRRCY, kRRM>),
                           gr, dim3(kRRCX * kRRCY * kRRM), 0, c->stream, vf, vc, W, (const T *)L[l].u,
                           (const T *)L[l].b, L[l + 1].b, kz);
    };
    if (op == OP_STENCIL) {
        go(IntC<OP_STENCIL>{}, BoolC<false>{});
    } else if (op == OP_REF) {
        if (c->has_reac) go(IntC<OP_REF>{}, BoolC<true>{});
        else go(IntC<OP_REF>{}, BoolC<false>{});
    } else {
        if (c->has_reac) go(IntC<OP_ALIGNED>{}, BoolC<true>{});
        else go(IntC<OP_ALIGNED>{}, BoolC<false>{});
    }
}

template <typename T>
void launch_prolong_tri(mps_ctx *c, std::vector<LevelBuf<T>> &L, int l)
{
    const Geom &g = c->geom[l];
    Lev<T> vf = view(c, L, l), vc = view(c, L, l + 1);
    // kz even and at most the kernel's chunk (kProlKZ planes)
    const int kz = std::max(2, z_chunk(g, kProlTX, kProlTY, 4096, kProlKZ) & ~1);
    dim3 gr((unsigned)((g.nx + kProlTX - 1) / kProlTX), (unsigned)((g.ny + kProlTY - 1) / kProlTY),
            (unsigned)((g.nz + kz - 1) / kz));
    hipLaunchKernelGGL((k_prolong_tri_planes<T, kProlTX, kProlTY, kProlKZ>), gr, dim3(kProlTX * kProlTY), 0,
                       c->stream, vf, vc, L[l].u, (const T *)L[l + 1].u, kz);
}

// most block partial sums any residual launch on level 0 writes
int residual_max_partials(const Geom &g)
{
    int kz;
    dim3 a = residual_planes_grid(g, kz), b = red_grid(g);
    return (int)std::max(a.x * a.y * a.z, b.x * b.y);
}

// r = b - A u on level l (r nullable) and, with partials, one partial sum of
// r^2 per block; returns the number of partials written. Rows without
// convection in fast arithmetic take the plane-marching kernel.
template <typename T, typename R_>
int launch_residual(mps_ctx *c, std::vector<LevelBuf<T>> &L, int l, const T *u, const T *b, R_ *r, double *partials,
                    int op, bool reduce_grid = false, const double *rscale = nullptr)
{
    Lev<T> v = view(c, L, l);
    const Geom &g = c->geom[l];
    KtScope kt(c, l == 0 ? 1 : -2, sizeof(T) == 4);
    if (v.fast && op != OP_STENCIL && !c->has_vel) {
        int kz;
        const dim3 gr = residual_planes_grid(g, kz);
        auto go = [&](auto O, auto Q) {
            hipLaunchKernelGGL(
                (k_residual_planes<T, R_, decltype(O)::value, decltype(Q)::value, kResTX, kResTY, kResRB>),
                               gr, dim3(kResTX * kResTY), 0, c->stream, v, u, b, r, partials, kz, rscale);
        };
        if (op == OP_REF) {
            if (c->has_reac) go(IntC<OP_REF>{}, BoolC<true>{});
            else go(IntC<OP_REF>{}, BoolC<false>{});
        } else {
            if (c->has_reac) go(IntC<OP_ALIGNED>{}, BoolC<true>{});
            else go(IntC<OP_ALIGNED>{}, BoolC<false>{});
        }
        return partials ? (int)(gr.x * gr.y * gr.z) : 0;
    }
    const dim3 gr = (reduce_grid || partials) ? red_grid(g) : cells_grid(g);
    with_op(op, c->has_vel, c->has_reac, [&](auto O, auto V, auto R) {
        hipLaunchKernelGGL((k_residual<T, R_, decltype(O)::value, decltype(V)::value, decltype(R)::value>), gr,
                           block(), 0, c->stream, v, u, b, r, partials, rscale);
    });
    return partials ? (int)(gr.x * gr.y) : 0;
}

// sum of squares of the residual of level 0 -> c->d_scalar (device)
template <typename T>
void residual_sumsq(mps_ctx *c, std::vector<LevelBuf<T>> &L, const T *u, const T *b, int op, T *r_out = nullptr)
{
    const int np = launch_residual<T, T>(c, L, 0, u, b, r_out, c->partials, op, true);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), block(), 0, c->stream, (const double *)c->partials, np,
                       c->d_scalar);
}

double read_scalar(mps_ctx *c)
{
    HIPCHK(hipMemcpyAsync(c->h_scalar, c->d_scalar, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return c->h_scalar[0];
}

template <typename T>
T *bottom_work(mps_ctx *c);
template <>
double *bottom_work<double>(mps_ctx *c) { return c->bottom64; }
template <>
float *bottom_work<float>(mps_ctx *c) { return c->bottom32; }

template <typename T>
void launch_bottom(mps_ctx *c, std::vector<LevelBuf<T>> &L, int l, int op, T *work)
{
    Lev<T> v = view(c, L, l);
    if (L[l].inv) {  // dense inverse (aligned hierarchy)
        hipLaunchKernelGGL(k_bottom_dense<T>, dim3(1), block(), 0, c->stream, v, L[l].u, (const T *)L[l].b, L[l].inv);
        return;
    }
    // err_tol = ERRTOLR*0.0001 with a default-real literal (mgridsteps.f90:305)
    const T tol = (T)(1e-11 * (double)0.0001f);
    const int m = (int)c->geom[l].n();
    const dim3 bb = block();
    with_op(op, c->has_vel, c->has_reac, [&](auto O, auto V, auto R) {
        if (m <= kBottomLdsM)
            hipLaunchKernelGGL((k_bottom<T, decltype(O)::value, decltype(V)::value, decltype(R)::value, true>),
                               dim3(1), bb, 0, c->stream, v, L[l].u, (const T *)L[l].b, work, m, 1, tol,
                               c->d_status);
        else
            hipLaunchKernelGGL((k_bottom<T, decltype(O)::value, decltype(V)::value, decltype(R)::value, false>),
                               dim3(1), block(), 0, c->stream, v, L[l].u, (const T *)L[l].b, work, m, 1, tol,
                               c->d_status);
    });
}

// fused red-black sweep u -> u2 (see mps3d_sweep.h)
#ifndef MPS_SWEEP_TX
#define MPS_SWEEP_TX 32
#endif
#ifndef MPS_SWEEP_TY
#define MPS_SWEEP_TY 8
#endif
// (the most planes per z-chunk; k_sweep_fast splits nz into ceil(nz / KZ)
// even chunks: at 257^3, 15 chunks of 17-18 planes, 3840 blocks, three
// rounds of the five blocks per CU that the FP32 sweep's LDS and VGPRs allow)
#ifndef MPS_SWEEP_KZ
#define MPS_SWEEP_KZ 18
#endif
constexpr int kSweepTX = MPS_SWEEP_TX, kSweepTY = MPS_SWEEP_TY, kSweepKZ = MPS_SWEEP_KZ;

// per-face multipliers of the staged ghost coefficient (k_sweep_ref,
// k_sweep2_fast): -1 on a no-flux face, g/h - 1 on a DIRC face, so that the
// face weight (D_c + D_ghost) h is zero or the wall term D_c g
template <typename T>
GhostD<T> ghost_multipliers(const mps_ctx *c, const Lev<T> &v)
{
    const double gf[6] = {v.gx, v.gx, v.gy, v.gy, v.gz, v.gz}, hf[6] = {v.hx, v.hx, v.hy, v.hy, v.hz, v.hz};
    GhostD<T> G{};
    for (int f = 0; f < 6; ++f) G.m[f] = c->bc[f] == MPS_BC_DIRC ? (T)(gf[f] / hf[f] - 1.0) : T(-1);
    return G;
}

// zero_in: sweep from a zero guess, uin unread (fused fast kernels only)
template <typename T>
void launch_sweep(mps_ctx *c, std::vector<LevelBuf<T>> &L, int l, const T *uin, T *uout, int op, bool zero_in = false)
{
    Lev<T> v = view(c, L, l);
    const Geom &g = c->geom[l];
    // (the zero-guess sweeps, which read no u, are not timed)
    KtScope kt(c, l == 0 && !zero_in ? 0 : -2, sizeof(T) == 4);
// (blocks a z-chunk length aims at: with 1536 the 129^3 level runs 1216
// blocks, one round of the resident blocks, where 2048 gave 1664)
#ifndef MPS_SWEEP_TARGET
#define MPS_SWEEP_TARGET 1536
#endif
    const int kz = z_chunk(g, kSweepTX, kSweepTY, MPS_SWEEP_TARGET, kSweepKZ);
    dim3 gr((unsigned)((g.nx + kSweepTX - 1) / kSweepTX), (unsigned)((g.ny + kSweepTY - 1) / kSweepTY),
            (unsigned)((g.nz + kz - 1) / kz));
    if (v.fast && op == OP_REF && !c->has_reac && c->sweep_ref && v.omega == T(1) &&
        g.ng() * (long long)sizeof(T) < (1LL << 31)) {
        // uniform face factors: boundary rows through staged ghost coefficients
        GhostD<T> G = ghost_multipliers(c, v);
        if (zero_in)
            hipLaunchKernelGGL((k_sweep_ref<T, kSweepTX, kSweepTY, true>), gr, dim3(kSweepTX * kSweepTY), 0,
                               c->stream, v, G, uin, uout, (const T *)L[l].b, kz);
        else
            hipLaunchKernelGGL((k_sweep_ref<T, kSweepTX, kSweepTY, false>), gr, dim3(kSweepTX * kSweepTY), 0,
                               c->stream, v, G, uin, uout, (const T *)L[l].b, kz);
        return;
    }
    // z ghost-plane multipliers of the fast kernel's OP_REF rows
    const GhostD<T> G = ghost_multipliers(c, v);
    const T zg0 = G.m[BACK], zg1 = G.m[FRONT];
    // k_sweep_fast: on k TX + 1 columns (rows) k tiles, the last extended
    // by the remaining one
    auto tiles = [](int n, int t) { return (n > t && n % t == 1) ? n / t : (n + t - 1) / t; };
    const dim3 grf((unsigned)tiles(g.nx, kSweepTX), (unsigned)tiles(g.ny, kSweepTY), gr.z);
    auto go = [&](auto O, auto R) {
        if (v.fast && zero_in)
            hipLaunchKernelGGL(
                (k_sweep_fast<T, decltype(O)::value, decltype(R)::value, kSweepTX, kSweepTY, true>), grf,
                dim3(kSweepTX * kSweepTY), 0, c->stream, v, uin, uout, (const T *)L[l].b, kz, zg0, zg1);
        else if (v.fast)
            hipLaunchKernelGGL(
                (k_sweep_fast<T, decltype(O)::value, decltype(R)::value, kSweepTX, kSweepTY, false>), grf,
                dim3(kSweepTX * kSweepTY), 0, c->stream, v, uin, uout, (const T *)L[l].b, kz, zg0, zg1);
        else
            hipLaunchKernelGGL((k_sweep<T, decltype(O)::value, decltype(R)::value, kSweepTX, kSweepTY>), gr,
                               dim3(kSweepTX * kSweepTY), 0, c->stream, v, uin, uout, (const T *)L[l].b, kz);
    };
    if (op == OP_STENCIL && L[l].fc[1]) {
        // the same rows from their face form (conductance coarse levels)
        if (zero_in)
            hipLaunchKernelGGL((k_sweep_fast<T, OP_FACES, false, kSweepTX, kSweepTY, true>), grf,
                               dim3(kSweepTX * kSweepTY), 0, c->stream, v, uin, uout, (const T *)L[l].b, kz, zg0, zg1);
        else
            hipLaunchKernelGGL((k_sweep_fast<T, OP_FACES, false, kSweepTX, kSweepTY, false>), grf,
                               dim3(kSweepTX * kSweepTY), 0, c->stream, v, uin, uout, (const T *)L[l].b, kz, zg0, zg1);
    } else if (op == OP_STENCIL) {
        // explicit rows carry their own reaction and time terms; the fast
        // kernel's arithmetic for them is the per-level kernels' in either
        // arithmetic mode
        if (zero_in)
            hipLaunchKernelGGL((k_sweep_fast<T, OP_STENCIL, false, kSweepTX, kSweepTY, true>), grf,
                               dim3(kSweepTX * kSweepTY), 0, c->stream, v, uin, uout, (const T *)L[l].b, kz, zg0, zg1);
        else
            hipLaunchKernelGGL((k_sweep_fast<T, OP_STENCIL, false, kSweepTX, kSweepTY, false>), grf,
                               dim3(kSweepTX * kSweepTY), 0, c->stream, v, uin, uout, (const T *)L[l].b, kz, zg0, zg1);
    } else if (op == OP_REF) {
        if (c->has_reac) go(IntC<OP_REF>{}, BoolC<true>{});
        else go(IntC<OP_REF>{}, BoolC<false>{});
    } else {
        if (c->has_reac) go(IntC<OP_ALIGNED>{}, BoolC<true>{});
        else go(IntC<OP_ALIGNED>{}, BoolC<false>{});
    }
}

template <typename T>
bool fused_ok(const mps_ctx *c, const std::vector<LevelBuf<T>> &L, int l, int op)
{
    // (the fused kernels address planes with 32-bit offsets, and the fast
    // sweep's buffer resources span a z-chunk and two planes either side in
    // fewer than 2^31 bytes). Explicit rows
    // -- conductance coarse rows, Scharfetter-Gummel rows with or without
    // drift, transposed rows -- carry everything a Gauss-Seidel update
    // needs; rows formed from coefficients are fused only without
    // convection (the fused kernels have no upwind terms)
    const Geom &g = c->geom[l];
    const long long slab = (long long)(kSweepKZ + 4) * (g.nx + 2) * (g.ny + 2) * (long long)sizeof(T);
    return (op == OP_STENCIL || !c->has_vel) && L[l].u2 && g.ng() < (1LL << 31) && slab < (1LL << 31);
}

// Two sweeps in one pass (k_sweep2_fast): FP32 fast rows from one
// coefficient field with uniform face factors (OP_REF: the finest level of
// either hierarchy), no convection or reaction term, fields below 2^31 bytes.
#ifndef MPS_SWEEP2_TX
#define MPS_SWEEP2_TX 32
#endif
#ifndef MPS_SWEEP2_TY
#define MPS_SWEEP2_TY 8
#endif
#ifndef MPS_SWEEP2_KZ
#define MPS_SWEEP2_KZ 48
#endif
#ifndef MPS_SWEEP2_TARGET
#define MPS_SWEEP2_TARGET 1536
#endif
#ifndef MPS_SWEEP2_PREC
#define MPS_SWEEP2_PREC 0   // bit 1: FP32, bit 2: FP64
#endif
constexpr int kSw2TX = MPS_SWEEP2_TX, kSw2TY = MPS_SWEEP2_TY;

template <typename T>
bool sweep2_ok(const mps_ctx *c, const std::vector<LevelBuf<T>> &L, int l, int op)
{
    if (!(c->sweep2 & (sizeof(T) == 4 ? 1 : 2)) || op != OP_REF || c->has_reac || c->has_vel || !L[l].u2)
        return false;
    if (c->geom[l].ng() * (long long)sizeof(T) >= (1LL << 31)) return false;
    // (the pair kernel sweeps with plain Gauss-Seidel only)
    const Lev<T> v = view(c, L, l);
    return v.fast != 0 && v.omega == T(1);
}

template <typename T>
void launch_sweep2(mps_ctx *c, std::vector<LevelBuf<T>> &L, int l, const T *uin, T *uout, bool zero_in)
{
    Lev<T> v = view(c, L, l);
    const Geom &g = c->geom[l];
    GhostD<T> G = ghost_multipliers(c, v);
    const int kz = z_chunk(g, kSw2TX, kSw2TY, MPS_SWEEP2_TARGET, MPS_SWEEP2_KZ);
    dim3 gr((unsigned)((g.nx + kSw2TX - 1) / kSw2TX), (unsigned)((g.ny + kSw2TY - 1) / kSw2TY),
            (unsigned)((g.nz + kz - 1) / kz));
    if (zero_in)
        hipLaunchKernelGGL((k_sweep2_fast<T, kSw2TX, kSw2TY, true>), gr, dim3(kSweep2Threads), 0, c->stream, v, G,
                           uin, uout, (const T *)L[l].b, kz);
    else
        hipLaunchKernelGGL((k_sweep2_fast<T, kSw2TX, kSw2TY, false>), gr, dim3(kSweep2Threads), 0, c->stream, v, G,
                           uin, uout, (const T *)L[l].b, kz);
}

// nu red-black sweeps on level l: pairs of fused sweeps u -> u2 -> u where
// the fused kernels apply, in-place colour passes otherwise. swap_odd (the
// cycle's pre- and post-smoothing counts add up to an even number): an odd
// last sweep is a fused sweep u -> u2 after which the level's u and u2
// swap roles, swapped back by the other odd smoothing of the same cycle.
template <typename T>
void smooth_level(mps_ctx *c, std::vector<LevelBuf<T>> &L, int l, int nu, int op, bool zero_first = false,
                  bool swap_odd = false)
{
    int s = 0;
    // pairs in one pass swap the level's u and u2: used when the cycle's
    // pre- and post-smoothing pairs add up to an even number of swaps
    if (sweep2_ok(c, L, l, op) && (c->opts.nu1 / 2 + c->opts.nu2 / 2) % 2 == 0) {
        for (; s + 2 <= nu; s += 2) {
            launch_sweep2<T>(c, L, l, L[l].u, L[l].u2, zero_first && s == 0);
            std::swap(L[l].u, L[l].u2);
        }
    }
    if (fused_ok(c, L, l, op)) {
        for (; s + 2 <= nu; s += 2) {
            launch_sweep<T>(c, L, l, L[l].u, L[l].u2, op, zero_first && s == 0);
            launch_sweep<T>(c, L, l, L[l].u2, L[l].u, op);
        }
        if (s < nu && swap_odd) {
            launch_sweep<T>(c, L, l, L[l].u, L[l].u2, op, zero_first && s == 0);
            std::swap(L[l].u, L[l].u2);
            ++s;
        }
    }
    const Geom &gs = c->geom[l];
    if (s < nu && gs.n() <= kSmallSmooth) {
        // the remaining sweeps of a small level in one launch
        Lev<T> v = view(c, L, l);
        with_op(op, c->has_vel, c->has_reac, [&](auto O, auto V, auto R) {
            hipLaunchKernelGGL((k_colour_small<T, decltype(O)::value, decltype(V)::value, decltype(R)::value>),
                               dim3(1), dim3(kSmallBlock), 0, c->stream, v, L[l].u, (const T *)L[l].b, nu - s);
        });
        s = nu;
    }
    for (; s < nu; ++s) {
        launch_colour<T>(c, L, l, 1, op);
        launch_colour<T>(c, L, l, 0, op);
    }
}

// A level whose correction starts at zero can skip zero-filling u when its
// first pre-smoothing sweep is a fused fast sweep from a zero guess (the
// ping-pong buffers' ghost layers are zeroed once at allocation and only
// interiors are written afterwards).
template <typename T>
bool zero_first_ok(const mps_ctx *c, const std::vector<LevelBuf<T>> &L, int l, int op)
{
    const bool first_fused = c->opts.nu1 >= 2 || (c->opts.nu1 == 1 && c->opts.nu2 % 2 == 1);
    return first_fused && fused_ok(c, L, l, op) && (op == OP_STENCIL || view(c, L, l).fast);
}

// ---------------------------------------------- LDS-resident small levels
// k_lds_cycle takes over the V-cycle from the first level whose levels down
// to the coarsest fit kLdsBudget bytes of LDS: aligned hierarchy with
// explicit-stencil (conductance) coarse rows and a dense coarsest solve.
constexpr long long kLdsBudget = 64 * 1024;

template <typename T>
int lds_cycle_from(const mps_ctx *c, const std::vector<LevelBuf<T>> &L)
{
    if (c->opts.hierarchy != MPS_HIER_ALIGNED || !L.back().inv) return -1;
    long long words = 0;
    int from = -1;
    for (int l = c->nlev() - 1; l >= 1; --l) {
        if (level_op(c, L, l) != OP_STENCIL) break;
        const Geom &g = c->geom[l];
        words += lds_level_words(g.nx, g.ny, g.nz);
        if (words * (long long)sizeof(T) > kLdsBudget || c->nlev() - l > kLdsMaxLev) break;
        from = l;
    }
    return (from >= 0 && from < c->nlev() - 1) ? from : -1;
}

template <typename T>
void launch_lds_cycle(mps_ctx *c, std::vector<LevelBuf<T>> &L, int l)
{
    LdsLevels<T> S{};
    S.nlev = c->nlev() - l;
    long long words = 0;
    for (int q = 0; q < S.nlev; ++q) {
        const Geom &g = c->geom[l + q];
        S.nx[q] = g.nx, S.ny[q] = g.ny, S.nz[q] = g.nz;
        for (int s = 0; s < ST_N; ++s) S.st[q][s] = L[l + q].st[s];
        S.W[q] = AlignedW{L[l + q].w[0], L[l + q].w[1], L[l + q].w[2]};
        words += lds_level_words(g.nx, g.ny, g.nz);
    }
    const Geom &g0 = c->geom[l];
    S.b0 = L[l].b;
    S.u0 = L[l].u;
    S.sy0 = g0.nx + 2;
    S.sz0 = (long long)(g0.nx + 2) * (g0.ny + 2);
    S.inv = L.back().inv;
    S.nu1 = c->opts.nu1;
    S.nu2 = c->opts.nu2;
    S.fast = view(c, L, l).fast;
    hipLaunchKernelGGL(k_lds_cycle<T>, dim3(1), dim3(kLdsBlock), (size_t)(words * sizeof(T)), c->stream, S);
}

// --------------------------------------------------------------- V-cycle
// One V-cycle from level l down (perform_vcycle, mgridsteps.f90:384-579).
// Level l's b must hold its right-hand side.
template <typename T>
// zero_u: level l's correction starts at zero and L[l].u is not yet
// zeroed (the correction scheme of the aligned hierarchy): the level zeroes
// it, or starts from a zero-guess sweep, or solves it outright.
void vcycle(mps_ctx *c, std::vector<LevelBuf<T>> &L, int l, T *bwork, bool homog, bool zero_u = false)
{
    const Geom &g = c->geom[l];
    const int op = level_op(c, L, l);
    if (g.bottom) {
        if (zero_u && !L[l].inv) fill<T>(c, L[l].u, g.ng(), T(0));  // the iterative bottom solve reads u
        launch_bottom<T>(c, L, l, op, bwork);
        return;
    }
    if (l > 0 && l == lds_cycle_from<T>(c, L)) {
        launch_lds_cycle<T>(c, L, l);
        return;
    }
    bool zg = false;
    if (zero_u) {
        zg = zero_first_ok<T>(c, L, l, op);
        if (!zg) fill<T>(c, L[l].u, g.ng(), T(0));
    }
    // odd smoothing counts on both sides: one swap each, so the level's
    // buffers are back in place when the cycle returns
    const bool swap_odd = c->opts.nu1 % 2 == 1 && c->opts.nu2 % 2 == 1;
    smooth_level<T>(c, L, l, c->opts.nu1, op, zg, swap_odd);
    const bool fused = resrestrict_ok<T>(c, L, l, op);
    if (!fused) launch_residual<T, T>(c, L, l, L[l].u, L[l].b, L[l].r, nullptr, op);
    const Geom &gc = c->geom[l + 1];
    Lev<T> vf = view(c, L, l), vc = view(c, L, l + 1);
    const int opc = level_op(c, L, l + 1);
    if (c->opts.hierarchy == MPS_HIER_REF) {
        {
            KtScope kt(c, l == 0 ? 2 : -2, sizeof(T) == 4);
            hipLaunchKernelGGL(k_restrict_fas<T>, cells_grid(gc), block(), 0, c->stream, vf, vc, (const T *)L[l].u,
                               (const T *)L[l].r, L[l + 1].uold, L[l + 1].u, L[l + 1].b);
        }
        with_op(opc, c->has_vel, c->has_reac, [&](auto O, auto V, auto R) {
            hipLaunchKernelGGL((k_tau<T, decltype(O)::value, decltype(V)::value, decltype(R)::value>),
                               cells_grid(gc), block(), 0, c->stream, vc, (const T *)L[l + 1].uold, L[l + 1].b);
        });
        vcycle<T>(c, L, l + 1, bwork, homog);
        FacesT<T> G{};
        for (int f = 0; f < 6; ++f) G.f[f] = homog ? nullptr : L[l + 1].faces[f];
        KtScope kt(c, l == 0 ? 3 : -2, sizeof(T) == 4);
        hipLaunchKernelGGL(k_prolong1_correct<T>, cells_grid(gc), block(), 0, c->stream, vf, vc, L[l].u,
                           (const T *)L[l + 1].u, (const T *)L[l + 1].uold, G);
    } else {
        {
            KtScope kt(c, l == 0 ? 2 : -2, sizeof(T) == 4);
            if (fused) launch_resrestrict<T>(c, L, l, op);
            else launch_restrict_fw<T>(c, L, l);
        }
        vcycle<T>(c, L, l + 1, bwork, homog, true);
        KtScope kt(c, l == 0 ? 3 : -2, sizeof(T) == 4);
        launch_prolong_tri<T>(c, L, l);
    }
    smooth_level<T>(c, L, l, c->opts.nu2, op, false, swap_odd);
}

// Full multigrid for A e = L[0].b on the aligned hierarchy (homogeneous
// rows, correction form): the right-hand side is restricted to every level
// down to the first one the V-cycle solves as a whole (the LDS-resident
// levels, else the coarsest), that level is solved by one V-cycle from zero,
// and on the way up each level starts from the trilinear interpolation of
// the coarser solution and runs one V-cycle. e ends in L[0].u. A level's
// right-hand side is still intact when its turn comes: the V-cycles of the
// coarser levels overwrite only the right-hand sides below them.
// The coarse levels' cycles of the full-multigrid pass smooth once before
// and once after (their errors are corrected again by the finest level's
// cycle): measured at 257^3, the pass reaches the same defect (4.7e-5
// relative) as with the context's two sweeps, 0.16 ms sooner.
#ifndef MPS_FMG_NU1C
#define MPS_FMG_NU1C 1   // pre-/post-smoothing of the coarse levels' cycles (-1: the context's)
#endif
#ifndef MPS_FMG_NU2C
#define MPS_FMG_NU2C 1
#endif
#ifndef MPS_FMG_NU10
#define MPS_FMG_NU10 -1  // and of the finest level's cycle
#endif
#ifndef MPS_FMG_NU20
#define MPS_FMG_NU20 -1
#endif
#ifndef MPS_FMG_CYCC
#define MPS_FMG_CYCC 1   // V-cycles per coarse level
#endif
template <typename T>
void fmg(mps_ctx *c, std::vector<LevelBuf<T>> &L, T *bwork)
{
    const int nl = c->nlev();
    int top = lds_cycle_from<T>(c, L);
    if (top < 0) top = nl - 1;
    for (int l = 0; l < top; ++l) launch_restrict_fw<T>(c, L, l, L[l].b);
    vcycle<T>(c, L, top, bwork, true, true);
    const int nu1 = c->opts.nu1, nu2 = c->opts.nu2;
    for (int l = top - 1; l >= 0; --l) {
        fill<T>(c, L[l].u, c->geom[l].ng(), T(0));
        launch_prolong_tri<T>(c, L, l);
        const int a = l ? MPS_FMG_NU1C : MPS_FMG_NU10, b = l ? MPS_FMG_NU2C : MPS_FMG_NU20;
        c->opts.nu1 = a >= 0 ? a : nu1;
        c->opts.nu2 = b >= 0 ? b : nu2;
        for (int q = 0; q < (l ? MPS_FMG_CYCC : 1); ++q) vcycle<T>(c, L, l, bwork, true, false);
        c->opts.nu1 = nu1;
        c->opts.nu2 = nu2;
    }
}

// ---------------------------------------------------------- coefficients
// coarse coefficients from the finer level (REF: 8-child mean as the
// reference does every cycle; ALIGNED: volume-weighted full weighting)
// (the coarse arrays' ghost layers are zeroed when they are allocated and
// never written: the restrictions write interiors only)
void derive_coarse(mps_ctx *c)
{
    auto &L = c->L64;
    for (int l = 1; l < c->nlev(); ++l) {
        const Geom &gc = c->geom[l];
        Lev<double> vf = view(c, L, l - 1), vc = view(c, L, l);
        FieldSet5 F{{L[l - 1].D, L[l - 1].vx, L[l - 1].vy, L[l - 1].vz, L[l - 1].reac},
                    {L[l].D, L[l].vx, L[l].vy, L[l].vz, L[l].reac}};
        if (c->opts.hierarchy == MPS_HIER_REF) {
            hipLaunchKernelGGL(k_restrict_mean5, cells_grid(gc), block(), 0, c->stream, vf, vc, F);
        } else {
            AlignedW W{L[l].w[0], L[l].w[1], L[l].w[2]};
            for (int q = 0; q < 5; ++q) {
                if (!F.src[q]) continue;
                hipLaunchKernelGGL(k_restrict_fw<double>, cells_grid(gc), block(), 0, c->stream, vf, vc, W,
                                   (const double *)F.src[q], F.dst[q]);
            }
        }
    }
}

#ifndef MPS_CONDUCTANCE
#define MPS_CONDUCTANCE 1
#endif
// Aligned hierarchy, rows without convection: the coarse levels take
// 7-point rows built from conductance-coarsened face coefficients
// (k_face_means, k_coarsen_faces, k_faces_stencil) instead of rows from
// restricted cell coefficients; other setups keep their level operators.
void conductance_coarse_levels(mps_ctx *c)
{
    auto &L = c->L64;
    const bool on = MPS_CONDUCTANCE && c->mg_ok && c->opts.hierarchy == MPS_HIER_ALIGNED &&
                    c->opts.flux != MPS_FLUX_SG && !c->has_vel && c->nlev() > 1;
    for (int l = 1; l < c->nlev(); ++l) {
        L[l].op = on ? OP_STENCIL : -1;
        if (l < (int)c->L32.size()) c->L32[l].op = L[l].op;
    }
    // the face form exists only while the levels hold these rows
    for (int l = 0; l < c->nlev(); ++l) {
        for (int q = 1; q < 4; ++q) {
            if (!on || l == 0) {
                dfree(L[l].fc[q]);
                L[l].fc[q] = nullptr;
            }
        }
        if (!on || l == 0) L[l].fc[0] = nullptr;
    }
    if (!on) return;
    std::vector<double *> f(3 * (size_t)c->nlev(), nullptr);
    try {
        for (int l = 0; l < c->nlev(); ++l) {
            for (int d = 0; d < 3; ++d) {
                f[3 * l + d] = dalloc<double>(c->geom[l].ng());
                fill<double>(c, f[3 * l + d], c->geom[l].ng(), 0.0);
            }
        }
        hipLaunchKernelGGL(k_face_means, cells_grid(c->geom[0]), block(), 0, c->stream, view(c, L, 0), f[0], f[1],
                           f[2]);
        for (int l = 1; l < c->nlev(); ++l) {
            const Geom &g = c->geom[l];
            AlignedW W{L[l].w[0], L[l].w[1], L[l].w[2]};
            hipLaunchKernelGGL(k_coarsen_faces, cells_grid(g), block(), 0, c->stream, view(c, L, l - 1), view(c, L, l),
                               W, (const double *)f[3 * l - 3], (const double *)f[3 * l - 2],
                               (const double *)f[3 * l - 1], f[3 * l], f[3 * l + 1], f[3 * l + 2]);
            LevelBuf<double> &B = L[l];
            for (int q = 0; q < ST_N; ++q) {
                if (!B.st[q]) B.st[q] = dalloc<double>(g.ng());
                fill<double>(c, B.st[q], g.ng(), 0.0);
            }
            // face form: conductances (ghost entries stay 0: the boundary
            // faces) and the diagonal of st
            for (int q = 1; q < 4; ++q) {
                if (!B.fc[q]) B.fc[q] = dalloc<double>(g.ng());
                fill<double>(c, B.fc[q], g.ng(), 0.0);
            }
            B.fc[0] = B.st[ST_C];
            hipLaunchKernelGGL(k_faces_stencil, cells_grid(g), block(), 0, c->stream, view(c, L, l),
                               (const double *)f[3 * l], (const double *)f[3 * l + 1], (const double *)f[3 * l + 2],
                               B.st[0], B.st[1], B.st[2], B.st[3], B.st[4], B.st[5], B.st[6], B.fc[1], B.fc[2],
                               B.fc[3]);
        }
        HIPCHK(hipStreamSynchronize(c->stream));
    } catch (...) {
        for (auto p : f) dfree(p);
        throw;
    }
    for (auto p : f) dfree(p);
}

void build_stencils(mps_ctx *c, std::vector<LevelBuf<double>> &L, bool transpose)
{
    for (int l = 0; l < c->nlev(); ++l) {
        const Geom &g = c->geom[l];
        LevelBuf<double> &B = L[l];
        for (int q = 0; q < ST_N; ++q) {
            if (!B.st[q]) B.st[q] = dalloc<double>(g.ng());
            fill<double>(c, B.st[q], g.ng(), 0.0);
        }
        // coefficient views come from the FP64 forward hierarchy
        Lev<double> v = view(c, c->L64, l);
        const bool aligned = c->opts.hierarchy == MPS_HIER_ALIGNED && l > 0;
        if (c->opts.flux == MPS_FLUX_SG) {
            if (aligned)
                hipLaunchKernelGGL(k_sg_stencil<OP_ALIGNED>, cells_grid(g), block(), 0, c->stream, v, B.st[0],
                                   B.st[1], B.st[2], B.st[3], B.st[4], B.st[5], B.st[6], transpose ? 1 : 0);
            else
                hipLaunchKernelGGL(k_sg_stencil<OP_REF>, cells_grid(g), block(), 0, c->stream, v, B.st[0], B.st[1],
                                   B.st[2], B.st[3], B.st[4], B.st[5], B.st[6], transpose ? 1 : 0);
        } else {
            const int op = aligned ? OP_ALIGNED : OP_REF;
            with_op(op, c->has_vel, c->has_reac, [&](auto O, auto V, auto R) {
                hipLaunchKernelGGL((k_build_stencil<double, decltype(O)::value, decltype(V)::value,
                                                    decltype(R)::value>),
                                   cells_grid(g), block(), 0, c->stream, v, B.st[0], B.st[1], B.st[2], B.st[3],
                                   B.st[4], B.st[5], B.st[6], transpose ? 1 : 0);
            });
        }
    }
}

// FP32 copies of the coefficient (or stencil) arrays for the inner cycles
void build_fp32(mps_ctx *c)
{
    auto &L = c->L64;
    auto &F = c->L32;
    for (int l = 0; l < c->nlev(); ++l) {
        const long long ng = c->geom[l].ng();
        convert<double, float>(c, L[l].D, F[l].D, ng);
        double *src[4] = {L[l].vx, L[l].vy, L[l].vz, L[l].reac};
        float **dst[4] = {&F[l].vx, &F[l].vy, &F[l].vz, &F[l].reac};
        for (int q = 0; q < 4; ++q) {
            if (!src[q]) continue;
            if (!*dst[q]) *dst[q] = dalloc<float>(ng);
            convert<double, float>(c, src[q], *dst[q], ng);
        }
        for (int q = 0; q < ST_N; ++q) {
            if (!L[l].st[q]) continue;
            if (!F[l].st[q]) F[l].st[q] = dalloc<float>(ng);
            convert<double, float>(c, L[l].st[q], F[l].st[q], ng);
        }
        // face form: FP32 conductances, the FP32 diagonal, and the FP32
        // off-diagonal rows rebuilt from them in FP32 (what the face-form
        // sweep forms), so all FP32 kernels of the level apply one operator
        if (L[l].fc[1]) {
            for (int q = 1; q < 4; ++q) {
                if (!F[l].fc[q]) F[l].fc[q] = dalloc<float>(ng);
                convert<double, float>(c, L[l].fc[q], F[l].fc[q], ng);
            }
            F[l].fc[0] = F[l].st[ST_C];
            hipLaunchKernelGGL(k_faces_rows<float>, cells_grid(c->geom[l]), block(), 0, c->stream, view(c, F, l),
                               F[l].st[1], F[l].st[2], F[l].st[3], F[l].st[4], F[l].st[5], F[l].st[6]);
        } else {
            for (int q = 1; q < 4; ++q) {
                dfree(F[l].fc[q]);
                F[l].fc[q] = nullptr;
            }
            F[l].fc[0] = nullptr;
        }
    }
}

// reference-mirroring entry points (find_AX, find_bvec, smoothing) always
// evaluate rows in the reference's arithmetic
struct ExactScope {
    mps_ctx *c;
    explicit ExactScope(mps_ctx *c_) : c(c_) { c->exact_now = true; }
    ~ExactScope() { c->exact_now = false; }
};

// --------------------------------------------------------------- helpers
// Orders the context's stream after the work already queued on the null
// stream (the default stream of frameworks such as torch), so that device
// arrays produced there are complete before this call reads them.
int after_null_stream(mps_ctx *c)
{
    hipError_t e = hipEventRecord(c->ev_in, nullptr);
    if (e == hipSuccess) e = hipStreamWaitEvent(c->stream, c->ev_in, 0);
    if (e != hipSuccess) return fail(MPS_ERR_HIP, "ordering after the null stream: %s", hipGetErrorString(e));
    return MPS_OK;
}

int guard(mps_ctx *c, bool need_coeffs = true)
{
    if (!c) return fail(MPS_ERR_ARG, "null context");
    if (need_coeffs && (!c->coeffs_set || !c->bc_set))
        return fail(MPS_ERR_STATE, "set boundary conditions and coefficients first");
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return fail(MPS_ERR_HIP, "hipSetDevice: %s", hipGetErrorString(e));
    return after_null_stream(c);
}

int guard_mg(mps_ctx *c)
{
    int st = guard(c);
    if (st) return st;
    if (!c->mg_ok) return fail(MPS_ERR_ARG, "no multigrid hierarchy for this grid: %s", c->mg_error.c_str());
    return MPS_OK;
}

// resolves the finest-level solution buffer: a device array is used in
// place; a host array is staged through the internal buffer
double *bind_solution(mps_ctx *c, double *soln, bool &staged)
{
    const long long ng = c->geom[0].ng();
    if (is_device_ptr(soln)) {
        staged = false;
        return soln;
    }
    staged = true;
    HIPCHK(hipMemcpyAsync(c->user_tmp, soln, ng * sizeof(double), hipMemcpyHostToDevice, c->stream));
    return c->user_tmp;
}

void unbind_solution(mps_ctx *c, double *soln, bool staged)
{
    if (!staged) return;
    HIPCHK(hipMemcpyAsync(soln, c->user_tmp, c->geom[0].ng() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
}

// b0 = find_bvec(applybc) + mgridsterm on the finest level
void compute_b0(mps_ctx *c, const double *sterm_dev)
{
    Lev<double> v = view(c, c->L64, 0);
    Faces F{};
    for (int f = 0; f < 6; ++f) F.f[f] = c->faces0[f];
    hipLaunchKernelGGL(k_bvec, cells_grid(c->geom[0]), block(), 0, c->stream, v, F, sterm_dev,
                       (const double *)c->source, 1, c->b0);
}

const double *stage_interior(mps_ctx *c, const double *a)
{
    if (!a) return nullptr;
    if (is_device_ptr(a)) return a;
    HIPCHK(hipMemcpyAsync(c->user_tmp2, a, c->geom[0].n() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    return c->user_tmp2;
}

// one FP64 V-cycle on u (level 0 bound to u) with right-hand side b;
// leaves the sum of squares of the final residual in d_scalar
void fp64_cycle(mps_ctx *c, double *u, double *b, bool homog)
{
    auto &L = c->L64;
    L[0].u = u;
    L[0].own_u = false;
    L[0].b = b;
    vcycle<double>(c, L, 0, c->bottom64, homog);
    if (!c->geom[0].bottom) residual_sumsq<double>(c, L, u, b, level_op(c, 0));
}
void fp64_cycle(mps_ctx *c, double *u) { fp64_cycle(c, u, c->b0, false); }


}  // namespace

// FP64 residual of the current iterate, scaled by rscale[0] (a power of two)
// and rounded to FP32 for the inner cycle
namespace mps {
template <int OP, bool VEL, bool REAC>
__global__ void __launch_bounds__(kBlock) k_defect32(Lev<double> L, const double *__restrict__ u,
                                                    const double *__restrict__ b, float *d32, double *partials,
                                                    const double *rscale)
{
    int i, j;
    double sq = 0.0;
    const double rs = rscale[0];
    if (cell_ij(L, i, j)) {
        for (int k = blockIdx.y + 1; k <= L.nz; k += gridDim.y) {
            const long long c = (long long)i + (long long)L.sy * j + L.sz * k;
            double dg, of;
            row<double, OP, VEL, REAC>(L, u, i, j, k, c, dg, of);
            double rv = b[c] - (dg * u[c] + of);
            d32[c] = (float)(rv * rs);
            sq += rv * rv;
        }
    }
    double s = block_sum(sq);
    if (threadIdx.x == 0) partials[blockIdx.x + (long long)gridDim.x * blockIdx.y] = s;
}
}  // namespace mps

namespace {

void defect_step(mps_ctx *c, double *u, bool do_cycle, const double *b = nullptr)
{
    auto &F = c->L32;
    const Geom &g = c->geom[0];
    Lev<double> v = view(c, c->L64, 0);
    if (!b) b = c->b0;
    int np;
    if (v.fast && level_op(c, 0) != OP_STENCIL && !c->has_vel) {
        np = launch_residual<double, float>(c, c->L64, 0, u, b, F[0].b, c->partials, level_op(c, 0), false,
                                            c->d_dcscale);
    } else {
        with_op(level_op(c, 0), c->has_vel, c->has_reac, [&](auto O, auto V, auto R) {
            hipLaunchKernelGGL((k_defect32<decltype(O)::value, decltype(V)::value, decltype(R)::value>),
                               red_grid(g), block(), 0, c->stream, v, (const double *)u, b, F[0].b, c->partials,
                               (const double *)c->d_dcscale);
        });
        dim3 gr = red_grid(g);
        np = (int)(gr.x * gr.y);
    }
    hipLaunchKernelGGL(k_reduce_final, dim3(1), block(), 0, c->stream, (const double *)c->partials, np,
                       c->d_scalar);
    if (!do_cycle) return;
    vcycle<float>(c, F, 0, c->bottom32, true, true);
    Lev<double> v64 = view(c, c->L64, 0);
    hipLaunchKernelGGL((k_add_interior<double, float>), cells_grid(g), block(), 0, c->stream, v64, u,
                       (const float *)F[0].u, (const double *)c->d_dcscale);
}

double measure(const mps_ctx *c, double sumsq, double bnorm2)
{
    const double N = (double)c->geom[0].n();
    if (c->opts.tol_kind == MPS_TOL_REL_L2) return bnorm2 > 0 ? std::sqrt(sumsq) / std::sqrt(bnorm2) : std::sqrt(sumsq);
    return std::sqrt(sumsq / N);
}

// squared 2-norm of b into d_scalar
void b_sumsq(mps_ctx *c, const double *b)
{
    // residual of u = 0 against b is b itself: reuse the residual kernel
    // with a zero solution held in the level-0 residual buffer
    auto &L = c->L64;
    fill<double>(c, L[0].r, c->geom[0].ng(), 0.0);
    residual_sumsq<double>(c, L, L[0].r, b, level_op(c, 0));
}

// One outer step of solve_core: the FP32 V-cycle on the defect, the FP64
// correction and the next FP64 defect (dc), or one FP64 V-cycle and its
// residual norm. The step's ~10 launches per level are captured once into a
// HIP graph and replayed, which takes the host launch cost off the critical
// path between the per-cycle convergence checks. Every node is a kernel
// (fills included), so replays do not depend on memset nodes.
void outer_step_body(mps_ctx *c, double *u, double *b, bool homog, bool dc, int inner)
{
    if (dc) {
        vcycle<float>(c, c->L32, 0, c->bottom32, true, true);
        // further FP32 cycles on the same defect from the current correction
        // (inner in all): the FP32 cycle's error reduction (~0.08 at 257^3)
        // stays far above FP32 rounding for a few cycles, so each FP64
        // defect carries several
        for (int q = 1; q < inner; ++q) vcycle<float>(c, c->L32, 0, c->bottom32, true, false);
        Lev<double> v64 = view(c, c->L64, 0);
        hipLaunchKernelGGL((k_add_interior<double, float>), cells_grid(c->geom[0]), block(), 0, c->stream, v64, u,
                           (const float *)c->L32[0].u, (const double *)c->d_dcscale);
        defect_step(c, u, false, b);
    } else {
        fp64_cycle(c, u, b, homog);
    }
}

// The first iteration of a solve with MPS_START_FMG: a full-multigrid pass
// on the current defect instead of one V-cycle, then, as outer_step_body,
// the FP64 residual of the corrected iterate into d_scalar. The FP64 form
// needs fmg_buffers first.
void fmg_step_body(mps_ctx *c, double *u, double *b, bool dc)
{
    const Geom &g = c->geom[0];
    Lev<double> v64 = view(c, c->L64, 0);
    if (dc) {
        // the defect (scaled FP32) is in L32[0].b from the previous defect_step
        fmg<float>(c, c->L32, c->bottom32);
        hipLaunchKernelGGL((k_add_interior<double, float>), cells_grid(g), block(), 0, c->stream, v64, u,
                           (const float *)c->L32[0].u, (const double *)c->d_dcscale);
        defect_step(c, u, false, b);
        return;
    }
    auto &L = c->L64;
    launch_residual<double, double>(c, L, 0, u, b, c->fmg_r, nullptr, level_op(c, 0));
    double *ub = L[0].u, *bb = L[0].b;
    L[0].u = c->fmg_e;
    L[0].b = c->fmg_r;
    fmg<double>(c, L, c->bottom64);
    L[0].u = ub;
    L[0].b = bb;
    hipLaunchKernelGGL((k_add_interior<double, double>), cells_grid(g), block(), 0, c->stream, v64, u,
                       (const double *)c->fmg_e, (const double *)nullptr);
    residual_sumsq<double>(c, L, u, b, level_op(c, 0));
}

// the FP64 defect and correction buffers of a full-multigrid start
void fmg_buffers(mps_ctx *c)
{
    const Geom &g = c->geom[0];
    for (double **p : {&c->fmg_r, &c->fmg_e}) {
        if (!*p) {
            *p = dalloc<double>(g.ng());
            fill<double>(c, *p, g.ng(), 0.0);  // ghost layers stay zero: interiors only are written
        }
    }
}

// One outer step (fmg_first: the full-multigrid first step; inner: FP32
// V-cycles of a defect-correction step), replayed from a HIP graph captured
// on first use.
void outer_step(mps_ctx *c, double *u, double *b, bool homog, bool dc, bool fmg_first = false, int inner = 1)
{
    if (fmg_first && !dc) fmg_buffers(c);
    auto body = [&]() {
        if (fmg_first) fmg_step_body(c, u, b, dc);
        else outer_step_body(c, u, b, homog, dc, inner);
    };
    if (!c->graphs) {
        body();
        return;
    }
    StepGraph &G = c->step[fmg_first ? 1 : dc ? 1 + inner : 0];
    const int kind = (dc ? 2 : 0) + (homog ? 1 : 0) + (c->exact_now ? 4 : 0);
    if (!G.exec || G.u != u || G.b != b || G.kind != kind || G.gen != c->gen) {
        if (G.exec) {
            HIPCHK(hipGraphExecDestroy(G.exec));
            G.exec = nullptr;
        }
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        try {
            body();
        } catch (...) {
            (void)hipStreamEndCapture(c->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        HIPCHK(hipStreamEndCapture(c->stream, &g));
        const hipError_t e = hipGraphInstantiate(&G.exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            G.exec = nullptr;
            throw HipError{e, "hipGraphInstantiate"};
        }
        G.u = u;
        G.b = b;
        G.kind = kind;
        G.gen = c->gen;
    }
    HIPCHK(hipGraphLaunch(G.exec, c->stream));
}

// FP32 V-cycles for the next defect-correction step after one of k cycles
// took the measure from prev to now: the fewest that reach tol at the
// step's per-cycle contraction rho = (now / prev)^(1/k), but no more than
// keep the step's reduction above 1e-5 (FP32 rounding bounds what one FP32
// correction of the defect can deliver), within 1..kDcInnerMax; two when
// the step did not contract.
int dc_inner_next(double prev, double now, int k, double tol)
{
    if (!(now > 0.0 && prev > now && tol > 0.0)) return 2;
    const double lr = std::log(now / prev) / k;   // log rho < 0
    if (!(lr < 0.0)) return 2;
    const double need = std::log(tol / now) / lr;  // cycles to reach tol
    const double room = std::log(1e-5) / lr;      // cycles before the FP32 limit
    int q = (int)std::ceil(std::min(need, kDcInnerMax + 1.0));
    q = std::min(q, std::max(1, (int)std::floor(std::min(room, kDcInnerMax + 1.0))));
    return std::max(1, std::min(q, kDcInnerMax));
}

// FP32 V-cycles for the first defect-correction step (no contraction seen
// yet): the fewest that would reach tol at a nominal 0.1 per cycle (a
// multigrid V-cycle's textbook contraction; the steps after it use the
// measured one), within 1..kDcInnerMax.
int dc_inner_first(double now, double tol)
{
    if (!(now > tol && tol > 0.0)) return 2;
    const double need = std::log(tol / now) / std::log(0.1);
    return std::max(1, std::min((int)std::ceil(std::min(need, kDcInnerMax + 1.0)), kDcInnerMax));
}

// The solve loop shared by the forward and the self-adjoint solves: V-cycles
// (FP64) or FP32 defect-correction steps on u for right-hand side b until
// the opts.tol_kind measure <= opts.tol. Returns the cycles run (outer
// steps); c->last_vcycles gets the V-cycles (a full-multigrid pass counts
// as one).
int solve_core(mps_ctx *c, double *u, double *b, bool homog, double *history, double &meas)
{
    struct RelaxScope {
        mps_ctx *c;
        explicit RelaxScope(mps_ctx *c_) : c(c_) { c->relax_now = true; }
        ~RelaxScope() { c->relax_now = false; }
    } relax(c);
    double bn2 = 0.0;
    const bool dc = c->opts.precision == MPS_PREC_FP32_DC;
    if (c->opts.tol_kind == MPS_TOL_REL_L2 || dc) {
        b_sumsq(c, b);
        // the FP32 defects of this solve are scaled by the binade of b's rms
        // (they stay between about 1 and the tolerance times that), so right-
        // hand sides far outside FP32's range, e.g. adjoint ones, still
        // correct in FP32
        if (dc)
            hipLaunchKernelGGL(k_dc_scale, dim3(1), dim3(64), 0, c->stream, (const double *)c->d_scalar,
                               (double)c->geom[0].n(), c->d_dcscale);
        if (c->opts.tol_kind == MPS_TOL_REL_L2) bn2 = read_scalar(c);
    }
    int it = 0;
    meas = 0.0;
    c->last_vcycles = 0;
    if (dc) {
        // defect correction: residual of the current iterate first
        defect_step(c, u, false, b);
        meas = measure(c, read_scalar(c), bn2);
        const bool adapt = c->opts.dc_inner == 0;
        int inner = adapt ? 2 : c->opts.dc_inner;
        bool contracted = false;   // a defect step's contraction has been seen
        if (!(meas <= c->opts.tol)) {
            for (it = 0; it < c->opts.max_cycles;) {
                // the FP32 cycles on the defect, the FP64 correction, then the
                // FP64 residual of the corrected iterate (the next defect)
                const bool fmg_first = it == 0 && c->opts.start == MPS_START_FMG;
                if (adapt && !fmg_first && !contracted) inner = dc_inner_first(meas, c->opts.tol);
                outer_step(c, u, b, homog, true, fmg_first, inner);
                c->last_vcycles += fmg_first ? 1 : inner;
                const double prev = meas;
                meas = measure(c, read_scalar(c), bn2);
                if (history) history[it] = meas;
                ++it;
                if (!std::isfinite(meas)) break;
                if (meas <= c->opts.tol) break;
                if (adapt && !fmg_first) {
                    inner = dc_inner_next(prev, meas, inner, c->opts.tol);
                    contracted = true;
                }
            }
        }
    } else {
        for (it = 0; it < c->opts.max_cycles;) {
            outer_step(c, u, b, homog, false, it == 0 && c->opts.start == MPS_START_FMG && !c->geom[0].bottom);
            ++c->last_vcycles;
            meas = c->geom[0].bottom ? 0.0 : measure(c, read_scalar(c), bn2);
            if (history) history[it] = meas;
            ++it;
            if (!std::isfinite(meas)) break;
            if (meas <= c->opts.tol) break;
        }
    }
    return it;
}

// coarsest-level dense inverse for the aligned hierarchy: assembles the
// coarsest operator from its 7-point rows and inverts it (Gauss-Jordan with
// partial pivoting, FP64) once per set of coefficients
// Reference hierarchy in fast arithmetic: the coarsest operator's inverse
// built on the device (k_dense_inverse) once per set of coefficients; the
// V-cycle's coarse GMRES becomes one matrix-vector product.
void build_bottom_inverse_device(mps_ctx *c)
{
    const int lb = c->nlev() - 1;
    const Geom &g = c->geom[lb];
    const int n = (int)g.n();
    if (n > kBottomMaxDense) return;   // k_bottom stays
    const long long ng = g.ng();
    for (int q = 0; q < ST_N; ++q)
        if (!c->bst[q]) {
            c->bst[q] = dalloc<double>(ng);
            fill<double>(c, c->bst[q], ng, 0.0);
        }
    if (!c->dinv) c->dinv = dalloc<double>((size_t)kBottomMaxDense * kBottomMaxDense);
    if (n > kDenseLdsN && !c->dwork) c->dwork = dalloc<double>((size_t)kBottomMaxDense * kBottomMaxDense);
    Lev<double> v = view(c, c->L64, lb);
    with_op(level_op(c, lb), c->has_vel, c->has_reac, [&](auto O, auto V, auto R) {
        hipLaunchKernelGGL((k_build_stencil<double, decltype(O)::value, decltype(V)::value, decltype(R)::value>),
                           cells_grid(g), block(), 0, c->stream, v, c->bst[0], c->bst[1], c->bst[2], c->bst[3],
                           c->bst[4], c->bst[5], c->bst[6], 0);
    });
#ifndef MPS_DENSE_LDS_THREADS
#define MPS_DENSE_LDS_THREADS 1024
#endif
    const unsigned nth = n <= kDenseLdsN ? MPS_DENSE_LDS_THREADS : kDenseThreads;
    hipLaunchKernelGGL(k_dense_inverse, dim3(1), dim3(nth), 0, c->stream, v, (const double *)c->bst[0],
                       (const double *)c->bst[1], (const double *)c->bst[2], (const double *)c->bst[3],
                       (const double *)c->bst[4], (const double *)c->bst[5], (const double *)c->bst[6], c->dwork,
                       c->dinv, (int *)nullptr);
    HIPCHK(hipGetLastError());
    c->L64[lb].inv = c->dinv;
    if (!c->L32.empty()) c->L32[lb].inv = c->dinv;
}

void build_bottom_inverse(mps_ctx *c)
{
    for (auto &B : c->L64) B.inv = nullptr;
    for (auto &B : c->L32) B.inv = nullptr;
    if (c->mg_ok && c->opts.hierarchy == MPS_HIER_REF && c->opts.arith == MPS_ARITH_FAST) {
        build_bottom_inverse_device(c);
        return;
    }
    if (!c->mg_ok || c->opts.hierarchy != MPS_HIER_ALIGNED) return;
    const int lb = c->nlev() - 1;
    const Geom &g = c->geom[lb];
    const int n = (int)g.n();
    if (n > kBottomMaxDense) throw HipError{hipErrorInvalidValue, "coarsest grid too large for a dense solve"};
    const long long ng = g.ng();
    std::vector<std::vector<double>> st(ST_N, std::vector<double>(ng, 0.0));
    auto &B = c->L64[lb];
    if (level_op(c, c->L64, lb) == OP_STENCIL) {
        for (int q = 0; q < ST_N; ++q)
            HIPCHK(hipMemcpyAsync(st[q].data(), B.st[q], ng * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    } else {
        double *tmp[ST_N];
        for (int q = 0; q < ST_N; ++q) {
            tmp[q] = dalloc<double>(ng);
            fill<double>(c, tmp[q], ng, 0.0);
        }
        Lev<double> v = view(c, c->L64, lb);
        with_op(level_op(c, lb), c->has_vel, c->has_reac, [&](auto O, auto V, auto R) {
            hipLaunchKernelGGL((k_build_stencil<double, decltype(O)::value, decltype(V)::value, decltype(R)::value>),
                               cells_grid(g), block(), 0, c->stream, v, tmp[0], tmp[1], tmp[2], tmp[3], tmp[4], tmp[5],
                               tmp[6], 0);
        });
        for (int q = 0; q < ST_N; ++q)
            HIPCHK(hipMemcpyAsync(st[q].data(), tmp[q], ng * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int q = 0; q < ST_N; ++q) dfree(tmp[q]);
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    std::vector<double> A((size_t)n * n, 0.0), I((size_t)n * n, 0.0);
    auto gi = [&](int i, int j, int k) { return (long long)i + (long long)(nx + 2) * (j + (long long)(ny + 2) * k); };
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                const int r = (i - 1) + nx * ((j - 1) + ny * (k - 1));
                const long long cc = gi(i, j, k);
                A[(size_t)r * n + r] = st[ST_C][cc];
                if (i > 1) A[(size_t)r * n + r - 1] = st[ST_XM][cc];
                if (i < nx) A[(size_t)r * n + r + 1] = st[ST_XP][cc];
                if (j > 1) A[(size_t)r * n + r - nx] = st[ST_YM][cc];
                if (j < ny) A[(size_t)r * n + r + nx] = st[ST_YP][cc];
                if (k > 1) A[(size_t)r * n + r - nx * ny] = st[ST_ZM][cc];
                if (k < nz) A[(size_t)r * n + r + nx * ny] = st[ST_ZP][cc];
            }
    for (int r = 0; r < n; ++r) I[(size_t)r * n + r] = 1.0;
    for (int col = 0; col < n; ++col) {
        int piv = col;
        for (int r = col + 1; r < n; ++r)
            if (std::fabs(A[(size_t)r * n + col]) > std::fabs(A[(size_t)piv * n + col])) piv = r;
        if (!(std::fabs(A[(size_t)piv * n + col]) > 0.0))
            throw HipError{hipErrorInvalidValue, "singular coarsest-level operator"};
        if (piv != col)
            for (int q = 0; q < n; ++q) {
                std::swap(A[(size_t)piv * n + q], A[(size_t)col * n + q]);
                std::swap(I[(size_t)piv * n + q], I[(size_t)col * n + q]);
            }
        const double d = 1.0 / A[(size_t)col * n + col];
        for (int q = 0; q < n; ++q) {
            A[(size_t)col * n + q] *= d;
            I[(size_t)col * n + q] *= d;
        }
        for (int r = 0; r < n; ++r) {
            if (r == col) continue;
            const double f = A[(size_t)r * n + col];
            if (f == 0.0) continue;
            for (int q = 0; q < n; ++q) {
                A[(size_t)r * n + q] -= f * A[(size_t)col * n + q];
                I[(size_t)r * n + q] -= f * I[(size_t)col * n + q];
            }
        }
    }
    std::vector<double> IT((size_t)n * n);
    for (int r = 0; r < n; ++r)
        for (int q = 0; q < n; ++q) IT[(size_t)q * n + r] = I[(size_t)r * n + q];
    if (!c->dinv) c->dinv = dalloc<double>((size_t)kBottomMaxDense * kBottomMaxDense);
    if (!c->dinvT) c->dinvT = dalloc<double>((size_t)kBottomMaxDense * kBottomMaxDense);
    HIPCHK(hipMemcpy(c->dinv, I.data(), (size_t)n * n * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->dinvT, IT.data(), (size_t)n * n * sizeof(double), hipMemcpyHostToDevice));
    c->L64[lb].inv = c->dinv;
    if (!c->L32.empty()) c->L32[lb].inv = c->dinv;
}

void free_gmres(GmresBufs &G);

// a bare level view of an nx x ny x nz ghosted grid (addressing only)
Lev<double> grid_view(int nx, int ny, int nz)
{
    Lev<double> v{};
    v.nx = nx;
    v.ny = ny;
    v.nz = nz;
    v.sy = nx + 2;
    v.sz = (long long)(nx + 2) * (ny + 2);
    return v;
}

}  // namespace

// =================================================================== C ABI
extern "C" {

void mps_default_opts(mps_opts *o)
{
    if (!o) return;
    o->hierarchy = MPS_HIER_REF;
    o->precision = MPS_PREC_FP64;
    o->flux = MPS_FLUX_UPWIND;
    o->nu1 = 2;
    o->nu2 = 2;
    o->max_cycles = 20;
    o->tol = 1e-11;
    o->tol_kind = MPS_TOL_RMS;
    o->arith = MPS_ARITH_EXACT;
    o->start = MPS_START_GUESS;
    o->dc_inner = 0;
    o->omega = 0.0;
}

const char *mps_error_string(int s)
{
    switch (s) {
    case MPS_OK: return "ok";
    case MPS_NOT_CONVERGED: return "not converged";
    case MPS_ERR_NAN: return "NaN or Inf encountered";
    case MPS_ERR_ARG: return "invalid argument";
    case MPS_ERR_HIP: return "HIP runtime error";
    case MPS_ERR_STATE: return "call out of order";
    case MPS_ERR_NOLIB: return "no usable GPU";
    default: return "unknown status";
    }
}

const char *mps_last_error(void) { return g_last_error.c_str(); }

void *mps_stream(mps_ctx *c) { return c ? (void *)c->stream : nullptr; }

int mps_destroy(mps_ctx *c)
{
    if (!c) return MPS_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto &G : c->step)
        if (G.exec) (void)hipGraphExecDestroy(G.exec);
    if (!c->L64.empty()) c->L64[0].b = nullptr;  // b0 is freed below
    if (!c->LT.empty()) c->LT[0].u = nullptr;
    c->L64.size() ? free_levels(c->L64) : void();
    c->L32.size() ? free_levels(c->L32) : void();
    c->LT.size() ? free_levels(c->LT) : void();
    for (auto p : c->faces0) dfree(p);
    dfree(c->source);
    dfree(c->b0);
    dfree(c->partials);
    dfree(c->d_scalar);
    dfree(c->d_dcscale);
    if (c->h_scalar) (void)hipHostFree(c->h_scalar);
    dfree(c->d_status);
    dfree(c->bottom64);
    dfree(c->bottom32);
    dfree(c->bottomT);
    dfree(c->user_tmp);
    dfree(c->user_tmp2);
    dfree(c->dinv);
    dfree(c->dinvT);
    for (auto p : c->bst) dfree(p);
    dfree(c->dwork);
    dfree(c->gbuf);
    dfree(c->fmg_r);
    dfree(c->fmg_e);
    free_gmres(c->gm);
    dfree(c->chem);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->ev_in) (void)hipEventDestroy(c->ev_in);
    for (auto e : c->kt_ev) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return MPS_OK;
}

int mps_create(mps_ctx **out, const mps_grid *grid, const mps_opts *opts)
{
    if (!out || !grid) return fail(MPS_ERR_ARG, "null argument");
    *out = nullptr;
    if (grid->nx < 1 || grid->ny < 1 || grid->nz < 1 || !(grid->lenx > 0) || !(grid->leny > 0) ||
        !(grid->lenz > 0))
        return fail(MPS_ERR_ARG, "bad grid");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        (void)hipGetLastError();
        return fail(MPS_ERR_NOLIB, "no HIP device");
    }
    mps_ctx *c = new mps_ctx();
    c->grid = *grid;
    if (opts) c->opts = *opts;
    else mps_default_opts(&c->opts);
    if (c->opts.nu1 < 0 || c->opts.nu2 < 0 || c->opts.max_cycles < 1) {
        delete c;
        return fail(MPS_ERR_ARG, "bad options");
    }
    if (c->opts.dc_inner < 0 || c->opts.dc_inner > kDcInnerMax) {
        delete c;
        return fail(MPS_ERR_ARG, "dc_inner must be 0 (chosen per step) or 1..4");
    }
    if (!(c->opts.omega >= 0.0 && c->opts.omega < 2.0)) {
        delete c;
        return fail(MPS_ERR_ARG, "omega must be 0 (default) or in (0, 2)");
    }
    // over-relaxed red-black Gauss-Seidel for the solves' fast sweeps (red-
    // black SOR smooths better than plain Gauss-Seidel for omega a little
    // above 1: I. Yavneh, SIAM J. Sci. Comput. 17 (1996) 180-192); on the
    // 257^3 var-coef Poisson 1.1, 1.15, 1.2, 1.3 took 7 -> 6, 5, 5, 5
    // V-cycles to 1e-10
    c->omega = c->opts.omega > 0.0 ? c->opts.omega : (c->opts.hierarchy == MPS_HIER_ALIGNED ? 1.15 : 1.0);
    if (c->opts.start != MPS_START_GUESS &&
        (c->opts.start != MPS_START_FMG || c->opts.hierarchy != MPS_HIER_ALIGNED)) {
        delete c;
        return fail(MPS_ERR_ARG, "start must be MPS_START_GUESS, or MPS_START_FMG with MPS_HIER_ALIGNED");
    }
    int st = build_geometry(c);
    if (st != MPS_OK) {
        delete c;
        return st;
    }
    try {
        HIPCHK(hipGetDevice(&c->device));
        HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        {
            const char *ng = std::getenv("MPS3D_NO_GRAPHS");
            c->graphs = !(ng && ng[0] == '1');
            const char *nf = std::getenv("MPS3D_NO_FUSED_RESTRICT");
            c->fused_rr = !(nf && nf[0] == '1');
            const char *sr = std::getenv("MPS3D_SWEEP_REF");
            c->sweep_ref = sr && sr[0] == '1';
            const char *om = std::getenv("MPS3D_OMEGA");
            if (om) {
                const double w = std::atof(om);
                if (w > 0.0 && w < 2.0) c->omega = w;
            }
            const char *n2 = std::getenv("MPS3D_SWEEP2");
            c->sweep2 = (n2 && n2[0] >= '0' && n2[0] <= '3') ? n2[0] - '0' : MPS_SWEEP2_PREC;
        }
        HIPCHK(hipEventCreate(&c->ev0));
        HIPCHK(hipEventCreate(&c->ev1));
        HIPCHK(hipEventCreateWithFlags(&c->ev_in, hipEventDisableTiming));
        alloc_levels<double>(c, c->L64, true);
        if (c->opts.precision == MPS_PREC_FP32_DC) alloc_levels<float>(c, c->L32, true);
        const Geom &g0 = c->geom[0];
        c->source = dalloc<double>(g0.ng());
        c->b0 = dalloc<double>(g0.ng());
        c->user_tmp = dalloc<double>(g0.ng());
        c->user_tmp2 = dalloc<double>(g0.ng());
        c->max_partials = residual_max_partials(g0);
        c->partials = dalloc<double>(c->max_partials);
        c->d_scalar = dalloc<double>(4);
        c->d_dcscale = dalloc<double>(2);
        fill<double>(c, c->d_dcscale, 2, 1.0);
        HIPCHK(hipHostMalloc((void **)&c->h_scalar, 4 * sizeof(double), hipHostMallocDefault));
        c->d_status = dalloc<int>(1);
        fill<int>(c, c->d_status, 1, 0);
        if (c->mg_ok) {
            const Geom &gb = c->geom.back();
            c->bottom64 = dalloc<double>(bottom_work_elems<double>((int)gb.n(), gb.ng()));
            if (c->opts.precision == MPS_PREC_FP32_DC)
                c->bottom32 = dalloc<float>(bottom_work_elems<float>((int)gb.n(), gb.ng()));
        }
        for (int l = 0; l < c->nlev(); ++l) {
            fill<double>(c, c->L64[l].u, c->geom[l].ng(), 0.0);
            fill<double>(c, c->L64[l].b, c->geom[l].ng(), 0.0);
            fill<double>(c, c->L64[l].r, c->geom[l].ng(), 0.0);
            if (c->L64[l].uold) fill<double>(c, c->L64[l].uold, c->geom[l].ng(), 0.0);
            if (c->opts.precision == MPS_PREC_FP32_DC) {
                fill<float>(c, c->L32[l].u, c->geom[l].ng(), 0.0f);
                fill<float>(c, c->L32[l].b, c->geom[l].ng(), 0.0f);
                fill<float>(c, c->L32[l].r, c->geom[l].ng(), 0.0f);
                if (c->L32[l].uold) fill<float>(c, c->L32[l].uold, c->geom[l].ng(), 0.0f);
            }
        }
        fill<double>(c, c->user_tmp, g0.ng(), 0.0);
        HIPCHK(hipStreamSynchronize(c->stream));
        // the finest u buffer is bound to the caller's array per call
        dfree(c->L64[0].u);
        c->L64[0].u = nullptr;
        c->L64[0].own_u = false;
        dfree(c->L64[0].b);
        c->L64[0].b = nullptr;
    } catch (const HipError &e) {
        int code = fail(MPS_ERR_HIP, "%s: %s", e.what, hipGetErrorString(e.e));
        mps_destroy(c);
        return code;
    }
    *out = c;
    return MPS_OK;
}

// mps_set_bc's device work in one launch, one workgroup per face: the
// caller's face (NULL: zeros) into the context's copy, then restrict_boundary
// (mgridsteps.f90:55-128) level by level, the 4-child mean in the host
// loop's order (as k_restrict_face), with a block barrier between levels
constexpr int kFaceLev = 12;
struct FaceSet {
    const double *src[6];
    double *d0[6];
    int n0[6];
    int nl;
    int fa[kFaceLev][6], ca[kFaceLev][6], cb[kFaceLev][6];
    double *c64[kFaceLev][6];
    float *c32[kFaceLev][6];
};
__global__ void __launch_bounds__(1024) k_set_faces(FaceSet S)
{
    const int f = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    const double *src = S.src[f];
    double *d0 = S.d0[f];
    for (int t = tid; t < S.n0[f]; t += nt) d0[t] = src ? src[t] : 0.0;
    for (int l = 1; l < S.nl; ++l) {
        __syncthreads();
        const double *F = l == 1 ? d0 : S.c64[l - 1][f];
        const int fa = S.fa[l][f], ca = S.ca[l][f], cb = S.cb[l][f];
        double *C = S.c64[l][f];
        float *c32 = S.c32[l][f];
        for (int t = tid; t < ca * cb; t += nt) {
            const int p = t % ca + 1, q = t / ca + 1;
            const double v = 0.25 * (F[(2 * p - 2) + (size_t)fa * (2 * q - 2)] +
                                     F[(2 * p - 1) + (size_t)fa * (2 * q - 2)] +
                                     F[(2 * p - 2) + (size_t)fa * (2 * q - 1)] +
                                     F[(2 * p - 1) + (size_t)fa * (2 * q - 1)]);
            C[t] = v;
            if (c32) c32[t] = (float)v;
        }
    }
}

// restrict_boundary (mgridsteps.f90:55-128) of one face on the device: the
// 4-child mean of an (fa, fb) face onto (ca, cb), in the host loop's order;
// c32 (nullable) gets the FP32 copy
__global__ void k_restrict_face(const double *__restrict__ F, int fa, double *__restrict__ C, int ca, int cb,
                                float *__restrict__ c32)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ca * cb) return;
    const int p = t % ca + 1, q = t / ca + 1;
    const double v = 0.25 * (F[(2 * p - 2) + (size_t)fa * (2 * q - 2)] + F[(2 * p - 1) + (size_t)fa * (2 * q - 2)] +
                             F[(2 * p - 2) + (size_t)fa * (2 * q - 1)] + F[(2 * p - 1) + (size_t)fa * (2 * q - 1)]);
    C[t] = v;
    if (c32) c32[t] = (float)v;
}

// mps_set_bc for faces that are all device arrays (or NULL): copies and
// coarse faces made on the context's stream, no host round trip
void set_bc_device(mps_ctx *c, const double *const faces[6])
{
    const int nl = c->nlev();
    const Geom &g0 = c->geom[0];
    const long long fsz[6] = {(long long)g0.ny * g0.nz, (long long)g0.ny * g0.nz, (long long)g0.nx * g0.nz,
                              (long long)g0.nx * g0.nz, (long long)g0.nx * g0.ny, (long long)g0.nx * g0.ny};
    const bool ref = c->opts.hierarchy == MPS_HIER_REF;
    if (nl <= kFaceLev && fsz[0] < (1LL << 31) && fsz[2] < (1LL << 31) && fsz[4] < (1LL << 31)) {
        // one launch: copies and every coarse level's faces
        FaceSet S{};
        S.nl = ref ? nl : 1;
        for (int f = 0; f < 6; ++f) {
            if (!c->faces0[f]) c->faces0[f] = dalloc<double>(fsz[f]);
            S.src[f] = faces ? faces[f] : nullptr;
            S.d0[f] = c->faces0[f];
            S.n0[f] = (int)fsz[f];
        }
        for (int l = 1; l < S.nl; ++l) {
            const Geom &gf = c->geom[l - 1], &gc = c->geom[l];
            const int fa[6] = {gf.ny, gf.ny, gf.nx, gf.nx, gf.nx, gf.nx};
            const int ca[6] = {gc.ny, gc.ny, gc.nx, gc.nx, gc.nx, gc.nx};
            const int cb[6] = {gc.nz, gc.nz, gc.nz, gc.nz, gc.ny, gc.ny};
            for (int f = 0; f < 6; ++f) {
                const size_t nc = (size_t)ca[f] * cb[f];
                auto &B64 = c->L64[l];
                if (!B64.faces[f]) B64.faces[f] = dalloc<double>(nc);
                float *c32 = nullptr;
                if (c->opts.precision == MPS_PREC_FP32_DC) {
                    auto &B32 = c->L32[l];
                    if (!B32.faces[f]) B32.faces[f] = dalloc<float>(nc);
                    c32 = B32.faces[f];
                }
                S.fa[l][f] = fa[f];
                S.ca[l][f] = ca[f];
                S.cb[l][f] = cb[f];
                S.c64[l][f] = B64.faces[f];
                S.c32[l][f] = c32;
            }
        }
        hipLaunchKernelGGL(k_set_faces, dim3(6), dim3(1024), 0, c->stream, S);
        HIPCHK(hipGetLastError());
        return;
    }
    for (int f = 0; f < 6; ++f) {
        if (!c->faces0[f]) c->faces0[f] = dalloc<double>(fsz[f]);
        if (faces && faces[f])
            HIPCHK(hipMemcpyAsync(c->faces0[f], faces[f], fsz[f] * sizeof(double), hipMemcpyDeviceToDevice,
                                  c->stream));
        else
            fill<double>(c, c->faces0[f], fsz[f], 0.0);
    }
    if (c->opts.hierarchy == MPS_HIER_REF) {
        for (int l = 1; l < nl; ++l) {
            const Geom &gf = c->geom[l - 1], &gc = c->geom[l];
            const int fa[6] = {gf.ny, gf.ny, gf.nx, gf.nx, gf.nx, gf.nx};
            const int ca[6] = {gc.ny, gc.ny, gc.nx, gc.nx, gc.nx, gc.nx};
            const int cb[6] = {gc.nz, gc.nz, gc.nz, gc.nz, gc.ny, gc.ny};
            for (int f = 0; f < 6; ++f) {
                const size_t nc = (size_t)ca[f] * cb[f];
                auto &B64 = c->L64[l];
                if (!B64.faces[f]) B64.faces[f] = dalloc<double>(nc);
                float *c32 = nullptr;
                if (c->opts.precision == MPS_PREC_FP32_DC) {
                    auto &B32 = c->L32[l];
                    if (!B32.faces[f]) B32.faces[f] = dalloc<float>(nc);
                    c32 = B32.faces[f];
                }
                const double *src = l == 1 ? c->faces0[f] : c->L64[l - 1].faces[f];
                hipLaunchKernelGGL(k_restrict_face, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, c->stream, src,
                                   fa[f], B64.faces[f], ca[f], cb[f], c32);
            }
        }
    }
    HIPCHK(hipGetLastError());
}

int mps_set_bc(mps_ctx *c, const int codes[6], const double *const faces[6])
{
    if (!c || !codes) return fail(MPS_ERR_ARG, "null argument");
    try {
        HIPCHK(hipSetDevice(c->device));
        for (int f = 0; f < 6; ++f) {
            if (codes[f] < 0 || codes[f] > 3) return fail(MPS_ERR_ARG, "bad BC code %d", codes[f]);
            c->bc[f] = codes[f];
        }
        bool all_dev = true;
        for (int f = 0; f < 6 && faces; ++f)
            if (faces[f] && !is_device_ptr(faces[f])) all_dev = false;
        if (faces && all_dev) {
            const int st = after_null_stream(c);
            if (st) return st;
            set_bc_device(c, faces);
            c->bc_set = true;
            touch_gen(c);
            c->transpose_ready = false;
            return MPS_OK;
        }
        const int nl = c->nlev();
        c->faces_host.assign((size_t)nl * 6, {});
        const Geom &g0 = c->geom[0];
        const long long fsz[6] = {(long long)g0.ny * g0.nz, (long long)g0.ny * g0.nz, (long long)g0.nx * g0.nz,
                                  (long long)g0.nx * g0.nz, (long long)g0.nx * g0.ny, (long long)g0.nx * g0.ny};
        for (int f = 0; f < 6; ++f) {
            auto &h = c->faces_host[f];
            h.assign(fsz[f], 0.0);
            if (faces && faces[f]) {
                if (is_device_ptr(faces[f]))
                    HIPCHK(hipMemcpy(h.data(), faces[f], fsz[f] * sizeof(double), hipMemcpyDeviceToHost));
                else
                    std::memcpy(h.data(), faces[f], fsz[f] * sizeof(double));
            }
            if (!c->faces0[f]) c->faces0[f] = dalloc<double>(fsz[f]);
            HIPCHK(hipMemcpy(c->faces0[f], h.data(), fsz[f] * sizeof(double), hipMemcpyHostToDevice));
        }
        // coarse faces: 4-child mean per level (restrict_boundary,
        // mgridsteps.f90:55-128), used by the ghost rule of prolongation1
        if (c->opts.hierarchy == MPS_HIER_REF) {
            for (int l = 1; l < nl; ++l) {
                const Geom &gf = c->geom[l - 1], &gc = c->geom[l];
                const int fa[6] = {gf.ny, gf.ny, gf.nx, gf.nx, gf.nx, gf.nx};
                const int fb[6] = {gf.nz, gf.nz, gf.nz, gf.nz, gf.ny, gf.ny};
                const int ca[6] = {gc.ny, gc.ny, gc.nx, gc.nx, gc.nx, gc.nx};
                const int cb[6] = {gc.nz, gc.nz, gc.nz, gc.nz, gc.ny, gc.ny};
                for (int f = 0; f < 6; ++f) {
                    const auto &F = c->faces_host[(size_t)(l - 1) * 6 + f];
                    auto &C = c->faces_host[(size_t)l * 6 + f];
                    C.assign((size_t)ca[f] * cb[f], 0.0);
                    for (int q = 1; q <= cb[f]; ++q)
                        for (int p = 1; p <= ca[f]; ++p)
                            C[(p - 1) + (size_t)ca[f] * (q - 1)] =
                                0.25 * (F[(2 * p - 2) + (size_t)fa[f] * (2 * q - 2)] +
                                        F[(2 * p - 1) + (size_t)fa[f] * (2 * q - 2)] +
                                        F[(2 * p - 2) + (size_t)fa[f] * (2 * q - 1)] +
                                        F[(2 * p - 1) + (size_t)fa[f] * (2 * q - 1)]);
                    auto &B64 = c->L64[l];
                    if (!B64.faces[f]) B64.faces[f] = dalloc<double>(C.size());
                    HIPCHK(hipMemcpy(B64.faces[f], C.data(), C.size() * sizeof(double), hipMemcpyHostToDevice));
                    if (c->opts.precision == MPS_PREC_FP32_DC) {
                        std::vector<float> cf(C.begin(), C.end());
                        auto &B32 = c->L32[l];
                        if (!B32.faces[f]) B32.faces[f] = dalloc<float>(cf.size());
                        HIPCHK(hipMemcpy(B32.faces[f], cf.data(), cf.size() * sizeof(float), hipMemcpyHostToDevice));
                    }
                }
            }
        }
        c->bc_set = true;
        touch_gen(c);
        c->transpose_ready = false;
    } catch (const HipError &e) {
        return fail(MPS_ERR_HIP, "%s: %s", e.what, hipGetErrorString(e.e));
    }
    return MPS_OK;
}

int mps_set_bc_str(mps_ctx *c, const char codes[24], const double *const faces[6])
{
    if (!codes) return fail(MPS_ERR_ARG, "null codes");
    int k[6];
    for (int f = 0; f < 6; ++f) {
        const char *s = codes + 4 * f;
        if (!std::strncmp(s, "DIRC", 4)) k[f] = MPS_BC_DIRC;
        else if (!std::strncmp(s, "ZGRD", 4)) k[f] = MPS_BC_ZGRD;
        else if (!std::strncmp(s, "FLUX", 4)) k[f] = MPS_BC_FLUX;
        else k[f] = MPS_BC_NONE;
    }
    return mps_set_bc(c, k, faces);
}

int mps_set_coeffs(mps_ctx *c, const double *dcoeff, const double *velx, const double *vely, const double *velz,
                   const double *reac, const double *source, int timederiv, double dt)
{
    if (!c) return fail(MPS_ERR_ARG, "null context");
    if (timederiv && !(dt != 0.0)) return fail(MPS_ERR_ARG, "timederiv with dt = 0");
    try {
        HIPCHK(hipSetDevice(c->device));
        if (int st = after_null_stream(c)) return st;
        auto &L = c->L64;
        const long long ng = c->geom[0].ng();
        c->has_vel = velx || vely || velz;
        c->has_reac = reac != nullptr;
        c->timederiv = timederiv != 0;
        c->dt = dt;
        const double *vsrc[3] = {velx, vely, velz};
        for (int l = 0; l < c->nlev(); ++l) {
            double **vd[4] = {&L[l].vx, &L[l].vy, &L[l].vz, &L[l].reac};
            const bool want[4] = {c->has_vel, c->has_vel, c->has_vel, c->has_reac};
            for (int q = 0; q < 4; ++q) {
                if (want[q] && !*vd[q]) {
                    *vd[q] = dalloc<double>(c->geom[l].ng());
                    fill<double>(c, *vd[q], c->geom[l].ng(), 0.0);
                }
                if (!want[q] && *vd[q]) {
                    dfree(*vd[q]);
                    *vd[q] = nullptr;
                    if (c->opts.precision == MPS_PREC_FP32_DC) {
                        float **fd[4] = {&c->L32[l].vx, &c->L32[l].vy, &c->L32[l].vz, &c->L32[l].reac};
                        dfree(*fd[q]);
                        *fd[q] = nullptr;
                    }
                }
            }
        }
        {
            // the finest level's arrays (absent ones zero) in one launch
            StageSet S(c, ng);
            S.add(dcoeff, L[0].D);
            if (c->has_vel) {
                S.add(vsrc[0], L[0].vx);
                S.add(vsrc[1], L[0].vy);
                S.add(vsrc[2], L[0].vz);
            }
            if (c->has_reac) S.add(reac, L[0].reac);
            S.add(source, c->source);
            S.flush();
        }
        derive_coarse(c);
        conductance_coarse_levels(c);
        for (int l = 0; l < c->nlev(); ++l) {
            if (level_op(c, l) == OP_STENCIL) {
                build_stencils(c, L, false);
                break;
            }
        }
        if (c->opts.precision == MPS_PREC_FP32_DC) build_fp32(c);
        build_bottom_inverse(c);
        HIPCHK(hipGetLastError());
        // host arrays were staged: complete before the call returns; device
        // arrays: the set-up stays ordered on the context's stream
        bool host_in = false;
        for (const double *p : {dcoeff, velx, vely, velz, reac, source})
            if (p && !is_device_ptr(p)) host_in = true;
        if (host_in) HIPCHK(hipStreamSynchronize(c->stream));
        c->coeffs_set = true;
        touch_gen(c);
        c->transpose_ready = false;
    } catch (const HipError &e) {
        return fail(MPS_ERR_HIP, "%s: %s", e.what, hipGetErrorString(e.e));
    }
    return MPS_OK;
}

int mps_set_source(mps_ctx *c, const double *source)
{
    if (!c) return fail(MPS_ERR_ARG, "null context");
    if (!c->coeffs_set) return fail(MPS_ERR_STATE, "set the coefficients first (mps_set_coeffs)");
    try {
        HIPCHK(hipSetDevice(c->device));
        if (int st = after_null_stream(c)) return st;
        const long long ng = c->geom[0].ng();
        if (source) stage_in(c, source, c->source, ng);
        else fill<double>(c, c->source, ng, 0.0);
        HIPCHK(hipGetLastError());
        if (source && !is_device_ptr(source)) HIPCHK(hipStreamSynchronize(c->stream));
    } catch (const HipError &e) {
        return fail(MPS_ERR_HIP, "%s: %s", e.what, hipGetErrorString(e.e));
    }
    return MPS_OK;
}

int mps_apply(mps_ctx *c, const double *x, double *ax, double *diag, double *offdiag)
{
    int st = guard(c);
    if (st) return st;
    ExactScope exact(c);
    if (!x) return fail(MPS_ERR_ARG, "null x");
    try {
        const Geom &g = c->geom[0];
        const double *xd = x;
        if (!is_device_ptr(x)) {
            stage_in(c, x, c->user_tmp, g.ng());
            xd = c->user_tmp;
        }
        double *outs[3] = {ax, diag, offdiag};
        double *dev[3] = {nullptr, nullptr, nullptr};
        std::vector<double *> tmp;
        for (int q = 0; q < 3; ++q) {
            if (!outs[q]) continue;
            if (is_device_ptr(outs[q])) dev[q] = outs[q];
            else {
                dev[q] = dalloc<double>(g.n());
                tmp.push_back(dev[q]);
            }
        }
        Lev<double> v = view(c, c->L64, 0);
        with_op(level_op(c, 0), c->has_vel, c->has_reac, [&](auto O, auto V, auto R) {
            hipLaunchKernelGGL((k_apply<double, decltype(O)::value, decltype(V)::value, decltype(R)::value>),
                               cells_grid(g), block(), 0, c->stream, v, xd, dev[0], dev[1], dev[2]);
        });
        for (int q = 0; q < 3; ++q)
            if (outs[q] && dev[q] != outs[q])
                HIPCHK(hipMemcpyAsync(outs[q], dev[q], g.n() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (auto p : tmp) dfree(p);
    } catch (const HipError &e) {
        return fail(MPS_ERR_HIP, "%s: %s", e.what, hipGetErrorString(e.e));
    }
    return MPS_OK;
}

int mps_bvec(mps_ctx *c, const double *sterm, int applybc, double *b)
{
    int st = guard(c);
    if (st) return st;
    ExactScope exact(c);
    if (!b) return fail(MPS_ERR_ARG, "null b");
    try {
        const Geom &g = c->geom[0];
        const double *sd = stage_interior(c, sterm);
        Lev<double> v = view(c, c->L64, 0);
        Faces F{};
        for (int f = 0; f < 6; ++f) F.f[f] = c->faces0[f];
        hipLaunchKernelGGL(k_bvec, cells_grid(g), block(), 0, c->stream, v, F, sd, (const double *)c->source,
                           applybc ? 1 : 0, c->b0);
        double *out = is_device_ptr(b) ? b : c->L64[0].r;
        hipLaunchKernelGGL(k_interior_out<double>, cells_grid(g), block(), 0, c->stream, v, (const double *)c->b0,
                           out);
        if (out != b)
            HIPCHK(hipMemcpyAsync(b, out, g.n() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    } catch (const HipError &e) {
        return fail(MPS_ERR_HIP, "%s: %s", e.what, hipGetErrorString(e.e));
    }
    return MPS_OK;
}

int mps_smooth(mps_ctx *c, double *soln, const double *sterm, int niter, double *res)
{
    int st = guard(c);
    if (st) return st;
    if (!soln || niter < 0) return fail(MPS_ERR_ARG, "bad arguments");
    try {
        const Geom &g = c->geom[0];
        bool staged;
        double *u = bind_solution(c, soln, staged);
        compute_b0(c, stage_interior(c, sterm));
        auto &L = c->L64;
        L[0].u = u;
        L[0].b = c->b0;
        const int op = level_op(c, 0);
        smooth_level<double>(c, L, 0, niter, op);
        if (L[0].u != u) {
            // an odd number of sweep pairs in one pass left the result in the
            // level's other buffer: its interior back into soln
            Lev<double> v = view(c, L, 0);
            hipLaunchKernelGGL(k_copy_interior<double>, cells_grid(g), block(), 0, c->stream, v,
                               (const double *)L[0].u, u);
            L[0].u2 = L[0].u;
            L[0].u = u;
        }
        if (res) {
            launch_residual<double, double>(c, L, 0, u, c->b0, L[0].r, nullptr, op);
            Lev<double> v = view(c, L, 0);
            double *out = is_device_ptr(res) ? res : c->user_tmp2;
            hipLaunchKernelGGL(k_interior_out<double>, cells_grid(g), block(), 0, c->stream, v,
                               (const double *)L[0].r, out);
            if (out != res)
                HIPCHK(hipMemcpyAsync(res, out, g.n() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        }
        unbind_solution(c, soln, staged);
        HIPCHK(hipStreamSynchronize(c->stream));
    } catch (const HipError &e) {
        return fail(MPS_ERR_HIP, "%s: %s", e.what, hipGetErrorString(e.e));
    }
    return MPS_OK;
}

int mps_fill_ghosts(mps_ctx *c, double *soln)
{
    int st = guard(c, false);
    if (st) return st;
    if (!c->bc_set) return fail(MPS_ERR_STATE, "set boundary conditions first");
    if (!soln) return fail(MPS_ERR_ARG, "null soln");
    try {
        const Geom &g = c->geom[0];
        bool staged;
        double *u = bind_solution(c, soln, staged);
        Lev<double> v = view(c, c->L64, 0);
        Faces F{};
        for (int f = 0; f < 6; ++f) F.f[f] = c->faces0[f];
        long long nt = (long long)g.ny * g.nz + (long long)g.nx * g.nz + (long long)g.nx * g.ny;
        hipLaunchKernelGGL(k_ghosts, dim3((unsigned)((nt + kBlock - 1) / kBlock)), block(), 0, c->stream, v, u, F);
        HIPCHK(hipGetLastError());
        unbind_solution(c, soln, staged);
    } catch (const HipError &e) {
        return fail(MPS_ERR_HIP, "%s: %s", e.what, hipGetErrorString(e.e));
    }
    return MPS_OK;
}

int mps_efield(mps_ctx *c, double *phi, double *efield, double scale, int limiter)
{
    int st = guard(c, false);
    if (st) return st;
    if (!c->bc_set) return fail(MPS_ERR_STATE, "set boundary conditions first");
    if (!phi || !efield) return fail(MPS_ERR_ARG, "null argument");
    try {
        const Geom &g = c->geom[0];
        const long long ng = g.ng();
        bool staged;
        double *u = bind_solution(c, phi, staged);
        const bool dev_out = is_device_ptr(efield);
        double *E = dev_out ? efield : dalloc<double>(3 * ng);
        if (!dev_out) HIPCHK(hipMemcpyAsync(E, efield, 3 * ng * sizeof(double), hipMemcpyHostToDevice, c->stream));
        Lev<double> v = view(c, c->L64, 0);
        Faces F{};
        for (int f = 0; f < 6; ++f) F.f[f] = c->faces0[f];
        const long long nt = (long long)g.ny * g.nz + (long long)g.nx * g.nz + (long long)g.nx * g.ny;
        const dim3 gg((unsigned)((nt + kBlock - 1) / kBlock));
        hipLaunchKernelGGL(k_ghosts, gg, block(), 0, c->stream, v, u, F);
        // interior and zero-gradient ghost faces of each component in one
        // pass over whole rows
        const long long plane = (long long)(g.nx + 2) * (g.ny + 2);
        const dim3 ge((unsigned)((plane + kBlock - 1) / kBlock), (unsigned)(g.nz + 2));
        hipLaunchKernelGGL(k_efield, ge, block(), 0, c->stream, v, (const double *)u, E, E + ng, E + 2 * ng, scale,
                           limiter);
        HIPCHK(hipGetLastError());
        unbind_solution(c, phi, staged);
        if (!dev_out) {
            HIPCHK(hipMemcpyAsync(efield, E, 3 * ng * sizeof(double), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            dfree(E);
        }
    } catch (const HipError &e) {
        return fail(MPS_ERR_HIP, "%s: %s", e.what, hipGetErrorString(e.e));
    }
    return MPS_OK;
}

// ------------------------------------------------------------------------
// transport assembly (plasma_pde/*_update_transport, update_electrontemp)
// ------------------------------------------------------------------------

// Host arrays are staged through device buffers for one call: inputs are
// copied in, outputs copied back when the call ends; device arrays are used
// in place.
struct Staging {
    mps_ctx *c;
    std::vector<void *> owned;
    std::vector<std::pair<double *, const double *>> back;   // host dst, device src
    std::vector<size_t> back_n;
    explicit Staging(mps_ctx *ctx) : c(ctx) {}
    Staging(const Staging &) = delete;
    Staging &operator=(const Staging &) = delete;
    ~Staging()
    {
        for (void *p : owned) dfree(p);
    }
    const double *in(const double *p, size_t n)
    {
        if (!p || is_device_ptr(p)) return p;
        double *d = dalloc<double>(n);
        owned.push_back(d);
        HIPCHK(hipMemcpyAsync(d, p, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
        return d;
    }
    // copy_in: the output's previous contents are kept where the kernel
    // does not write (in/out arrays, unwritten ghost cells)
    double *out(double *p, size_t n, bool copy_in)
    {
        if (!p || is_device_ptr(p)) return p;
        double *d = dalloc<double>(n);
        owned.push_back(d);
        if (copy_in) HIPCHK(hipMemcpyAsync(d, p, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
        back.push_back({p, d});
        back_n.push_back(n);
        return d;
    }
    // host outputs copied back and the stream synchronized; with device
    // arrays only, the results stay ordered on the context's stream
    void finish()
    {
        for (size_t q = 0; q < back.size(); ++q)
            HIPCHK(hipMemcpyAsync(back[q].first, back[q].second, back_n[q] * sizeof(double), hipMemcpyDeviceToHost,
                                  c->stream));
        if (!back.empty() || !owned.empty()) HIPCHK(hipStreamSynchronize(c->stream));
    }
};

dim3 ghosted_grid(const Geom &g)
{
    return dim3((unsigned)(((g.nx + 2) * (g.ny + 2) + 255) / 256), (unsigned)(g.nz + 2));
}

int transport_guard(mps_ctx *c, const double scaling[4], int timederiv, const double *soln, double dt)
{
    int st = guard(c, false);
    if (st) return st;
    if (!c->chem) return fail(MPS_ERR_STATE, "set the chemistry first (mps_set_chemistry)");
    if (!scaling) return fail(MPS_ERR_ARG, "null scaling");
    if (timederiv && (!soln || dt == 0.0)) return fail(MPS_ERR_ARG, "the time term needs soln and dt != 0");
    return MPS_OK;
}

TransportIn transport_in(const double scaling[4], int timederiv, double dt, double Tg, double Pg)
{
    TransportIn in{};
    in.vs = scaling[0];
    in.ds = scaling[1];
    in.rs = scaling[2];
    in.ss = scaling[3];
    in.timederiv = timederiv ? 1 : 0;
    in.dt = dt;
    in.Tg = Tg;
    in.Pg = Pg;
    return in;
}

extern "C" int mps_set_chemistry(mps_ctx *c, const mps_chem *chem)
{
    int st = guard(c, false);
    if (st) return st;
    if (!chem) return fail(MPS_ERR_ARG, "null chemistry");
    const int ns = chem->nspecies, nr = chem->nreac;
    if (ns < 1 || ns > kChemMaxSpecies) return fail(MPS_ERR_ARG, "nspecies must be 1..%d", kChemMaxSpecies);
    if (nr < 0 || nr > kChemMaxReac) return fail(MPS_ERR_ARG, "nreac must be 0..%d", kChemMaxReac);
    auto sp_ok = [&](int s) { return s >= 1 && s <= ns; };
    if (!sp_ok(chem->especnum) || !sp_ok(chem->bgspecnum)) return fail(MPS_ERR_ARG, "especnum/bgspecnum out of range");
    const bool no_ions = chem->ionspecmax < chem->ionspecmin;
    if (!no_ions && (!sp_ok(chem->ionspecmin) || !sp_ok(chem->ionspecmax)))
        return fail(MPS_ERR_ARG, "ion species range out of range");
    if (!chem->mobility || !chem->dcoeff || !chem->molmass || !chem->charge)
        return fail(MPS_ERR_ARG, "null species table");
    if (nr > 0 && (!chem->reactants || !chem->products || !chem->k_arrh || !chem->elecenergy))
        return fail(MPS_ERR_ARG, "null reaction table");
    ChemDev h{};
    h.ns = ns;
    h.nr = nr;
    h.e = chem->especnum - 1;
    h.bg = chem->bgspecnum - 1;
    h.imin = no_ions ? 0 : chem->ionspecmin - 1;
    h.imax = no_ions ? -1 : chem->ionspecmax - 1;
    for (int r = 0; r < nr; ++r) {
        for (int j = 0; j < ns; ++j) {
            // reactants(nreac, nspecies) in Fortran order
            const int a = chem->reactants[r + (size_t)nr * j], b = chem->products[r + (size_t)nr * j];
            if (a < 0 || b < 0 || a > 16 || b > 16) return fail(MPS_ERR_ARG, "stoichiometry must be 0..16");
            h.react[r][j] = a;
            h.nu[r][j] = b - a;
        }
        h.A[r] = chem->k_arrh[3 * r];
        h.alpha[r] = chem->k_arrh[3 * r + 1];
        h.Ea[r] = chem->k_arrh[3 * r + 2];
        h.eel[r] = chem->elecenergy[r];
    }
    for (int j = 0; j < ns; ++j) {
        h.mob[j] = chem->mobility[j];
        h.dco[j] = chem->dcoeff[j];
        h.mass[j] = chem->molmass[j];
        h.charge[j] = chem->charge[j];
    }
    const bool models = chem->mob_kind || chem->mob_terms || chem->mob_coef || chem->dcoeff_kind;
    if (models && !(chem->mob_kind && chem->mob_terms && chem->mob_coef && chem->dcoeff_kind))
        return fail(MPS_ERR_ARG, "mob_kind, mob_terms, mob_coef and dcoeff_kind go together");
    for (int s2 = 0; models && s2 < ns; ++s2) {
        const int mk = chem->mob_kind[s2], dk = chem->dcoeff_kind[s2];
        if (mk < 0 || mk > 2 || dk < 0 || dk > 2) return fail(MPS_ERR_ARG, "mob_kind and dcoeff_kind must be 0..2");
        if (dk != 0 && chem->charge[s2] == 0.0) return fail(MPS_ERR_ARG, "Einstein diffusion of a neutral species");
        h.mkind[s2] = mk;
        h.dkind[s2] = dk;
        int nterms = 0;
        for (int j = 0; j < ns; ++j) {
            h.mterm[s2][j] = chem->mob_terms[j + (size_t)ns * s2] != 0;
            nterms += h.mterm[s2][j];
            for (int q = 0; q < 3; ++q) h.mco[s2][j][q] = chem->mob_coef[q + 3 * (j + (size_t)ns * s2)];
        }
        if (mk != 0 && nterms == 0) return fail(MPS_ERR_ARG, "a mobility fit needs at least one partner species");
    }
    if ((chem->rate_kind == nullptr) != (chem->rate_coef == nullptr))
        return fail(MPS_ERR_ARG, "rate_kind and rate_coef go together");
    for (int r = 0; r < nr; ++r) {
        const int rk = chem->rate_kind ? chem->rate_kind[r] : 0;
        if (rk < 0 || rk > 2) return fail(MPS_ERR_ARG, "rate_kind must be 0..2");
        h.rkind[r] = rk;
        for (int q = 0; q < 5; ++q) h.rco[r][q] = chem->rate_coef ? chem->rate_coef[q + 5 * (size_t)r] : 0.0;
    }
    try {
        if (!c->chem) c->chem = dalloc<ChemDev>(1);
        c->chem_host = h;
        HIPCHK(hipMemcpyAsync(c->chem, &c->chem_host, sizeof(ChemDev), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    } catch (const HipError &e) {
        return fail(MPS_ERR_HIP, "%s: %s", e.what, hipGetErrorString(e.e));
    }
    return MPS_OK;
}

extern "C" int mps_species_transport(mps_ctx *c, int specnum, double dt, int timederiv, const double *efield,
                                     const double *electemp, const double *numden, double Tg, double Pg,
                                     const double scaling[4], const double *soln, double *vel, double *dcoeff,
                                     double *reac, double *source, double *prod)
{
    int st = transport_guard(c, scaling, timederiv, soln, dt);
    if (st) return st;
    if (specnum < 1 || specnum > c->chem_host.ns) return fail(MPS_ERR_ARG, "specnum out of range");
    if (!efield || !electemp || !numden) return fail(MPS_ERR_ARG, "null input");
    try {
        const Geom &g = c->geom[0];
        const size_t ng = (size_t)g.ng();
        Staging S(c);
        TransportIn in = transport_in(scaling, timederiv, dt, Tg, Pg);
        const double *E = S.in(efield, 3 * ng);
        in.ex = E;
        in.ey = E + ng;
        in.ez = E + 2 * ng;
        in.te = S.in(electemp, ng);
        in.nd = S.in(numden, (size_t)c->chem_host.ns * ng);
        in.soln = timederiv ? S.in(soln, ng) : nullptr;
        TransportOut out{};
        double *V = S.out(vel, 3 * ng, false);
        if (V) {
            out.vx = V;
            out.vy = V + ng;
            out.vz = V + 2 * ng;
        }
        out.dcoeff = S.out(dcoeff, ng, false);
        out.reac = S.out(reac, ng, false);
        out.source = S.out(source, ng, false);
        out.prod = S.out(prod, ng, true);   // interior only
        Lev<double> v = view(c, c->L64, 0);
        hipLaunchKernelGGL(k_transport<0>, ghosted_grid(g), dim3(256), 0, c->stream, v, (const ChemDev *)c->chem,
                           specnum - 1, in, out);
        HIPCHK(hipGetLastError());
        S.finish();
    } catch (const HipError &e) {
        return fail(MPS_ERR_HIP, "%s: %s", e.what, hipGetErrorString(e.e));
    }
    return MPS_OK;
}

extern "C" int mps_eenergy_transport(mps_ctx *c, double dt, int timederiv, double lscale, const double *efield,
                                     const double *electemp, const double *numden, double Tg, double Pg,
                                     const double scaling[4], const double *soln, double *vel, double *dcoeff,
                                     double *reac, double *source)
{
    int st = transport_guard(c, scaling, timederiv, soln, dt);
    if (st) return st;
    if (!efield || !electemp || !numden) return fail(MPS_ERR_ARG, "null input");
    if (!(lscale > 0.0)) return fail(MPS_ERR_ARG, "lscale must be > 0");
    try {
        const Geom &g = c->geom[0];
        const size_t ng = (size_t)g.ng();
        Staging S(c);
        TransportIn in = transport_in(scaling, timederiv, dt, Tg, Pg);
        const double *E = S.in(efield, 3 * ng);
        in.ex = E;
        in.ey = E + ng;
        in.ez = E + 2 * ng;
        in.te = S.in(electemp, ng);
        in.nd = S.in(numden, (size_t)c->chem_host.ns * ng);
        in.soln = timederiv ? S.in(soln, ng) : nullptr;
        in.dxl[0] = g.dx * lscale;
        in.dxl[1] = g.dy * lscale;
        in.dxl[2] = g.dz * lscale;
        TransportOut out{};
        double *V = S.out(vel, 3 * ng, false);
        if (V) {
            out.vx = V;
            out.vy = V + ng;
            out.vz = V + 2 * ng;
        }
        out.dcoeff = S.out(dcoeff, ng, false);
        out.reac = S.out(reac, ng, false);
        out.source = S.out(source, ng, false);
        Lev<double> v = view(c, c->L64, 0);
        hipLaunchKernelGGL(k_transport<1>, ghosted_grid(g), dim3(256), 0, c->stream, v, (const ChemDev *)c->chem,
                           0, in, out);
        HIPCHK(hipGetLastError());
        S.finish();
    } catch (const HipError &e) {
        return fail(MPS_ERR_HIP, "%s: %s", e.what, hipGetErrorString(e.e));
    }
    return MPS_OK;
}

extern "C" int mps_poisson_transport(mps_ctx *c, double dt, int timederiv, const double *elecden,
                                     const double *iondens, const double scaling[4], const double *soln,
                                     double *vel, double *dcoeff, double *reac, double *source)
{
    int st = transport_guard(c, scaling, timederiv, soln, dt);
    if (st) return st;
    const int nions = c->chem_host.imax - c->chem_host.imin + 1;
    if (!elecden || (nions > 0 && !iondens)) return fail(MPS_ERR_ARG, "null input");
    try {
        const Geom &g = c->geom[0];
        const size_t ng = (size_t)g.ng();
        Staging S(c);
        TransportIn in = transport_in(scaling, timederiv, dt, 0.0, 0.0);
        const double *ne = S.in(elecden, ng);
        const double *ions = nions > 0 ? S.in(iondens, (size_t)nions * ng) : nullptr;
        in.soln = timederiv ? S.in(soln, ng) : nullptr;
        TransportOut out{};
        double *V = S.out(vel, 3 * ng, false);
        if (V) {
            out.vx = V;
            out.vy = V + ng;
            out.vz = V + 2 * ng;
        }
        out.dcoeff = S.out(dcoeff, ng, false);
        out.reac = S.out(reac, ng, false);
        out.source = S.out(source, ng, false);
        Lev<double> v = view(c, c->L64, 0);
        hipLaunchKernelGGL(k_poisson_transport, ghosted_grid(g), dim3(256), 0, c->stream, v,
                           (const ChemDev *)c->chem, ne, ions, in, out);
        HIPCHK(hipGetLastError());
        S.finish();
    } catch (const HipError &e) {
        return fail(MPS_ERR_HIP, "%s: %s", e.what, hipGetErrorString(e.e));
    }
    return MPS_OK;
}

extern "C" int mps_ion_drift_faces(mps_ctx *c, int specnum, const double *efield, const double *numden,
                                   const double *ionsoln, double velscale, double *const faces[6])
{
    int st = guard(c, false);
    if (st) return st;
    if (!c->chem) return fail(MPS_ERR_STATE, "set the chemistry first (mps_set_chemistry)");
    if (!efield || !numden || !ionsoln || !faces) return fail(MPS_ERR_ARG, "null argument");
    const int sp = specnum - 1;
    if (sp < 0 || sp >= c->chem_host.ns) return fail(MPS_ERR_ARG, "species %d out of range", specnum);
    const int kind = c->chem_host.mkind[sp];
    if (kind != 0 && kind != 2)
        return fail(MPS_ERR_ARG, "boundary mobility of kind %d (constant or reduced ion mobility only)", kind);
    const Geom &g = c->geom[0];
    const size_t ng = (size_t)g.ng();
    try {
        Staging S(c);
        const double *E = S.in(efield, 3 * ng);
        const double *nd = S.in(numden, (size_t)c->chem_host.ns * ng);
        const double *ni = S.in(ionsoln, ng);
        const size_t fsz[6] = {(size_t)g.ny * g.nz, (size_t)g.ny * g.nz, (size_t)g.nx * g.nz,
                               (size_t)g.nx * g.nz, (size_t)g.nx * g.ny, (size_t)g.nx * g.ny};
        IonFaces F{};
        size_t most = 0;
        for (int f = 0; f < 6; ++f) {
            F.f[f] = S.out(faces[f], fsz[f], false);
            if (faces[f]) most = std::max(most, fsz[f]);
        }
        if (most > 0) {
            Lev<double> v = view(c, c->L64, 0);
            hipLaunchKernelGGL(k_ion_drift_faces, dim3((unsigned)((most + 255) / 256), 6), dim3(256), 0, c->stream,
                               v, (const ChemDev *)c->chem, sp, E, nd, ni, velscale, F);
            HIPCHK(hipGetLastError());
        }
        S.finish();
    } catch (const HipError &e) {
        return fail(MPS_ERR_HIP, "%s: %s", e.what, hipGetErrorString(e.e));
    }
    return MPS_OK;
}

// the boundary updaters' faces (k_thermal_faces): kind as the kernel's
static int thermal_faces(mps_ctx *c, int kind, const int codes[6], const double *n, const double *T, double c1,
                         double Tg, double escale, const double twall[6], double vscale, double *const faces[6])
{
    int st = guard(c, false);
    if (st) return st;
    if (!codes || !faces || !n || (kind != 1 && !T)) return fail(MPS_ERR_ARG, "null argument");
    for (int f = 0; f < 6; ++f)
        if (codes[f] < 0 || codes[f] > 3) return fail(MPS_ERR_ARG, "bad BC code %d", codes[f]);
    const Geom &g = c->geom[0];
    const size_t ng = (size_t)g.ng();
    try {
        Staging S(c);
        const double *nd = S.in(n, ng);
        const double *Td = T ? S.in(T, ng) : nullptr;
        const size_t fsz[6] = {(size_t)g.ny * g.nz, (size_t)g.ny * g.nz, (size_t)g.nx * g.nz,
                               (size_t)g.nx * g.nz, (size_t)g.nx * g.ny, (size_t)g.nx * g.ny};
        ThermalFaces F{};
        size_t most = 0;
        for (int f = 0; f < 6; ++f) {
            F.code[f] = codes[f];
            F.twall[f] = twall ? twall[f] : 0.0;
            // faces of other codes keep their values: in/out
            F.f[f] = S.out(faces[f], fsz[f], true);
            if (faces[f]) most = std::max(most, fsz[f]);
        }
        if (most > 0) {
            Lev<double> v = view(c, c->L64, 0);
            hipLaunchKernelGGL(k_thermal_faces, dim3((unsigned)((most + 255) / 256), 6), dim3(256), 0, c->stream, v,
                               kind, nd, Td, c1, Tg, escale, vscale, F);
            HIPCHK(hipGetLastError());
        }
        S.finish();
    } catch (const HipError &e) {
        return fail(MPS_ERR_HIP, "%s: %s", e.what, hipGetErrorString(e.e));
    }
    return MPS_OK;
}

// const_1 = 8 k_B / PI / m of the updaters, PI the reference's default-real
// literal (plasma_pde/fundconstants.f90:8)
static double thermal_c1(double mass)
{
    const double pi_ref = (double)3.14159265f;
    return 8.0 * kBoltz / pi_ref / mass;
}

extern "C" int mps_eden_faces(mps_ctx *c, const int codes[6], const double *elecden, const double *electemp,
                              double velscale, double *const faces[6])
{
    return thermal_faces(c, 0, codes, elecden, electemp, thermal_c1(9.109382e-31), 0.0, 0.0, nullptr, velscale,
                         faces);
}

extern "C" int mps_neutral_faces(mps_ctx *c, const int codes[6], const double *neutralden, double molmass,
                                 double Tg, double velscale, double *const faces[6])
{
    if (!(molmass > 0.0)) return fail(MPS_ERR_ARG, "molmass must be > 0");
    return thermal_faces(c, 1, codes, neutralden, nullptr, thermal_c1(molmass), Tg, 0.0, nullptr, velscale, faces);
}

extern "C" int mps_eenergy_faces(mps_ctx *c, const int codes[6], const double *elecenergy, const double *elecden,
                                 double eenergyscale, const double bcparams[6], double velscale,
                                 double *const faces[6])
{
    if (!bcparams) return fail(MPS_ERR_ARG, "null bcparams");
    return thermal_faces(c, 2, codes, elecenergy, elecden, thermal_c1(9.109382e-31), 0.0, eenergyscale, bcparams,
                         velscale, faces);
}

extern "C" int mps_floor(mps_ctx *c, double *soln, double vmin)
{
    int st = guard(c, false);
    if (st) return st;
    if (!c->bc_set) return fail(MPS_ERR_STATE, "set boundary conditions first");
    if (!soln) return fail(MPS_ERR_ARG, "null soln");
    try {
        const Geom &g = c->geom[0];
        bool staged;
        double *u = bind_solution(c, soln, staged);
        Lev<double> v = view(c, c->L64, 0);
        hipLaunchKernelGGL(k_floor, cells_grid(g), block(), 0, c->stream, v, u, vmin);
        Faces F{};
        for (int f = 0; f < 6; ++f) F.f[f] = c->faces0[f];
        const long long nt = (long long)g.ny * g.nz + (long long)g.nx * g.nz + (long long)g.nx * g.ny;
        hipLaunchKernelGGL(k_ghosts, dim3((unsigned)((nt + kBlock - 1) / kBlock)), block(), 0, c->stream, v, u, F);
        HIPCHK(hipGetLastError());
        unbind_solution(c, soln, staged);
    } catch (const HipError &e) {
        return fail(MPS_ERR_HIP, "%s: %s", e.what, hipGetErrorString(e.e));
    }
    return MPS_OK;
}

extern "C" int mps_electron_temp(mps_ctx *c, double *eenergysoln, const double *elecden, double eenergyscale,
                                 double etempmin, double *electemp)
{
    int st = guard(c, false);
    if (st) return st;
    if (!eenergysoln || !elecden || !electemp) return fail(MPS_ERR_ARG, "null argument");
    try {
        const Geom &g = c->geom[0];
        const size_t ng = (size_t)g.ng();
        Staging S(c);
        const double *ne = S.in(elecden, ng);
        double *eps = S.out(eenergysoln, ng, true);
        double *te = S.out(electemp, ng, true);   // ghosts are not written
        Lev<double> v = view(c, c->L64, 0);
        hipLaunchKernelGGL(k_electron_temp, cells_grid(g), block(), 0, c->stream, v, eps, ne, eenergyscale, etempmin,
                           te);
        HIPCHK(hipGetLastError());
        S.finish();
    } catch (const HipError &e) {
        return fail(MPS_ERR_HIP, "%s: %s", e.what, hipGetErrorString(e.e));
    }
    return MPS_OK;
}

int mps_vcycle(mps_ctx *c, double *soln, const double *sterm, double *resnorm)
{
    int st = guard_mg(c);
    if (st) return st;
    if (!soln) return fail(MPS_ERR_ARG, "null soln");
    try {
        bool staged;
        double *u = bind_solution(c, soln, staged);
        fill<int>(c, c->d_status, 1, 0);
        compute_b0(c, stage_interior(c, sterm));
        fp64_cycle(c, u);
        double ss = c->geom[0].bottom ? 0.0 : read_scalar(c);
        HIPCHK(hipMemcpyAsync(c->h_scalar + 1, c->d_status, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        unbind_solution(c, soln, staged);
        HIPCHK(hipStreamSynchronize(c->stream));
        int bstat;
        std::memcpy(&bstat, c->h_scalar + 1, sizeof(int));
        if (resnorm) *resnorm = std::sqrt(ss / (double)c->geom[0].n());
        if (bstat != 0) return fail(MPS_ERR_NAN, "NaN or Inf in the coarse-grid GMRES");
        if (!std::isfinite(ss)) return fail(MPS_ERR_NAN, "residual is not finite");
    } catch (const HipError &e) {
        return fail(MPS_ERR_HIP, "%s: %s", e.what, hipGetErrorString(e.e));
    }
    return MPS_OK;
}

int mps_solve(mps_ctx *c, double *soln, int *ncycles, double *resnorm, double *history)
{
    int st = guard_mg(c);
    if (st) return st;
    if (!soln) return fail(MPS_ERR_ARG, "null soln");
    int status = MPS_NOT_CONVERGED;
    try {
        bool staged;
        double *u = bind_solution(c, soln, staged);
        HIPCHK(hipEventRecord(c->ev0, c->stream));
        fill<int>(c, c->d_status, 1, 0);
        compute_b0(c, nullptr);
        double meas = 0.0;
        const int it = solve_core(c, u, c->b0, false, history, meas);
        HIPCHK(hipEventRecord(c->ev1, c->stream));
        HIPCHK(hipMemcpyAsync(c->h_scalar + 1, c->d_status, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        unbind_solution(c, soln, staged);
        HIPCHK(hipStreamSynchronize(c->stream));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        c->last_ms = ms;
        c->last_cycles = it;
        if (ncycles) *ncycles = it;
        if (resnorm) *resnorm = meas;
        int bstat;
        std::memcpy(&bstat, c->h_scalar + 1, sizeof(int));
        if (bstat != 0) return fail(MPS_ERR_NAN, "NaN or Inf in the coarse-grid GMRES");
        if (!std::isfinite(meas)) return fail(MPS_ERR_NAN, "residual is not finite");
        status = meas <= c->opts.tol ? MPS_OK : MPS_NOT_CONVERGED;
        if (status == MPS_NOT_CONVERGED) fail(MPS_NOT_CONVERGED, "tolerance %g not met after %d cycles (%g)",
                                              c->opts.tol, it, meas);
    } catch (const HipError &e) {
        return fail(MPS_ERR_HIP, "%s: %s", e.what, hipGetErrorString(e.e));
    }
    return status;
}

int mps_solve_rhs(mps_ctx *c, const double *bsrc, double *soln, int *ncycles, double *resnorm, double *history)
{
    int st = guard_mg(c);
    if (st) return st;
    if (!bsrc || !soln) return fail(MPS_ERR_ARG, "null argument");
    int status = MPS_NOT_CONVERGED;
    try {
        const Geom &g = c->geom[0];
        bool staged;
        double *u = bind_solution(c, soln, staged);
        if (!c->gbuf) {
            c->gbuf = dalloc<double>(g.ng());
            fill<double>(c, c->gbuf, g.ng(), 0.0);  // ghost layers stay zero
        }
        const double *bd = stage_interior(c, bsrc);
        Lev<double> v = view(c, c->L64, 0);
        hipLaunchKernelGGL(k_interior_in<double>, cells_grid(g), block(), 0, c->stream, v, bd, c->gbuf);
        HIPCHK(hipEventRecord(c->ev0, c->stream));
        fill<int>(c, c->d_status, 1, 0);
        double meas = 0.0;
        const int it = solve_core(c, u, c->gbuf, true, history, meas);
        HIPCHK(hipEventRecord(c->ev1, c->stream));
        HIPCHK(hipMemcpyAsync(c->h_scalar + 1, c->d_status, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        unbind_solution(c, soln, staged);
        HIPCHK(hipStreamSynchronize(c->stream));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        c->last_ms = ms;
        c->last_cycles = it;
        if (ncycles) *ncycles = it;
        if (resnorm) *resnorm = meas;
        int bstat;
        std::memcpy(&bstat, c->h_scalar + 1, sizeof(int));
        if (bstat != 0) return fail(MPS_ERR_NAN, "NaN or Inf in the coarse-grid GMRES");
        if (!std::isfinite(meas)) return fail(MPS_ERR_NAN, "residual is not finite");
        status = meas <= c->opts.tol ? MPS_OK : MPS_NOT_CONVERGED;
        if (status == MPS_NOT_CONVERGED) fail(MPS_NOT_CONVERGED, "tolerance %g not met after %d cycles (%g)",
                                              c->opts.tol, it, meas);
    } catch (const HipError &e) {
        return fail(MPS_ERR_HIP, "%s: %s", e.what, hipGetErrorString(e.e));
    }
    return status;
}

int mps_restrict(mps_ctx *c, const double *fine, double *coarse)
{
    int st = guard(c, false);
    if (st) return st;
    if (!fine || !coarse) return fail(MPS_ERR_ARG, "null argument");
    const Geom &g = c->geom[0];
    if ((g.nx & 1) || (g.ny & 1) || (g.nz & 1)) return fail(MPS_ERR_ARG, "restriction needs even nx, ny, nz");
    std::vector<double *> tmp;
    try {
        const Lev<double> vf = grid_view(g.nx, g.ny, g.nz), vc = grid_view(g.nx / 2, g.ny / 2, g.nz / 2);
        const long long ngc = (long long)(vc.nx + 2) * (vc.ny + 2) * (vc.nz + 2), nc = (long long)vc.nx * vc.ny * vc.nz;
        for (int q = 0; q < 4; ++q) tmp.push_back(dalloc<double>(q == 0 ? g.ng() : ngc));
        fill<double>(c, tmp[0], g.ng(), 0.0);
        hipLaunchKernelGGL(k_interior_in<double>, cells_grid(g), block(), 0, c->stream, vf, stage_interior(c, fine),
                           tmp[0]);
        // the V-cycle's kernel: u and r restricted together (here both are fine)
        Geom gc{};
        gc.nx = vc.nx, gc.ny = vc.ny, gc.nz = vc.nz;
        hipLaunchKernelGGL(k_restrict_fas<double>, cells_grid(gc), block(), 0, c->stream, vf, vc,
                           (const double *)tmp[0], (const double *)tmp[0], tmp[1], tmp[2], tmp[3]);
        double *out = is_device_ptr(coarse) ? coarse : c->user_tmp2;
        hipLaunchKernelGGL(k_interior_out<double>, cells_grid(gc), block(), 0, c->stream, vc, (const double *)tmp[2],
                           out);
        if (out != coarse)
            HIPCHK(hipMemcpyAsync(coarse, out, nc * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
    } catch (const HipError &e) {
        for (auto p : tmp) dfree(p);
        return fail(MPS_ERR_HIP, "%s: %s", e.what, hipGetErrorString(e.e));
    }
    for (auto p : tmp) dfree(p);
    return MPS_OK;
}

int mps_prolong1(mps_ctx *c, const double *coarse, double *fine)
{
    int st = guard(c, false);
    if (st) return st;
    if (!fine || !coarse) return fail(MPS_ERR_ARG, "null argument");
    const Geom &g = c->geom[0];
    if ((g.nx & 1) || (g.ny & 1) || (g.nz & 1)) return fail(MPS_ERR_ARG, "prolongation needs even nx, ny, nz");
    std::vector<double *> tmp;
    try {
        const Lev<double> vf = grid_view(g.nx, g.ny, g.nz), vc = grid_view(g.nx / 2, g.ny / 2, g.nz / 2);
        const long long ngc = (long long)(vc.nx + 2) * (vc.ny + 2) * (vc.nz + 2);
        for (int q = 0; q < 3; ++q) tmp.push_back(dalloc<double>(q == 0 ? g.ng() : ngc));
        fill<double>(c, tmp[0], g.ng(), 0.0);   // fine, zero: the kernel adds the interpolant
        fill<double>(c, tmp[2], ngc, 0.0);      // the "old" coarse values: zero
        stage_in(c, coarse, tmp[1], ngc);
        Geom gc{};
        gc.nx = vc.nx, gc.ny = vc.ny, gc.nz = vc.nz;
        const FacesT<double> none{};
        hipLaunchKernelGGL((k_prolong1_correct<double, true>), cells_grid(gc), block(), 0, c->stream, vf, vc, tmp[0],
                           (const double *)tmp[1], (const double *)tmp[2], none);
        double *out = is_device_ptr(fine) ? fine : c->user_tmp2;
        hipLaunchKernelGGL(k_interior_out<double>, cells_grid(g), block(), 0, c->stream, vf, (const double *)tmp[0],
                           out);
        if (out != fine) HIPCHK(hipMemcpyAsync(fine, out, g.n() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
    } catch (const HipError &e) {
        for (auto p : tmp) dfree(p);
        return fail(MPS_ERR_HIP, "%s: %s", e.what, hipGetErrorString(e.e));
    }
    for (auto p : tmp) dfree(p);
    return MPS_OK;
}

int mps_get_grid(mps_ctx *c, mps_grid *grid, int *nspecies)
{
    if (!c || !grid) return fail(MPS_ERR_ARG, "null argument");
    *grid = c->grid;
    if (nspecies) *nspecies = c->chem ? c->chem_host.ns : 0;
    return MPS_OK;
}

int mps_last_timing(mps_ctx *c, double *total_ms, int *ncycles)
{
    if (!c) return fail(MPS_ERR_ARG, "null context");
    if (total_ms) *total_ms = c->last_ms;
    if (ncycles) *ncycles = c->last_cycles;
    return MPS_OK;
}

int mps_last_vcycles(mps_ctx *c, int *nvcycles)
{
    if (!c || !nvcycles) return fail(MPS_ERR_ARG, "null argument");
    *nvcycles = c->last_vcycles;
    return MPS_OK;
}

}  // extern "C"

// ======================================================= Krylov and adjoint
namespace {

// lazily built transposed hierarchy (explicit A^T stencils on every level)
void ensure_transpose(mps_ctx *c)
{
    if (c->transpose_ready) return;
    if (c->LT.empty()) {
        c->LT.assign(c->geom.size(), LevelBuf<double>{});
        for (int l = 0; l < c->nlev(); ++l) {
            const long long ng = c->geom[l].ng();
            LevelBuf<double> &B = c->LT[l];
            B.u = dalloc<double>(ng);
            B.b = dalloc<double>(ng);
            B.r = dalloc<double>(ng);
            if (l > 0) B.uold = dalloc<double>(ng);
            fill<double>(c, B.u, ng, 0.0);
            fill<double>(c, B.b, ng, 0.0);
            fill<double>(c, B.r, ng, 0.0);
            if (B.uold) fill<double>(c, B.uold, ng, 0.0);
            for (int d = 0; d < 3; ++d) B.w[d] = nullptr;
            B.op = OP_STENCIL;
        }
        // restriction weights are shared with the forward hierarchy
        for (int l = 0; l < c->nlev(); ++l)
            for (int d = 0; d < 3; ++d) c->LT[l].w[d] = nullptr;
        dfree(c->LT[0].u);
        c->LT[0].u = nullptr;
        c->LT[0].own_u = false;
    }
    build_stencils(c, c->LT, true);
    c->LT.back().inv = c->opts.hierarchy == MPS_HIER_ALIGNED ? c->dinvT : nullptr;
    HIPCHK(hipStreamSynchronize(c->stream));
    c->transpose_ready = true;
}

// aligned weights live on the forward levels; the transposed cycle borrows them
void borrow_weights(mps_ctx *c)
{
    for (int l = 0; l < c->nlev(); ++l)
        for (int d = 0; d < 3; ++d) c->LT[l].w[d] = c->L64[l].w[d];
}

void release_weights(mps_ctx *c)
{
    for (int l = 0; l < c->nlev(); ++l)
        for (int d = 0; d < 3; ++d) c->LT[l].w[d] = nullptr;
}

// one V-cycle preconditioner application z = M^-1 v (homogeneous BCs,
// zero initial guess), in FP64 or with FP32 inner cycles
void precondition(mps_ctx *c, const double *v, double *z)
{
    const Geom &g = c->geom[0];
    if (c->opts.precision == MPS_PREC_FP32_DC) {
        auto &F = c->L32;
        convert<double, float>(c, v, F[0].b, g.ng());
        fill<float>(c, F[0].u, g.ng(), 0.0f);
        vcycle<float>(c, F, 0, c->bottom32, true);
        convert<float, double>(c, F[0].u, z, g.ng());
    } else {
        auto &L = c->L64;
        fill<double>(c, z, g.ng(), 0.0);
        double *ub = L[0].u, *bb = L[0].b;
        L[0].u = z;
        L[0].b = const_cast<double *>(v);
        vcycle<double>(c, L, 0, c->bottom64, true);
        L[0].u = ub;
        L[0].b = bb;
    }
}

void matvec(mps_ctx *c, std::vector<LevelBuf<double>> &L, const double *u, double *out)
{
    Lev<double> v = view(c, L, 0);
    with_op(level_op(c, L, 0), c->has_vel, c->has_reac, [&](auto O, auto V, auto R) {
        hipLaunchKernelGGL((k_matvec<double, decltype(O)::value, decltype(V)::value, decltype(R)::value>),
                           cells_grid(c->geom[0]), block(), 0, c->stream, v, u, out);
    });
}

// sum_q of (w . v_q) for q < nv into host h[0..nv)
void multidot(mps_ctx *c, GmresBufs &G, const double *w, double *const *vs, int nv, double *h)
{
    Lev<double> v = view(c, c->L64, 0);
    VecList VL{};
    for (int q = 0; q < nv; ++q) VL.v[q] = vs[q];
    hipLaunchKernelGGL(k_multidot, red_grid(c->geom[0]), block(), 0, c->stream, v, w, VL, nv, G.dpart);
    hipLaunchKernelGGL(k_reduce_cols, dim3(nv), block(), 0, c->stream, (const double *)G.dpart, G.nblocks, G.dh);
    HIPCHK(hipMemcpyAsync(G.hh, G.dh, nv * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    std::memcpy(h, G.hh, nv * sizeof(double));
}

// multidot with the products left on the device in out[0..nv)
void multidot_dev(mps_ctx *c, GmresBufs &G, const double *w, double *const *vs, int nv, double *out)
{
    Lev<double> v = view(c, c->L64, 0);
    VecList VL{};
    for (int q = 0; q < nv; ++q) VL.v[q] = vs[q];
    hipLaunchKernelGGL(k_multidot, red_grid(c->geom[0]), block(), 0, c->stream, v, w, VL, nv, G.dpart);
    hipLaunchKernelGGL(k_reduce_cols, dim3(nv), block(), 0, c->stream, (const double *)G.dpart, G.nblocks, out);
}

// w -= sum_q h[q] v_q with device coefficients h (the first min(nv, *nvd)
// when nvd is given); |w|^2 into *norm2 (device) when norm2 is given
void multiaxpy_dev(mps_ctx *c, double *w, double *const *vs, int nv, const double *h, const int *nvd,
                   double *norm2)
{
    Lev<double> v = view(c, c->L64, 0);
    VecList VL{};
    for (int q = 0; q < nv; ++q) VL.v[q] = vs[q];
    dim3 gr = red_grid(c->geom[0]);
    hipLaunchKernelGGL(k_multiaxpy, gr, block(), 0, c->stream, v, w, VL, h, nv, norm2 ? c->partials : nullptr, nvd);
    if (norm2)
        hipLaunchKernelGGL(k_reduce_final, dim3(1), block(), 0, c->stream, (const double *)c->partials,
                           (int)(gr.x * gr.y), norm2);
}

// performgmres with the reference's preconditioner on the finest level (at
// most kBottomMax cells): b (interior, device) into the level's zero-ghost
// right-hand side buffer, then k_bottom on x (ghosted, device) with `work`
// (bottom_work_elems doubles; k_bottom initialises every part it reads).
// Queues the copy of the kernel's NaN status for gmres_ref_status.
void gmres_ref_launch(mps_ctx *c, const double *bd, double *u, int m, int max_restarts, double tol, double *work)
{
    const Geom &g = c->geom[0];
    double *bg = c->L64[0].r;
    fill<double>(c, bg, g.ng(), 0.0);
    Lev<double> v = view(c, c->L64, 0);
    hipLaunchKernelGGL(k_interior_in<double>, cells_grid(g), block(), 0, c->stream, v, bd, bg);
    with_op(level_op(c, 0), c->has_vel, c->has_reac, [&](auto O, auto V, auto R) {
        if (m <= kBottomLdsM)
            hipLaunchKernelGGL((k_bottom<double, decltype(O)::value, decltype(V)::value, decltype(R)::value, true>),
                               dim3(1), block(), 0, c->stream, v, u, (const double *)bg, work, m, max_restarts, tol,
                               c->d_status);
        else
            hipLaunchKernelGGL((k_bottom<double, decltype(O)::value, decltype(V)::value, decltype(R)::value, false>),
                               dim3(1), block(), 0, c->stream, v, u, (const double *)bg, work, m, max_restarts, tol,
                               c->d_status);
    });
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(c->h_scalar + 1, c->d_status, sizeof(int), hipMemcpyDeviceToHost, c->stream));
}

// status of the last gmres_ref_launch (after the stream is synchronized)
int gmres_ref_status(mps_ctx *c)
{
    int bstat;
    std::memcpy(&bstat, c->h_scalar + 1, sizeof(int));
    if (bstat != 0) return fail(MPS_ERR_NAN, "NaN or Inf in Arnoldi");
    return MPS_OK;
}

void free_gmres(GmresBufs &G)
{
    for (auto p : G.V) dfree(p);
    for (auto p : G.Z) dfree(p);
    dfree(G.w);
    dfree(G.bg);
    dfree(G.dh);
    dfree(G.dpart);
    dfree(G.darn);
    dfree(G.dstate);
    if (G.hh) (void)hipHostFree(G.hh);
    if (G.hflag) (void)hipHostFree(G.hflag);
    for (auto e : G.ev) (void)hipEventDestroy(e);
    G = GmresBufs{};
}

// sizes of the device Arnoldi block (doubles) for m <= kMaxVec - 1
constexpr int kArnLd = kMaxVec;                    // leading dimension of H (m + 1 <= kMaxVec)
constexpr size_t kArnH = (size_t)kArnLd * kMaxVec; // H
constexpr size_t kArnDoubles = kArnH + 7 * (size_t)kMaxVec + 1;

ArnoldiDev arnoldi_view(GmresBufs &G)
{
    double *p = G.darn;
    ArnoldiDev A;
    A.H = p;
    A.cs = p + kArnH;
    A.sn = A.cs + kMaxVec;
    A.gv = A.sn + kMaxVec;
    A.y = A.gv + kMaxVec;
    A.hn = A.y + kMaxVec;   // 1 double (padded to kMaxVec)
    A.state = G.dstate;
    return A;
}
// the second pass's products, the subtracted coefficients and |w|^2
double *arn_h2(GmresBufs &G) { return G.darn + kArnH + 5 * (size_t)kMaxVec; }
double *arn_ny(GmresBufs &G) { return G.darn + kArnH + 6 * (size_t)kMaxVec; }
double *arn_wn2(GmresBufs &G) { return G.darn + kArnH + 7 * (size_t)kMaxVec; }

// the context's FGMRES(m) vectors, (re)allocated when m grows. Every vector
// starts with zero ghosts and the Krylov kernels write interiors only, so
// the ghosts stay zero from call to call.
GmresBufs &gmres_bufs(mps_ctx *c, int m)
{
    GmresBufs &G = c->gm;
    if (G.m >= m) return G;
    HIPCHK(hipStreamSynchronize(c->stream));
    free_gmres(G);
    const Geom &g = c->geom[0];
    const dim3 gr = red_grid(g);
    G.nblocks = (int)(gr.x * gr.y);
    for (int q = 0; q <= m; ++q) G.V.push_back(dalloc<double>(g.ng()));
    for (int q = 0; q < m; ++q) G.Z.push_back(dalloc<double>(g.ng()));
    G.w = dalloc<double>(g.ng());
    G.bg = dalloc<double>(g.ng());
    G.dh = dalloc<double>(kMaxVec);
    G.dpart = dalloc<double>((size_t)G.nblocks * kMaxVec);
    HIPCHK(hipHostMalloc((void **)&G.hh, kMaxVec * sizeof(double), hipHostMallocDefault));
    G.darn = dalloc<double>(kArnDoubles);
    G.dstate = dalloc<int>(4);
    HIPCHK(hipHostMalloc((void **)&G.hflag, (kMaxVec + 4) * sizeof(int), hipHostMallocDefault));
    G.ev.resize(kMaxVec);
    for (auto &e : G.ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    fill<double>(c, G.darn, kArnDoubles, 0.0);
    for (auto p : G.V) fill<double>(c, p, g.ng(), 0.0);
    for (auto p : G.Z) fill<double>(c, p, g.ng(), 0.0);
    fill<double>(c, G.w, g.ng(), 0.0);
    fill<double>(c, G.bg, g.ng(), 0.0);
    G.m = m;
    return G;
}

// v = w / s[0] with s on the device (nothing when s[0] <= 0)
void scale_from(mps_ctx *c, const double *w, double *v, const double *s)
{
    Lev<double> vw = view(c, c->L64, 0);
    hipLaunchKernelGGL(k_scale, cells_grid(c->geom[0]), block(), 0, c->stream, vw, w, v, s);
}

}  // namespace

extern "C" {

int mps_gmres(mps_ctx *c, const double *b, double *x, int m, int max_restarts, double tol, int precond, int *iters,
              double *relres)
{
    int st = guard(c);
    if (st) return st;
    if (!b || !x || m < 1 || max_restarts < 1 || !(tol >= 0)) return fail(MPS_ERR_ARG, "bad arguments");
    if (precond != 0 && precond != 1) return fail(MPS_ERR_ARG, "precond must be 0 or 1");
    const Geom &g = c->geom[0];
    try {
        bool staged;
        double *u = bind_solution(c, x, staged);
        if (precond == 0) {
            // the reference's performgmres: one workgroup, n <= kBottomMax
            if (g.n() > kBottomMax)
                return fail(MPS_ERR_ARG, "precond 0 (reference GMRES) supports at most %d cells", kBottomMax);
            if (m > 4 * kBottomMax) return fail(MPS_ERR_ARG, "m too large");
            double *work = dalloc<double>(bottom_work_elems<double>(m, g.ng()));
            try {
                gmres_ref_launch(c, stage_interior(c, b), u, m, max_restarts, tol, work);
                unbind_solution(c, x, staged);
                HIPCHK(hipStreamSynchronize(c->stream));
            } catch (...) {
                dfree(work);
                throw;
            }
            dfree(work);
            if (iters) *iters = -1;
            if (relres) *relres = -1.0;
            return gmres_ref_status(c);
        }
        // ---------------- FGMRES(m), right-preconditioned by one V-cycle
        if (m + 1 > kMaxVec) return fail(MPS_ERR_ARG, "m at most %d", kMaxVec - 1);
        if (!c->mg_ok) return fail(MPS_ERR_ARG, "no multigrid hierarchy for this grid: %s", c->mg_error.c_str());
        GmresBufs &G = gmres_bufs(c, m);
        const double *bd = stage_interior(c, b);
        Lev<double> v0 = view(c, c->L64, 0);
        hipLaunchKernelGGL(k_interior_in<double>, cells_grid(g), block(), 0, c->stream, v0, bd, G.bg);
        double bn2;
        {
            double *one[1] = {G.bg};
            multidot(c, G, G.bg, one, 1, &bn2);
        }
        const double bnorm = std::sqrt(bn2);
        auto true_residual = [&]() {  // w = b - A x, returns ||w||
            const int np = launch_residual<double, double>(c, c->L64, 0, u, G.bg, G.w, c->partials, level_op(c, 0));
            hipLaunchKernelGGL(k_reduce_final, dim3(1), block(), 0, c->stream, (const double *)c->partials, np,
                               c->d_scalar);
            return std::sqrt(read_scalar(c));
        };
        int total = 0;
        double beta = true_residual();
        int status = MPS_NOT_CONVERGED;
        if (bnorm == 0.0) {
            fill<double>(c, u, g.ng(), 0.0);
            beta = 0.0;
            status = MPS_OK;
        }
        // The Gram-Schmidt products stay on the device, k_arnoldi_step
        // updates H and the Givens rotations there and marks the restart
        // done; per Arnoldi step the host waits once, for that 4-byte flag
        // (copied to pinned memory behind an event), before it queues the
        // next step, and reads the residual once per restart. (Checking the
        // previous step's flag instead, so that the next step is always
        // queued, costs an extra preconditioned step per solve, and the
        // coupled solves converge in 1-3 steps.)
        ArnoldiDev A = arnoldi_view(G);
        double *dh2 = arn_h2(G), *dny = arn_ny(G), *dwn2 = arn_wn2(G);
        const double tolb = tol * bnorm;
        for (int rs = 0; rs < max_restarts && status != MPS_OK; ++rs) {
            if (!std::isfinite(beta)) {
                status = MPS_ERR_NAN;
                break;
            }
            if (beta <= tolb) {
                status = MPS_OK;
                break;
            }
            // c->d_scalar still holds |b - A x|^2 from true_residual
            hipLaunchKernelGGL(k_arnoldi_begin, dim3(1), dim3(64), 0, c->stream, A, (const double *)c->d_scalar);
            scale_from(c, G.w, G.V[0], A.gv);
            int queued = 0;
            for (int j = 0; j < m; ++j) {
                precondition(c, G.V[j], G.Z[j]);
                matvec(c, c->L64, G.Z[j], G.w);
                // classical Gram-Schmidt with one re-orthogonalisation pass
                multidot_dev(c, G, G.w, G.V.data(), j + 1, G.dh);
                multiaxpy_dev(c, G.w, G.V.data(), j + 1, G.dh, nullptr, nullptr);
                multidot_dev(c, G, G.w, G.V.data(), j + 1, dh2);
                multiaxpy_dev(c, G.w, G.V.data(), j + 1, dh2, nullptr, dwn2);
                hipLaunchKernelGGL(k_arnoldi_step, dim3(1), dim3(64), 0, c->stream, A, j, kArnLd,
                                   (const double *)G.dh, (const double *)dh2, (const double *)dwn2, tolb);
                scale_from(c, G.w, G.V[j + 1], A.hn);
                HIPCHK(hipMemcpyAsync(G.hflag + j, A.state, sizeof(int), hipMemcpyDeviceToHost, c->stream));
                HIPCHK(hipEventRecord(G.ev[j], c->stream));
                queued = j + 1;
                HIPCHK(hipEventSynchronize(G.ev[j]));
                if (G.hflag[j]) break;
            }
            hipLaunchKernelGGL(k_arnoldi_solve, dim3(1), dim3(64), 0, c->stream, A, kArnLd, dny);
            multiaxpy_dev(c, u, G.Z.data(), queued, dny, A.state + 1, nullptr);
            HIPCHK(hipMemcpyAsync(G.hflag + kMaxVec, A.state, 3 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
            beta = true_residual();
            if (G.hflag[kMaxVec + 2]) {
                status = MPS_ERR_NAN;
                break;
            }
            total += G.hflag[kMaxVec + 1];
            if (beta <= tolb) status = MPS_OK;
        }
        if (iters) *iters = total;
        if (relres) *relres = bnorm > 0 ? beta / bnorm : beta;
        unbind_solution(c, x, staged);
        HIPCHK(hipStreamSynchronize(c->stream));
        if (status == MPS_ERR_NAN) return fail(MPS_ERR_NAN, "NaN or Inf in GMRES");
        if (status == MPS_NOT_CONVERGED)
            fail(MPS_NOT_CONVERGED, "GMRES: relative residual %g above %g", bnorm > 0 ? beta / bnorm : beta, tol);
        return status;
    } catch (const HipError &e) {
        return fail(MPS_ERR_HIP, "%s: %s", e.what, hipGetErrorString(e.e));
    }
}

}  // extern "C"

// ------------------------------------------------ no-allocation reference GMRES
// Work space of mps_gmres_noalloc: one array in the library's code object,
// so each GPU the library runs on holds its own copy from the first launch
// until the process ends. Its address differs per device; host code uses
// the address hipGetSymbolAddress returns on that device, never the
// array's name (a host-side stand-in for it).
constexpr long long kNoAllocWork = 1LL << 19;  // doubles: 4 MiB
static __device__ __align__(256) double scratch[kNoAllocWork];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {};

// the device address of `scratch` on device dev (the current device),
// looked up on the first call per device and kept
int scratch_on(int dev, double **out)
{
    if (dev < 0 || dev >= kMaxDevices) return fail(MPS_ERR_ARG, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(MPS_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch[dev] = static_cast<double *>(p);
    }
    *out = g_scratch[dev];
    return MPS_OK;
}

}  // namespace

extern "C" {

int mps_gmres_noalloc(mps_ctx *c, const double *b, double *x, int m, int max_restarts, double tol)
{
    int st = guard(c);
    if (st) return st;
    if (!b || !x || m < 1 || max_restarts < 1 || !(tol >= 0)) return fail(MPS_ERR_ARG, "bad arguments");
    if (!is_device_ptr(b) || !is_device_ptr(x)) return fail(MPS_ERR_ARG, "mps_gmres_noalloc takes device arrays");
    const Geom &g = c->geom[0];
    if (g.n() > kBottomMax)
        return fail(MPS_ERR_ARG, "mps_gmres_noalloc supports at most %d cells (grid has %lld)", kBottomMax, g.n());
    const long long need = bottom_work_elems<double>(m, g.ng());
    if (need > kNoAllocWork)
        return fail(MPS_ERR_ARG,
                    "GMRES(%d) on a %dx%dx%d grid needs %lld doubles of work space; the static work space holds "
                    "%lld",
                    m, g.nx, g.ny, g.nz, need, kNoAllocWork);
    double *work = nullptr;
    st = scratch_on(c->device, &work);
    if (st) return st;
    try {
        gmres_ref_launch(c, b, x, m, max_restarts, tol, work);
        HIPCHK(hipStreamSynchronize(c->stream));
    } catch (const HipError &e) {
        return fail(MPS_ERR_HIP, "%s: %s", e.what, hipGetErrorString(e.e));
    }
    return gmres_ref_status(c);
}

int mps_apply_transpose(mps_ctx *c, const double *x, double *atx)
{
    int st = guard(c);
    if (st) return st;
    if (!x || !atx) return fail(MPS_ERR_ARG, "null argument");
    try {
        ensure_transpose(c);
        const Geom &g = c->geom[0];
        const double *xd = x;
        if (!is_device_ptr(x)) {
            stage_in(c, x, c->user_tmp, g.ng());
            xd = c->user_tmp;
        }
        double *out = is_device_ptr(atx) ? atx : c->user_tmp2;
        Lev<double> v = view(c, c->LT, 0);
        hipLaunchKernelGGL((k_apply<double, OP_STENCIL, false, false>), cells_grid(g), block(), 0, c->stream, v, xd,
                           out, (double *)nullptr, (double *)nullptr);
        if (out != atx) HIPCHK(hipMemcpyAsync(atx, out, g.n() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    } catch (const HipError &e) {
        return fail(MPS_ERR_HIP, "%s: %s", e.what, hipGetErrorString(e.e));
    }
    return MPS_OK;
}

int mps_solve_transpose(mps_ctx *c, const double *gsrc, double *lambda, int *ncycles, double *resnorm)
{
    int st = guard_mg(c);
    if (st) return st;
    if (!gsrc || !lambda) return fail(MPS_ERR_ARG, "null argument");
    int status = MPS_NOT_CONVERGED;
    // A = A^T for the reference diffusion operator without convection on a
    // uniform finest grid: the adjoint solve is the forward solve with the
    // adjoint source (the V-cycle is then self-adjoint as well)
    const bool symmetric = c->opts.flux == MPS_FLUX_UPWIND && !c->has_vel && level_op(c, 0) == OP_REF;
    if (symmetric) {
        try {
            const Geom &g = c->geom[0];
            bool staged;
            double *u = bind_solution(c, lambda, staged);
            if (!c->gbuf) {
                c->gbuf = dalloc<double>(g.ng());
                fill<double>(c, c->gbuf, g.ng(), 0.0);
            }
            const double *gd = stage_interior(c, gsrc);
            Lev<double> v = view(c, c->L64, 0);
            hipLaunchKernelGGL(k_interior_in<double>, cells_grid(g), block(), 0, c->stream, v, gd, c->gbuf);
            HIPCHK(hipEventRecord(c->ev0, c->stream));
            double meas = 0.0;
            const int it = solve_core(c, u, c->gbuf, true, nullptr, meas);
            HIPCHK(hipEventRecord(c->ev1, c->stream));
            unbind_solution(c, lambda, staged);
            HIPCHK(hipStreamSynchronize(c->stream));
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
            c->last_ms = ms;
            c->last_cycles = it;
            if (ncycles) *ncycles = it;
            if (resnorm) *resnorm = meas;
            if (!std::isfinite(meas)) return fail(MPS_ERR_NAN, "residual is not finite");
            return meas <= c->opts.tol ? MPS_OK : MPS_NOT_CONVERGED;
        } catch (const HipError &e) {
            return fail(MPS_ERR_HIP, "%s: %s", e.what, hipGetErrorString(e.e));
        }
    }
    try {
        ensure_transpose(c);
        borrow_weights(c);
        const Geom &g = c->geom[0];
        bool staged;
        double *u = bind_solution(c, lambda, staged);
        auto &L = c->LT;
        L[0].u = u;
        const double *gd = stage_interior(c, gsrc);
        Lev<double> v = view(c, L, 0);
        fill<double>(c, L[0].b, g.ng(), 0.0);
        hipLaunchKernelGGL(k_interior_in<double>, cells_grid(g), block(), 0, c->stream, v, gd, L[0].b);
        HIPCHK(hipEventRecord(c->ev0, c->stream));
        double bn2 = 0.0;
        if (c->opts.tol_kind == MPS_TOL_REL_L2) {
            fill<double>(c, L[0].r, g.ng(), 0.0);
            residual_sumsq<double>(c, L, L[0].r, L[0].b, OP_STENCIL);
            bn2 = read_scalar(c);
        }
        int it = 0;
        double meas = 0.0;
        for (it = 0; it < c->opts.max_cycles;) {
            vcycle<double>(c, L, 0, c->bottom64, true);
            residual_sumsq<double>(c, L, u, L[0].b, OP_STENCIL);
            meas = measure(c, read_scalar(c), bn2);
            ++it;
            if (!std::isfinite(meas) || meas <= c->opts.tol) break;
        }
        HIPCHK(hipEventRecord(c->ev1, c->stream));
        unbind_solution(c, lambda, staged);
        HIPCHK(hipStreamSynchronize(c->stream));
        release_weights(c);
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        c->last_ms = ms;
        c->last_cycles = it;
        if (ncycles) *ncycles = it;
        if (resnorm) *resnorm = meas;
        if (!std::isfinite(meas)) return fail(MPS_ERR_NAN, "residual is not finite");
        status = meas <= c->opts.tol ? MPS_OK : MPS_NOT_CONVERGED;
    } catch (const HipError &e) {
        return fail(MPS_ERR_HIP, "%s: %s", e.what, hipGetErrorString(e.e));
    }
    return status;
}

}  // extern "C"

// ============================================================ measurement
namespace {

template <typename T>
int time_kernel_t(mps_ctx *c, std::vector<LevelBuf<T>> &L, int kernel, int reps, double *avg_ms, double *bytes,
                  std::string &name)
{
    const Geom &g = c->geom[0];
    const double w = sizeof(T);
    const double N = (double)g.n();
    const int op = level_op(c, L, 0);
    // coefficient fields read by the operator: D (+ 3 velocities + reac),
    // or the 7 stencil rows
    double cf = op == OP_STENCIL ? 7.0 : 1.0 + (c->has_vel ? 3.0 : 0.0) + (c->has_reac ? 1.0 : 0.0);
    T *u = L[0].u, *b = L[0].b;
    const bool bind64 = std::is_same<T, double>::value;
    if (bind64) {
        L[0].u = (T *)c->user_tmp;
        L[0].b = (T *)c->b0;
    }
    std::function<void()> launch;
    if (kernel == 0 && sweep2_ok(c, L, 0, op)) {
        // two red-black sweeps in one pass u -> u2: 2 (3 + 2c) w per cell
        launch = [&]() { launch_sweep2<T>(c, L, 0, L[0].u, L[0].u2, false); };
        *bytes = 2.0 * (3.0 + 2.0 * cf) * w * N;
        name = "k_sweep2_fast";
    } else if (kernel == 0 && fused_ok(c, L, 0, op)) {
        // one fused red+black sweep u -> u2 (both colours: (3 + 2c) w per cell)
        launch = [&]() { launch_sweep<T>(c, L, 0, L[0].u, L[0].u2, op); };
        *bytes = (3.0 + 2.0 * cf) * w * N;
        const bool ref = view(c, L, 0).fast && op == OP_REF && !c->has_reac && c->sweep_ref;
        name = ref ? "k_sweep_ref" : view(c, L, 0).fast ? "k_sweep_fast" : "k_sweep";
    } else if (kernel == 0) {
        launch = [&]() { launch_colour<T>(c, L, 0, 1, op); };
        *bytes = (1.5 + cf) * w * N;  // other colour u (1/2), own b (1/2), coefficients; own u write (1/2)
        name = "k_colour";
    } else if (kernel == 1) {
        launch = [&]() { launch_residual<T, T>(c, L, 0, L[0].u, L[0].b, L[0].r, nullptr, op); };
        *bytes = (cf + 3.0) * w * N;
        name = (view(c, L, 0).fast && op != OP_STENCIL && !c->has_vel) ? "k_residual_planes" : "k_residual";
    } else if (kernel == 2) {
        if (c->nlev() < 2) return fail(MPS_ERR_ARG, "no coarse level");
        const Geom &gc = c->geom[1];
        Lev<T> vf = view(c, L, 0), vc = view(c, L, 1);
        if (c->opts.hierarchy == MPS_HIER_REF) {
            launch = [&, vf, vc]() {
                hipLaunchKernelGGL(k_restrict_fas<T>, cells_grid(gc), block(), 0, c->stream, vf, vc,
                                   (const T *)L[0].u, (const T *)L[0].r, L[1].uold, L[1].u, L[1].b);
            };
            *bytes = (2.0 * N + 3.0 * (double)gc.n()) * w;
            name = "k_restrict_fas";
        } else {
            launch = [&]() { launch_restrict_fw<T>(c, L, 0); };
            *bytes = (N + (double)gc.n()) * w;
            name = "k_restrict_fw_planes";
        }
    } else if (kernel == 3) {
        if (c->nlev() < 2) return fail(MPS_ERR_ARG, "no coarse level");
        const Geom &gc = c->geom[1];
        Lev<T> vf = view(c, L, 0), vc = view(c, L, 1);
        if (c->opts.hierarchy == MPS_HIER_REF) {
            FacesT<T> G{};
            launch = [&, vf, vc, G]() {
                hipLaunchKernelGGL(k_prolong1_correct<T>, cells_grid(gc), block(), 0, c->stream, vf, vc, L[0].u,
                                   (const T *)L[1].u, (const T *)L[1].uold, G);
            };
            *bytes = (2.0 * N + 2.0 * (double)gc.n()) * w;
            name = "k_prolong1_correct";
        } else {
            launch = [&]() { launch_prolong_tri<T>(c, L, 0); };
            *bytes = (2.0 * N + (double)gc.n()) * w;
            name = "k_prolong_tri_planes";
        }
    } else {
        return fail(MPS_ERR_ARG, "unknown kernel %d", kernel);
    }
    // Each timed launch starts with cold caches: 1 GiB read in between (four
    // times the 256 MB Infinity Cache) evicts what the previous launch left
    // (its dirty lines are written back during the read, not during the
    // next timed launch), so a kernel whose working set would fit the cache
    // still reads it from HBM, as it does inside a solve. Launches are timed
    // one by one.
    constexpr long long kEvict4 = (1LL << 30) / 16;
    const long long per_block = (long long)kStreamBlock * kStreamUnroll;
    const unsigned nbe = (unsigned)((kEvict4 + per_block - 1) / per_block);
    float4 *evict = dalloc<float4>((size_t)kEvict4);
    double *epart = nullptr;
    double total = 0.0;
    try {
        epart = dalloc<double>(nbe);
        HIPCHK(hipMemsetAsync(evict, 0, (size_t)kEvict4 * 16, c->stream));
        launch();  // warm
        for (int q = 0; q < reps; ++q) {
            hipLaunchKernelGGL(k_stream<0>, dim3(nbe), dim3(kStreamBlock), 0, c->stream, (const float4 *)evict,
                               (const float4 *)nullptr, (const float4 *)nullptr, (float4 *)nullptr, kEvict4, 0.f,
                               epart);
            HIPCHK(hipEventRecord(c->ev0, c->stream));
            launch();
            HIPCHK(hipEventRecord(c->ev1, c->stream));
            HIPCHK(hipEventSynchronize(c->ev1));
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
            total += ms;
        }
    } catch (...) {
        dfree(evict);
        dfree(epart);
        throw;
    }
    dfree(evict);
    dfree(epart);
    *avg_ms = total / reps;
    if (bind64) {
        L[0].u = u;
        L[0].b = b;
    }
    return MPS_OK;
}

}  // namespace

extern "C" int mps_time_kernel(mps_ctx *c, int kernel, int fp32, int reps, double *avg_ms, double *alg_bytes,
                               char *name, int name_len)
{
    int st = guard_mg(c);
    if (st) return st;
    if (reps < 1 || !avg_ms || !alg_bytes) return fail(MPS_ERR_ARG, "bad arguments");
    if (fp32 && c->opts.precision != MPS_PREC_FP32_DC) return fail(MPS_ERR_ARG, "no FP32 hierarchy");
    try {
        std::string nm;
        st = fp32 ? time_kernel_t<float>(c, c->L32, kernel, reps, avg_ms, alg_bytes, nm)
                  : time_kernel_t<double>(c, c->L64, kernel, reps, avg_ms, alg_bytes, nm);
        if (name && name_len > 0) {
            std::strncpy(name, nm.c_str(), name_len - 1);
            name[name_len - 1] = 0;
        }
        return st;
    } catch (const HipError &e) {
        return fail(MPS_ERR_HIP, "%s: %s", e.what, hipGetErrorString(e.e));
    }
}

extern "C" int mps_time_in_solve(mps_ctx *c, int kernel, int fp32, double *soln, int *ncycles, double *avg_ms,
                                 int *launches)
{
    int st = guard_mg(c);
    if (st) return st;
    if (kernel < 0 || kernel > 3 || !soln || !avg_ms || !launches) return fail(MPS_ERR_ARG, "bad arguments");
    if (fp32 && c->opts.precision != MPS_PREC_FP32_DC) return fail(MPS_ERR_ARG, "no FP32 hierarchy");
    try {
        if (c->kt_ev.empty()) {
            c->kt_ev.resize(2 * kKtMax);
            for (auto &e : c->kt_ev) HIPCHK(hipEventCreate(&e));
        }
        HIPCHK(hipStreamSynchronize(c->stream));
        c->kt_kernel = kernel;
        c->kt_fp32 = fp32 != 0;
        c->kt_n = 0;
        // the solve's launches one by one (no graph replay), so that each
        // timed launch is bracketed by its own events on the stream
        const bool graphs = c->graphs;
        c->graphs = false;
        st = mps_solve(c, soln, ncycles, nullptr, nullptr);
        c->graphs = graphs;
        c->kt_kernel = -1;
        HIPCHK(hipStreamSynchronize(c->stream));
        if (st != MPS_OK && st != MPS_NOT_CONVERGED) return st;
        double total = 0.0;
        for (int q = 0; q < c->kt_n; ++q) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, c->kt_ev[2 * q], c->kt_ev[2 * q + 1]));
            total += ms;
        }
        *launches = c->kt_n;
        *avg_ms = c->kt_n ? total / c->kt_n : 0.0;
        return st;
    } catch (const HipError &e) {
        c->kt_kernel = -1;
        return fail(MPS_ERR_HIP, "%s: %s", e.what, hipGetErrorString(e.e));
    }
}

extern "C" int mps_time_stream(int kind, long long bytes_per_array, int reps, double *avg_ms, double *bytes)
{
    if (kind < 0 || kind > 4 || reps < 1 || bytes_per_array < 16 || !avg_ms || !bytes)
        return fail(MPS_ERR_ARG, "bad arguments");
    const long long n4 = bytes_per_array / 16;
    const int narr = kind == 0 ? 1 : kind == 1 ? 1 : kind == 2 ? 2 : kind == 3 ? 3 : 4;
    float4 *arr[4] = {};
    double *part = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int status = MPS_OK;
    try {
        HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        for (int q = 0; q < narr; ++q) {
            arr[q] = dalloc<float4>((size_t)n4);
            HIPCHK(hipMemsetAsync(arr[q], 0, (size_t)n4 * 16, st));
        }
        const long long per_block = (long long)kStreamBlock * kStreamUnroll;
        const unsigned nb = (unsigned)((n4 + per_block - 1) / per_block);
        part = dalloc<double>(nb);
        // inputs first, the output last: read (a), write (o), copy (a -> o),
        // triad (a, b -> o), sweep pattern (a, b, c -> o)
        const float4 *a = arr[0], *b = narr > 2 ? arr[1] : nullptr, *cc = narr > 3 ? arr[2] : nullptr;
        float4 *o = kind == 0 ? nullptr : arr[narr - 1];
        if (kind == 1) a = nullptr;
        auto launch = [&]() {
            switch (kind) {
            case 0: hipLaunchKernelGGL(k_stream<0>, dim3(nb), dim3(kStreamBlock), 0, st, a, b, cc, o, n4, 0.5f, part); break;
            case 1: hipLaunchKernelGGL(k_stream<1>, dim3(nb), dim3(kStreamBlock), 0, st, a, b, cc, o, n4, 0.5f, part); break;
            case 2: hipLaunchKernelGGL(k_stream<2>, dim3(nb), dim3(kStreamBlock), 0, st, a, b, cc, o, n4, 0.5f, part); break;
            case 3: hipLaunchKernelGGL(k_stream<3>, dim3(nb), dim3(kStreamBlock), 0, st, a, b, cc, o, n4, 0.5f, part); break;
            default: hipLaunchKernelGGL(k_stream<4>, dim3(nb), dim3(kStreamBlock), 0, st, a, b, cc, o, n4, 0.5f, part); break;
            }
        };
        launch();  // warm
        HIPCHK(hipEventRecord(e0, st));
        for (int q = 0; q < reps; ++q) launch();
        HIPCHK(hipEventRecord(e1, st));
        HIPCHK(hipEventSynchronize(e1));
        HIPCHK(hipGetLastError());
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        *avg_ms = ms / reps;
        *bytes = (double)narr * (double)n4 * 16.0;
    } catch (const HipError &e) {
        status = fail(MPS_ERR_HIP, "%s: %s", e.what, hipGetErrorString(e.e));
    }
    if (st) (void)hipStreamSynchronize(st);
    for (auto p : arr) dfree(p);
    dfree(part);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    return status;
}

extern "C" int mps_grad_velocity(mps_ctx *c, const double *lam, const double *x, double *gvx, double *gvy,
                                 double *gvz)
{
    int st = guard(c);
    if (st) return st;
    if (!lam || !x || !gvx || !gvy || !gvz) return fail(MPS_ERR_ARG, "null argument");
    if (!c->has_vel) return fail(MPS_ERR_STATE, "the operator has no velocity field");
    if (!is_device_ptr(lam) || !is_device_ptr(x) || !is_device_ptr(gvx) || !is_device_ptr(gvy) ||
        !is_device_ptr(gvz))
        return fail(MPS_ERR_ARG, "mps_grad_velocity takes device arrays");
    try {
        ExactScope exact(c);
        const Geom &g = c->geom[0];
        Lev<double> v = view(c, c->L64, 0);
        if (c->opts.flux == MPS_FLUX_SG)
            hipLaunchKernelGGL((k_velocity_grad<OP_REF, MPS_FLUX_SG_K>), cells_grid(g), block(), 0, c->stream, v, lam,
                               x, gvx, gvy, gvz);
        else
            hipLaunchKernelGGL((k_velocity_grad<OP_REF, MPS_FLUX_UPWIND_K>), cells_grid(g), block(), 0, c->stream, v,
                               lam, x, gvx, gvy, gvz);
        HIPCHK(hipStreamSynchronize(c->stream));
    } catch (const HipError &e) {
        return fail(MPS_ERR_HIP, "%s: %s", e.what, hipGetErrorString(e.e));
    }
    return MPS_OK;
}
