/* mps3d_amd: MI355X (gfx950) implicit field-solve library -- C ABI.
 *
 * Drop-in for the multigrid call surface of mps3d's mgridlib. Each entry
 * point names the reference routine it replaces (file:line in the
 * reference tree). Plain pointers and sizes only; callable from Fortran via
 * ISO_C_BINDING (mps3d_amd/fortran/mps3d_gpu.f90), from Python via ctypes
 * (mps3d_amd/_lib.py) or from C/C++.
 *
 * Layouts. Solution and coefficient arrays are the reference's ghosted
 * Fortran arrays, (0:nx+1, 0:ny+1, 0:nz+1) real*8, i fastest
 * (mgridlib/discretepde.f90:75-104). Interior arrays (mgridsterm, AX, diag,
 * offdiag, residual) are (nx,ny,nz). BC faces follow
 * mgridlib/globalvars.f90:57-66: left/right (ny,nz), bottom/top (nx,nz),
 * back/front (nx,ny). Any array pointer may be host or device memory; the
 * library detects which (hipPointerGetAttributes) and works in place on
 * device arrays.
 *
 * Status codes replace the reference's abort()/MPI_Abort paths
 * (mgridlib/gmres.f90:396-400, mgridlib/mgridsteps.f90:357-360).
 *
 * Threading: a context owns one HIP stream and its device buffers; it is not
 * re-entrant. Use one host thread per context.
 *
 * Ordering of device arrays: every call first orders the context's stream
 * after the work already queued on the null (default) stream, so device
 * arrays written on the null stream before a call need nothing more
 * (arrays produced on another non-blocking stream must be ordered before
 * the call by the caller, e.g. hipStreamWaitEvent on mps_stream(ctx)). A
 * call that reads or writes any host array returns with its results
 * complete. A call on device arrays only may return with its work still
 * queued on the context's stream (mps_set_bc, mps_set_coeffs,
 * mps_set_source, mps_fill_ghosts, mps_floor, mps_efield, the transport and
 * boundary-face entry points): later calls on the same context are ordered
 * after it; other readers order after mps_stream(ctx) (or synchronize it).
 * The solve entry points return with their results complete: their
 * stopping test reads the residual norm.
 */
#ifndef MPS3D_H
#define MPS3D_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define MPS_OK 0
#define MPS_NOT_CONVERGED 1   /* tolerance not met within max_cycles      */
#define MPS_ERR_NAN (-1)      /* NaN/Inf met (gmres.f90:113-120)          */
#define MPS_ERR_ARG (-2)      /* bad arguments or unsupported grid        */
#define MPS_ERR_HIP (-3)      /* HIP runtime error                         */
#define MPS_ERR_STATE (-4)    /* call out of order (e.g. no coefficients)  */
#define MPS_ERR_NOLIB (-5)    /* no usable gfx950 device                   */

/* ---- BC codes: character(4) 'DIRC','ZGRD','FLUX' of the reference;
 *      any other code is a no-flux face with no row term ---------------- */
enum { MPS_BC_NONE = 0, MPS_BC_DIRC = 1, MPS_BC_ZGRD = 2, MPS_BC_FLUX = 3 };

/* ---- solver forms ----------------------------------------------------- */
enum {
    /* the reference FAS V-cycle: 8-child mean restriction, prolongation1,
     * rediscretised coarse operators, GMRES bottom solve on the coarse grid
     * fixed by the process grid (mgridsteps.f90:384-579). Needs per-rank
     * sizes nx/px.. that halve evenly down to 1. */
    MPS_HIER_REF = 0,
    /* cell-centred grids of m*2^k+1 cells: coarse cells sit on every other
     * fine cell centre (65 -> 33 -> 17 -> 9 -> 5 -> 3), volume-weighted
     * full-weighting restriction, trilinear prolongation, correction scheme,
     * direct bottom solve. No reference counterpart (the reference cannot
     * coarsen odd sizes, mgridsteps.f90:21-23). */
    MPS_HIER_ALIGNED = 1
};
enum {
    MPS_PREC_FP64 = 0,    /* every kernel in FP64 (reference arithmetic)   */
    MPS_PREC_FP32_DC = 1  /* FP32 smoothing/restriction/prolongation inside
                             an FP64 defect-correction loop; the residual is
                             always evaluated in FP64                        */
};
enum {
    MPS_FLUX_UPWIND = 0,  /* reference face fluxes: arithmetic-mean D,
                             first-order upwind convection
                             (discretepde.f90:620-670)                      */
    MPS_FLUX_SG = 1       /* Scharfetter-Gummel exponential fluxes with the
                             harmonic-mean D of mps1d/source/convdiff.f90:10-64,
                             applied per face in 3D                         */
};
enum {
    MPS_ARITH_EXACT = 0,  /* rows evaluated in the reference's order and
                             rounding (mgridlib/discretepde.f90:620-1119)   */
    MPS_ARITH_FAST = 1    /* the same operator with precomputed face factors
                             and one reciprocal per row, for rows without
                             convection; rows with convection stay exact.
                             The find_AX/find_bvec entry points are always
                             exact; mps_smooth follows the context's mode  */
};
enum {
    MPS_TOL_RMS = 0,      /* RMS of the absolute residual (globalvars.f90:136-159) */
    MPS_TOL_REL_L2 = 1    /* ||b - A u||_2 / ||b||_2                              */
};
enum {
    MPS_START_GUESS = 0,  /* every iteration of mps_solve is one V-cycle (or
                             defect-correction step) from the current iterate,
                             as the reference's caller loops                */
    MPS_START_FMG = 1     /* MPS_HIER_ALIGNED only: the first iteration is a
                             full-multigrid pass on the initial defect (its
                             right-hand side restricted to every level,
                             coarse-to-fine nested iteration with trilinear
                             interpolation and one V-cycle per level), the
                             rest are V-cycles. No reference counterpart.  */
};

typedef struct {
    int nx, ny, nz;             /* global cells (g_Nx, g_Ny, g_Nz)             */
    double lenx, leny, lenz;    /* domain lengths (g_xlen, g_ylen, g_zlen)     */
    int px, py, pz;             /* process grid (g_px..): fixes the coarsest
                                   level of MPS_HIER_REF; 1,1,1 otherwise      */
} mps_grid;

typedef struct {
    int hierarchy;              /* MPS_HIER_*                                  */
    int precision;              /* MPS_PREC_*                                  */
    int flux;                   /* MPS_FLUX_*                                  */
    int nu1, nu2;               /* pre/post RB-GS sweeps (reference: 2, 2,
                                   mgridsteps.f90:473)                         */
    int max_cycles;             /* V-cycles (or defect-correction steps) per
                                   mps_solve (reference loops: 20)             */
    double tol;                 /* stopping tolerance for mps_solve            */
    int tol_kind;               /* MPS_TOL_*                                   */
    int arith;                  /* MPS_ARITH_* for the solver's own rows       */
    int start;                  /* MPS_START_*                                 */
    int dc_inner;               /* FP32 V-cycles per FP64 defect-correction
                                   step under MPS_PREC_FP32_DC, 1..4, or 0
                                   (default): chosen per step, the fewest
                                   that reach tol at the contraction the
                                   previous step showed, at most 4 and no
                                   more than keep a step's reduction above
                                   1e-5; the first step assumes 0.1 per
                                   cycle, a step after one that did not
                                   contract takes 2                            */
    double omega;               /* over-relaxation of the red-black sweeps of
                                   mps_solve's V-cycles in MPS_ARITH_FAST:
                                   u = u_old + omega (u_GS - u_old); 0
                                   (default) = 1.15 on MPS_HIER_ALIGNED, 1 on
                                   MPS_HIER_REF; else 0 < omega < 2. Exact
                                   arithmetic, mps_smooth and mps_vcycle
                                   always sweep with 1 (the reference's
                                   Gauss-Seidel)                               */
} mps_opts;

typedef struct mps_ctx mps_ctx;

/* Fills reference defaults: REF hierarchy, FP64, upwind, nu1=nu2=2,
 * max_cycles=20, tol=1e-11 RMS (globalvars.f90:29), exact arithmetic,
 * V-cycles from the caller's guess. */
void mps_default_opts(mps_opts *opts);

/* Allocates a context for one grid on the current HIP device. Replaces the
 * setup of domaindecompose (par_decompose.f90:7-85): the whole domain lives
 * on one GPU and the halo exchange becomes in-device ghost layers. */
int mps_create(mps_ctx **ctx, const mps_grid *grid, const mps_opts *opts);
int mps_destroy(mps_ctx *ctx);

/* Boundary conditions: codes[6] (MPS_BC_*) and six face arrays (NULL = 0),
 * as bc_codes/bcvals of perform_vcycle (mgridsteps.f90:392-393). */
int mps_set_bc(mps_ctx *ctx, const int codes[6], const double *const faces[6]);
/* Same, with the Fortran character(4) codes packed as 24 chars. */
int mps_set_bc_str(mps_ctx *ctx, const char codes[24], const double *const faces[6]);

/* Operator coefficients, ghosted layout, NULL = 0: dcoeff, vel (x,y,z),
 * reac, source, and the time-derivative flag and step of perform_vcycle's
 * argument list (mgridsteps.f90:384-415). Coarse-level coefficients are
 * derived here once per call (the reference re-restricts them on every
 * V-cycle, mgridsteps.f90:533-539). */
int mps_set_coeffs(mps_ctx *ctx, const double *dcoeff, const double *velx, const double *vely,
                   const double *velz, const double *reac, const double *source, int timederiv,
                   double dt);

/* The source of the last mps_set_coeffs replaced, the other coefficients
 * and the hierarchy built from them kept: the Poisson equation's per-step
 * update (poisson_update_transport, poissoneq.f90:335-385, changes only its
 * space-charge source; the coarse levels use no source,
 * discretepde.f90:611-616). source ghosted, NULL = 0. */
int mps_set_source(mps_ctx *ctx, const double *source);

/* perform_vcycle (mgridsteps.f90:384-579) on the finest level: one V-cycle
 * on soln (ghosted, in/out) with multigrid source mgridsterm (interior,
 * NULL = 0); resnorm gets the RMS residual after the cycle. Ghost values of
 * soln are left as the caller set them, as in the reference. */
int mps_vcycle(mps_ctx *ctx, double *soln, const double *mgridsterm, double *resnorm);

/* The solve loop of the equation modules (poissoneq.f90:427-450,
 * pdedefs.f90:378-402): up to opts.max_cycles V-cycles (or FP32
 * defect-correction steps), stopping when the opts.tol_kind residual measure
 * <= opts.tol. history (nullable, max_cycles entries) gets the measure after
 * each cycle. Returns MPS_OK or MPS_NOT_CONVERGED. */
int mps_solve(mps_ctx *ctx, double *soln, int *ncycles, double *resnorm, double *history);

/* The solve of mps_solve for a right-hand side given directly: A x = b with
 * b the whole right-hand side (interior array: the source and any boundary
 * terms, as find_bvec returns them) and homogeneous boundary data in the
 * cycle; the context's source and faces are not used. For callers whose
 * right-hand side changes from call to call while the coefficients stay,
 * e.g. the Poisson solve of every time step (poisson_solve's loop,
 * poissoneq.f90:427-450, with only the space charge changed): no
 * coefficient set-up per call. Stops as mps_solve. */
int mps_solve_rhs(mps_ctx *ctx, const double *b, double *soln, int *ncycles, double *resnorm,
                  double *history);

/* find_AX (discretepde.f90:672-1119): AX, diag and offdiag (interior arrays,
 * each nullable) for the ghosted input x. */
int mps_apply(mps_ctx *ctx, const double *x, double *ax, double *diag, double *offdiag);

/* Transposed operator: ATX = A^T x (interior output) for the adjoint. */
int mps_apply_transpose(mps_ctx *ctx, const double *x, double *atx);

/* find_bvec (discretepde.f90:382-618): b = BC terms + mgridsterm + source
 * (BC and source terms only when applybc != 0). */
int mps_bvec(mps_ctx *ctx, const double *mgridsterm, int applybc, double *b);

/* perform_gseidel_smoothing (discretepde.f90:132-236) on the finest level:
 * niter red-black sweeps, then res = b - A soln (interior, nullable), in the
 * context's arithmetic mode (bit-exact with the reference in EXACT). */
int mps_smooth(mps_ctx *ctx, double *soln, const double *mgridsterm, int niter, double *res);

/* restriction (mgridsteps.f90:11-53): coarse (nx/2, ny/2, nz/2) = the
 * 8-child mean of fine (nx, ny, nz), both interior arrays, in the
 * reference's summation order, by the kernel the reference-hierarchy
 * V-cycle runs. The caller loops use it for output as well
 * (plasma_pde/solve_manager.f90:573). Needs even nx, ny, nz. */
int mps_restrict(mps_ctx *ctx, const double *fine, double *coarse);

/* prolongation1 (mgridsteps.f90:170-216): fine (nx, ny, nz interior, out) =
 * the central-slope interpolation of coarse, a ghosted
 * (nx/2+2, ny/2+2, nz/2+2) array whose ghost layer is read as given, by
 * the kernel of the reference-hierarchy V-cycle's correction step. Needs
 * even nx, ny, nz. */
int mps_prolong1(mps_ctx *ctx, const double *coarse, double *fine);

/* update_boundary_ghostvalues (discretepde.f90:1121-1267): DIRC ghost =
 * 2 g - u, other codes ghost = u, on the six faces of soln. */
int mps_fill_ghosts(mps_ctx *ctx, double *soln);

/* update_efield (plasma_pde/poissoneq.f90:565-680): the ghost values of phi
 * (ghosted, in/out) from the context's boundary conditions as
 * update_boundary_ghostvalues sets them, then the cell-centre field
 * E = -scale * 0.5 (phi(+1) - phi(-1)) / d per direction, with the
 * reference's minmod limiter (poissoneq.f90:549-563, 613-656) when limiter
 * != 0 (the reference's default is off), and zero-gradient ghost faces of
 * each component. efield is the reference's efield(0:nx+1,0:ny+1,0:nz+1,3)
 * (three ghosted components one after the other; ghost edges and corners
 * are not written, as in the reference); scale = potscale / lscale. */
int mps_efield(mps_ctx *ctx, double *phi, double *efield, double scale, int limiter);

/* ---- Transport assembly (the step before and after each solve) ----------
 * Per-cell coefficients and sources of the plasma equations, computed on the
 * device in the ghosted layout mps_set_coeffs takes, for chemistries given
 * as tables: Arrhenius rates in T_e, and per species a constant mobility and
 * diffusion (plasma_pde/chemistry/ion_elec_chem.f90) or the field-dependent
 * mobility fits and Einstein diffusion of H2_O2_chem.f90. Inputs and
 * outputs are host or device arrays (host arrays are staged for the call);
 * numden is the reference's numden(0:nx+1,0:ny+1,0:nz+1,nspecies), efield
 * and vel three ghosted components one after the other. scaling[4] =
 * {vel, dcoeff, reac, source} scalings of the equation module
 * (edeneq.f90:82-85: tscale/lscale, tscale/lscale^2, tscale,
 * tscale/varscale); with timederiv != 0, source += soln/dt over the whole
 * ghosted array, as the reference does. Output pointers are nullable. The
 * reference's physical constants (plasma_pde/fundconstants.f90:3-9) apply. */
typedef struct {
    int nspecies, nreac;
    int especnum, bgspecnum;        /* 1-based species numbers */
    int ionspecmin, ionspecmax;     /* 1-based; ionspecmax < ionspecmin: no ions */
    const int *reactants;           /* reactants(nreac, nspecies), Fortran order */
    const int *products;            /* products(nreac, nspecies), Fortran order */
    const double *k_arrh;           /* k_arrh(3, nreac): A, alpha, Ea (K) */
    const double *elecenergy;       /* (nreac) electron energy change, eV */
    const double *mobility;         /* (nspecies) signed mobility */
    const double *dcoeff;           /* (nspecies) diffusion coefficient */
    const double *molmass;          /* (nspecies) kg */
    const double *charge;           /* (nspecies) charge number */
    /* Field-dependent transport (getspecmobility / getspecdcoeff of
     * plasma_pde/chemistry/H2_O2_chem.f90:532-672); all four NULL = the
     * constants above. mob_kind(s): 0 mobility(s); 1 electron fit
     * mobility(s) / sum_j n_j / exp(a_j L^2 + b_j L + c_j), L = ln(T_e / 11604.5 K);
     * 2 reduced ion mobility mobility(s) N_atm / sum_j n_j / K_j, N_atm =
     * 101325 Pa / k_B / 300 K. The sums run over the partners j with
     * mob_terms(j, s) != 0, in species order, with mob_coef(1:3, j, s) =
     * (a, b, c) or (K, -, -). dcoeff_kind(s): 0 dcoeff(s); 1 k_B T_e mu /
     * (Z e); 2 k_B T_g mu / (Z e). */
    const int *mob_kind;            /* (nspecies) */
    const int *mob_terms;           /* (nspecies, nspecies), Fortran order */
    const double *mob_coef;         /* (3, nspecies, nspecies), Fortran order */
    const int *dcoeff_kind;         /* (nspecies) */
    /* Reaction rates (getspecproduction / getelectroninelasticterm with
     * isratearrh and getcustomrate); both NULL = every reaction Arrhenius
     * (getarrhrate from k_arrh). rate_kind(r): 0 Arrhenius; 1 the fit of
     * O2_chem.f90:413-448, c0 exp(c1/T + c2/T**2 + c3/T**3 + c4/T**4)
     * with rate_coef(1:5, r) = (c0, c1, c2, c3, c4), signed; 2
     * alpha |E| exp(-E0/|E|) of H2_chem.f90:423-437 with rate_coef(1:2, r)
     * = (alpha, E0) and |E| the cell's field magnitude. T is T_e. */
    const int *rate_kind;           /* (nreac) */
    const double *rate_coef;        /* (5, nreac), Fortran order */
} mps_chem;

/* initializechemistry (ion_elec_chem.f90:104-109, H2_O2_chem.f90:433-438,
 * with the tables of assignreactions and setspecparams): copies the tables
 * to the device (at most 16 species, 64 reactions, stoichiometry 0..16). */
int mps_set_chemistry(mps_ctx *ctx, const mps_chem *chem);

/* eden_update_transport (edeneq.f90:360-444) for specnum = especnum, with
 * prod = elecprod (interior, unscaled production), and ions_update_transport
 * (ionseq.f90:478-558) for an ion's specnum (prod NULL): vel = mu E, dcoeff =
 * D, source = getspecproduction (ion_elec_chem.f90:111-156), all scaled;
 * reac = 0. Pg is accepted for the reference's argument list; the table
 * chemistries do not use it. */
int mps_species_transport(mps_ctx *ctx, int specnum, double dt, int timederiv, const double *efield,
                          const double *electemp, const double *numden, double Tg, double Pg,
                          const double scaling[4], const double *soln, double *vel, double *dcoeff,
                          double *reac, double *source, double *prod);

/* eenergy_update_transport (eenergyeq.f90:679-774): vel = 5/3 mu_e E,
 * dcoeff = 5/3 D_e, source = Joule heating (jouleheatingterm, upwinded face
 * fluxes per axis clipped at zero, :394-555) - elastic loss (elastic_colterm,
 * :557-612) + inelastic term (inelastic_colterm, :614-677), all scaled.
 * lscale turns the context's cell widths into the Joule term's dx. */
int mps_eenergy_transport(mps_ctx *ctx, double dt, int timederiv, double lscale, const double *efield,
                          const double *electemp, const double *numden, double Tg, double Pg,
                          const double scaling[4], const double *soln, double *vel, double *dcoeff,
                          double *reac, double *source);

/* poisson_update_transport (poissoneq.f90:335-385): source = e/eps0
 * (sum_ions Z n - n_e), dcoeff = 1, both scaled; vel = reac = 0. iondens
 * holds the ion species ionspecmin..ionspecmax one after the other. */
int mps_poisson_transport(mps_ctx *ctx, double dt, int timederiv, const double *elecden,
                          const double *iondens, const double scaling[4], const double *soln,
                          double *vel, double *dcoeff, double *reac, double *source);

/* The FLUX faces of ions_update_bcs (ionseq.f90:247-476) for ion species
 * specnum (1-based): on the boundary cells of each face whose pointer is not
 * null, the drift outflow mu n E_normal velscale where mu E_normal points
 * out of the domain, else 0. efield: 3 ghosted components; numden: every
 * species, ghosted, for the mobility (getspecmobility: constant or reduced
 * ion mobility); ionsoln: the ion's scaled solution, ghosted. faces[6]:
 * left/right (ny,nz), bottom/top (nx,nz), back/front (nx,ny), first index
 * fastest, as mps_set_bc takes them. */
int mps_ion_drift_faces(mps_ctx *ctx, int specnum, const double *efield, const double *numden,
                        const double *ionsoln, double velscale, double *const faces[6]);

/* The FLUX faces of eden_update_bcs (edeneq.f90:203-358): on the faces
 * whose code (codes[6], the equation's MPS_BC_*) is FLUX, the thermal
 * outflow 0.25 n_e cbar velscale, cbar = sqrt(8 k_B T_e / (PI m_e)), from
 * the boundary cells of elecden (scaled) and electemp (ghosted); faces of
 * other codes, and null face pointers, are left as they are. faces[6] in
 * the layout mps_set_bc takes. */
int mps_eden_faces(mps_ctx *ctx, const int codes[6], const double *elecden, const double *electemp,
                   double velscale, double *const faces[6]);

/* The FLUX faces of neutrals_update_bcs (neutralseq.f90:206-358): 0.25 n
 * cbar velscale with cbar = sqrt(8 k_B Tg / (PI molmass)). */
int mps_neutral_faces(mps_ctx *ctx, const int codes[6], const double *neutralden, double molmass, double Tg,
                      double velscale, double *const faces[6]);

/* eenergy_update_bcs (eenergyeq.f90:208-392): DIRC faces
 * 1.5 k_B n_e bcparams[f] / eenergyscale, FLUX faces 4/3 n_eps cbar
 * velscale with cbar from T = n_eps eenergyscale / (1.5 n_e k_B);
 * elecenergy scaled, elecden physical, both ghosted. */
int mps_eenergy_faces(mps_ctx *ctx, const int codes[6], const double *elecenergy, const double *elecden,
                      double eenergyscale, const double bcparams[6], double velscale, double *const faces[6]);

/* The end of eden_solve (edeneq.f90:509-523): interior values of soln below
 * vmin become vmin, then the ghost values from the context's boundary
 * conditions (update_boundary_ghostvalues). */
int mps_floor(mps_ctx *ctx, double *soln, double vmin);

/* update_electrontemp (eenergyeq.f90:842-877): electemp = eenergysoln *
 * eenergyscale / (3/2 n_e k_B) on the interior; below etempmin the
 * temperature is etempmin and eenergysoln (in/out) is reset to match, in the
 * same cell (the reference's reset lands one cell off, eenergyeq.f90:868). */
int mps_electron_temp(mps_ctx *ctx, double *eenergysoln, const double *elecden, double eenergyscale,
                      double etempmin, double *electemp);

/* performgmres (gmres.f90:197-434) on the finest level: restarted GMRES(m)
 * for A x = b (b interior, x ghosted in/out).
 *   precond = 0: the reference's method, left-preconditioned by one
 *     red-black sweep from zero (gseidel_precond), classical Gram-Schmidt,
 *     converged when the preconditioned residual <= tol relative to
 *     ||M^-1 b||; one workgroup, at most 512 cells (coarse grids).
 *   precond = 1: flexible GMRES(m), right-preconditioned by one V-cycle
 *     (the FP32 V-cycle under MPS_PREC_FP32_DC, Arnoldi and Givens in FP64),
 *     Gram-Schmidt with one re-orthogonalisation pass; converged when
 *     ||b - A x|| <= tol ||b|| (the Givens estimate inside a cycle, as the
 *     in-loop exit of mps1d/source/gmres.f90:343, confirmed by the true
 *     residual at its end). The m + 1 Krylov and m preconditioned vectors
 *     are allocated on the first call with a given m and kept by the
 *     context until mps_destroy.
 * iters and relres are nullable. */
int mps_gmres(mps_ctx *ctx, const double *b, double *x, int m, int max_restarts, double tol,
              int precond, int *iters, double *relres);

/* mps_gmres with precond = 0 (performgmres, gmres.f90:197-434, as above)
 * that allocates nothing: no reference counterpart. For callers that cannot
 * call hipMalloc (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation). b (interior) and x (ghosted, in/out)
 * must be device arrays of the context's grid. The GMRES work space is a
 * static device array of 2^19 doubles (4 MiB) per GPU, held from the first
 * launch until the process ends: GMRES(m) on the grid must fit in it
 * ((m + 9) ghosted vectors plus (m + 1)(m + 4) scalars; e.g. m = 64 on 8^3
 * cells), else MPS_ERR_ARG. Every call on one GPU shares that array, so
 * calls that overlap on one GPU (contexts on different host threads or
 * streams) race; calls in sequence are independent (the kernel initialises
 * every part of the work space it reads). Results equal mps_gmres's with
 * precond = 0 bit for bit. */
int mps_gmres_noalloc(mps_ctx *ctx, const double *b, double *x, int m, int max_restarts, double tol);

/* Adjoint solve: A^T lambda = g (g interior, lambda ghosted out) with the
 * transposed operator and the same hierarchy; stops as mps_solve. */
int mps_solve_transpose(mps_ctx *ctx, const double *g, double *lambda, int *ncycles,
                        double *resnorm);

/* Adjoint sensitivity to the cell velocities (no reference counterpart):
 * gv_d(c) = -d(lam^T A(v) x)/dv_d(c) for d = x, y, z (interior outputs),
 * with lam and x ghosted. For x = A^-1 b and lam = A^-T dJ/dx this is dJ/dv
 * through the operator (upwind or SG fluxes, boundary rows included). */
int mps_grad_velocity(mps_ctx *ctx, const double *lam, const double *x, double *gvx, double *gvy,
                      double *gvz);

/* The grid a context was created for and the number of species of its
 * chemistry (0 before mps_set_chemistry; nspecies nullable). No reference
 * counterpart: callers that size arrays themselves (the Fortran wrappers,
 * whose arrays are dimensioned from lx, ly, lz) check them against the
 * context before passing them. */
int mps_get_grid(mps_ctx *ctx, mps_grid *grid, int *nspecies);

/* Timing of the last mps_solve / mps_solve_transpose on the context's stream
 * (HIP events): total milliseconds and per-cycle milliseconds. */
int mps_last_timing(mps_ctx *ctx, double *total_ms, int *ncycles);

/* V-cycles of the last solve on the context (no reference counterpart, whose
 * solves count V-cycles only): under MPS_PREC_FP32_DC the FP32 V-cycles of
 * all its defect-correction steps, which differ from step to step when
 * opts.dc_inner is 0; otherwise the V-cycles. A full-multigrid start counts
 * as one. For callers that report work per solve. */
int mps_last_vcycles(mps_ctx *ctx, int *nvcycles);

/* Measurement hook (no reference counterpart): times `reps` launches of one
 * hot kernel of the current hierarchy with HIP events on the context's
 * stream, after the context has solved once. Each launch is timed alone,
 * after 1 GiB of a scratch buffer (allocated for the call) has been read,
 * so it starts with the Infinity Cache holding none of its data, as inside
 * a solve.
 *   kernel 0  smoothing on the finest level: the two red-black sweeps of a
 *             smoothing step in one pass where that kernel applies (FP32
 *             rows from one coefficient field without convection or
 *             reaction), else one fused red-black sweep where the operator
 *             has no convection term, else one colour pass
 *   kernel 1  residual b - A u on the finest level
 *   kernel 2  restriction of the finest residual to level 1
 *   kernel 3  prolongation + correction from level 1 to the finest level
 * fp32 selects the FP32 inner hierarchy of MPS_PREC_FP32_DC. avg_ms gets
 * the mean launch time, alg_bytes the algorithmic HBM bytes of one launch
 * (each input and output element counted once), name (nullable, name_len
 * bytes) the kernel's symbol for matching profiler output. */
int mps_time_kernel(mps_ctx *ctx, int kernel, int fp32, int reps, double *avg_ms, double *alg_bytes,
                    char *name, int name_len);

/* Measurement hook (no reference counterpart): runs mps_solve on soln (as
 * mps_solve; ncycles nullable) with every launch of one hot kernel on the
 * finest level timed by its own pair of HIP events on the context's stream,
 * and returns the mean launch time and the number of launches timed (at
 * most 256). kernel: 0 the red-black sweep (mps_time_kernel's kernel 0
 * where it applies; sweeps from a zero guess, which read no iterate, are
 * not timed), 1 the residual (for MPS_PREC_FP32_DC with fp32 = 0 the FP64
 * defect), 2 the V-cycle's restriction from the finest level (with the
 * residual where the two are one kernel), 3 the V-cycle's prolongation to
 * the finest level;
 * fp32 selects the FP32 inner hierarchy. The solve replays no graphs while
 * timed. Returns mps_solve's status. */
int mps_time_in_solve(mps_ctx *ctx, int kernel, int fp32, double *soln, int *ncycles, double *avg_ms, int *launches);

/* Measurement hook (no reference counterpart): the HBM rate a hand-written
 * streaming kernel reaches on the current device, 16-byte accesses per lane,
 * arrays of bytes_per_array each (allocated for the call; take them well
 * above the 256 MB Infinity Cache), timed over `reps` launches with HIP
 * events. kind 0 read, 1 write, 2 copy, 3 triad o = a + s b, 4 the
 * red-black sweep's pattern o = a + s (b + c) (three streams read, one
 * written). avg_ms gets the mean launch time, bytes the bytes one launch
 * reads and writes. */
int mps_time_stream(int kind, long long bytes_per_array, int reps, double *avg_ms, double *bytes);

/* The context's HIP stream (hipStream_t), for callers that order work
 * against it. */
void *mps_stream(mps_ctx *ctx);

const char *mps_error_string(int status);
/* Text of the most recent failure on this thread. */
const char *mps_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* MPS3D_H */
