/* CPU restatement of the mps3d reference multigrid path (TEST INFRASTRUCTURE).
 *
 * This library is the checker for the HIP implementation: only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg load it. The
 * product path never calls it.
 *
 * It restates, on one global cell-centred grid, the algorithm of
 *   mgridlib/discretepde.f90   (find_AX, find_bvec, perform_gseidel_smoothing,
 *                               gseidel_precond, update_boundary_ghostvalues)
 *   mgridlib/mgridsteps.f90    (restriction, restrict_boundary, prolongation1,
 *                               bottomsolve, perform_vcycle)
 *   mgridlib/gmres.f90         (performgmres and helpers)
 *   mgridlib/globalvars.f90    (compute_norm)
 *   mps1d/source/convdiff.f90  (find_sg_exponential_flux)
 * with the MPI block decomposition replaced by a "virtual process grid"
 * (px,py,pz) that fixes the coarsest multigrid level exactly as the
 * reference's rank layout does (mgridsteps.f90:512, :301-303).
 *
 * Arrays are Fortran-ordered with one ghost layer: element (i,j,k),
 * 0 <= i <= nx+1, is at i + (nx+2)*(j + (ny+2)*k). Interior-only arrays
 * (AX, diag, residual, mgridsterm) are nx*ny*nz, i fastest.
 */
#ifndef MPS_ORACLE_H
#define MPS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* BC codes (mgridlib/discretepde.f90 string tests 'DIRC','FLUX','ZGRD';
 * anything else acts as a no-flux face with no row term). */
enum { ORC_BC_NONE = 0, ORC_BC_DIRC = 1, ORC_BC_ZGRD = 2, ORC_BC_FLUX = 3 };

typedef struct {
    int nx, ny, nz;          /* global cells */
    int px, py, pz;          /* virtual process grid */
    double lenx, leny, lenz; /* domain lengths */
    int bc[6];               /* left,right,bottom,top,back,front */
    const double *bcv[6];    /* face arrays (ny,nz),(ny,nz),(nx,nz),(nx,nz),(nx,ny),(nx,ny) */
    const double *D, *vx, *vy, *vz, *reac, *source; /* ghosted; NULL = 0 */
    int timederiv;
    double dt;
} orc_problem;

/* find_AX on the finest level: AX, diag, offdiag (interior arrays). */
void orc_apply(const orc_problem *p, const double *u, double *ax, double *diag, double *off);
/* find_bvec on the finest level. */
void orc_bvec(const orc_problem *p, const double *sterm, int applybc, double *b);
/* perform_gseidel_smoothing on the finest level (applybcflag = true). */
void orc_smooth(const orc_problem *p, double *u, const double *sterm, int niter, double *res);
/* restriction (8-child mean): fine interior (nx,ny,nz) -> coarse (nx/2,..) */
void orc_restrict(int nx, int ny, int nz, const double *fine, double *coarse);
/* prolongation1: coarse ghosted (cx+2)(cy+2)(cz+2) -> fine interior (2cx,2cy,2cz) */
void orc_prolong1(int cx, int cy, int cz, const double *coarse_g, double *fine);
/* update_boundary_ghostvalues with the problem's BC codes and face arrays. */
void orc_ghosts(const orc_problem *p, double *u);

/* update_efield, plasma_pde/poissoneq.f90:565-680 (efield: 3 ghosted components) */
void orc_efield(const orc_problem *p, double *phi, double *efield, double scale, int limiter);
/* Chemistry tables of a chem_module (plasma_pde/chemistry/ion_elec_chem.f90:
 * reactants/products(nreac, nspecies) and k_arrh(3, nreac) in Fortran
 * order, 1-based species numbers). */
typedef struct {
    int nspecies, nreac;
    int especnum, bgspecnum, ionspecmin, ionspecmax;
    const int *reactants, *products;
    const double *k_arrh, *elecenergy;
    const double *mobility, *dcoeff, *molmass, *charge;
    /* field-dependent transport of H2_O2_chem.f90:532-672 (NULL: constants):
     * mob_kind (nspecies), mob_terms (nspecies, nspecies), mob_coef
     * (3, nspecies, nspecies), dcoeff_kind (nspecies), as in mps_chem */
    const int *mob_kind, *mob_terms;
    const double *mob_coef;
    const int *dcoeff_kind;
    /* reaction rates (NULL: every reaction getarrhrate): rate_kind (nreac)
     * 0 Arrhenius, 1 getcustomrate of O2_chem.f90:413-448, 2 getcustomrate
     * of H2_chem.f90:423-437; rate_coef (5, nreac), as in mps_chem */
    const int *rate_kind;
    const double *rate_coef;
} orc_chem;

/* Outputs of an *_update_transport: vel (3 ghosted components), dcoeff,
 * reac, source (ghosted), prod (interior cells of a ghosted array). */
typedef struct {
    double *vel, *dcoeff, *reac, *source, *prod;
} orc_transport_out;

/* eden_update_transport (plasma_pde/edeneq.f90:360-444) and
 * ions_update_transport (plasma_pde/ionseq.f90:478-558) for species specnum;
 * scaling = {vel, dcoeff, reac, source}. */
void orc_species_transport(const orc_problem *p, const orc_chem *ch, int specnum, double dt, int timederiv,
                           const double *efield, const double *electemp, const double *numden, double Tg,
                           const double *scaling, const double *soln, orc_transport_out out);
/* eenergy_update_transport (plasma_pde/eenergyeq.f90:394-774). */
void orc_eenergy_transport(const orc_problem *p, const orc_chem *ch, double dt, int timederiv, double lscale,
                           const double *efield, const double *electemp, const double *numden, double Tg,
                           const double *scaling, const double *soln, orc_transport_out out);
/* poisson_update_transport (plasma_pde/poissoneq.f90:335-385). */
void orc_poisson_transport(const orc_problem *p, const orc_chem *ch, double dt, int timederiv,
                           const double *elecden, const double *iondens, const double *scaling,
                           const double *soln, orc_transport_out out);
/* update_electrontemp (plasma_pde/eenergyeq.f90:842-877), the clamp's
 * energy reset in the clamped cell itself. */
void orc_electron_temp(const orc_problem *p, double *eenergysoln, const double *elecden, double eenergyscale,
                       double etempmin, double *electemp);
/* restrict_boundary for one face array of (a,b) -> (a/2,b/2). */
void orc_restrict_face(int a, int b, const double *f, double *c);
/* One perform_vcycle on the finest level; returns the RMS residual. */
double orc_vcycle(const orc_problem *p, double *u, const double *sterm);
/* Loop of V-cycles as generic_pde/pdedefs.f90:378-402: stops after a cycle
 * whose RMS residual <= tol; history gets one entry per cycle run.
 * Returns the number of cycles run. */
int orc_solve(const orc_problem *p, double *u, int max_cycles, double tol, double *history);
/* performgmres with find_AX and gseidel_precond on the whole grid
 * (single rank). x is ghosted, in/out. Returns 0 ok, -1 NaN. */
int orc_gmres(const orc_problem *p, const double *b, double *x, int maxkspdim, int nrestarts,
              double tol);
/* RMS norm of an interior array as compute_norm. */
double orc_rms(int n, const double *r);
/* find_sg_exponential_flux */
void orc_sg_flux(double cL, double cR, double DL, double DR, double uL, double uR, double dx,
                 double *wL, double *wR, double *flux);

#ifdef __cplusplus
}
#endif
#endif
