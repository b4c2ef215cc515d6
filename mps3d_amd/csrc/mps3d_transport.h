// Transport assembly of the plasma equations on the device: the per-cell
// coefficients and sources each implicit solve of a time step takes
// (plasma_pde/{edeneq,ionseq,eenergyeq,poissoneq}.f90, *_update_transport,
// update_electrontemp), for chemistries given as tables: Arrhenius rates in
// T_e, and per species a constant mobility and diffusion
// (plasma_pde/chemistry/ion_elec_chem.f90) or the field-dependent mobility
// fits and Einstein diffusion of plasma_pde/chemistry/H2_O2_chem.f90,
// evaluated per cell (the solved neutral species of neutrals_update_transport
// take the species form with their zero mobility).
//
// Every kernel is a one-pass, HBM-bound stream over the ghosted grid (the
// Joule term reads the six neighbours of n_e and E as well, which the L2
// serves): one thread per cell, lanes along x, blockIdx.y walking the
// ghosted z planes. Arithmetic follows the reference's evaluation order
// term by term, and the library is built without FMA contraction, so every
// output that does not pass through exp/pow matches a sequential FP64
// evaluation bit for bit.
#pragma once

#include "mps3d_kernels.h"

namespace mps {

constexpr int kChemMaxSpecies = 16;
constexpr int kChemMaxReac = 64;

// the reference's physical constants (plasma_pde/fundconstants.f90:3-9)
constexpr double kEcharge = 1.602176e-19;
constexpr double kEps0 = 8.854188e-12;
constexpr double kBoltz = 1.380650e-23;

// One chemistry (chem_module's tables) in device memory. Species and
// reaction indices are 0-based here.
struct ChemDev {
    int ns, nr;
    int e, bg;          // electron and background-gas species
    int imin, imax;     // ion species range
    int nu[kChemMaxReac][kChemMaxSpecies];      // products - reactants
    int react[kChemMaxReac][kChemMaxSpecies];   // reactant moles
    double A[kChemMaxReac], alpha[kChemMaxReac], Ea[kChemMaxReac];
    double eel[kChemMaxReac];                   // electron energy change, eV
    double mob[kChemMaxSpecies], dco[kChemMaxSpecies];
    double mass[kChemMaxSpecies], charge[kChemMaxSpecies];
    // field-dependent transport: mobility and diffusion models per species
    int mkind[kChemMaxSpecies];                     // 0 constant, 1 electron fit, 2 reduced ion mobility
    int mterm[kChemMaxSpecies][kChemMaxSpecies];    // [s][j]: partner j enters s's sum
    double mco[kChemMaxSpecies][kChemMaxSpecies][3];
    int dkind[kChemMaxSpecies];                     // 0 constant, 1 Einstein with T_e, 2 with T_g
    // reaction rates: 0 getarrhrate; 1 the exponential fit in 1/T of
    // getcustomrate (O2_chem.f90:413-448); 2 alpha |E| exp(-E0 / |E|) of
    // getcustomrate (H2_chem.f90:423-437), |E| the cell's field magnitude
    int rkind[kChemMaxReac];
    double rco[kChemMaxReac][5];
};

constexpr double kEvToK = 11604.5;                  // plasma_pde/fundconstants.f90:9
constexpr double kNatm = 101325.0 / kBoltz / 300.0;  // Patm / k_BMAN / Troom of getspecmobility

// getspecmobility (H2_O2_chem.f90:600-672; ion_elec_chem.f90:222-242 for
// the constants) of species s at cell c
__device__ __forceinline__ double spec_mobility(const ChemDev &C, int s, const double *__restrict__ nd, long long ng,
                                               long long c, double Te)
{
    const int kind = C.mkind[s];
    if (kind == 0) return C.mob[s];
    double sum = 0.0;
    bool first = true;
    const double L = kind == 1 ? log(Te / kEvToK) : 0.0;
    for (int j = 0; j < C.ns; ++j) {
        if (!C.mterm[s][j]) continue;
        const double nj = nd[(long long)j * ng + c];
        const double inv = kind == 1 ? nj / exp(C.mco[s][j][0] * (L * L) + C.mco[s][j][1] * L + C.mco[s][j][2])
                                     : nj / C.mco[s][j][0];
        sum = first ? inv : sum + inv;
        first = false;
    }
    const double num = kind == 2 ? C.mob[s] * kNatm : C.mob[s];
    return num / sum;
}

// getspecdcoeff: constant, or the Einstein relation k_B T mu / (Z e)
__device__ __forceinline__ double spec_dcoeff(const ChemDev &C, int s, double mu, double Te, double Tg)
{
    const int kind = C.dkind[s];
    if (kind == 0) return C.dco[s];
    const double T = kind == 1 ? Te : Tg;
    return kBoltz * T * mu / (C.charge[s] * kEcharge);
}

// x**m for an integer m >= 0 by repeated squaring (the integer power of
// Fortran's real**integer: x**0 = 1, x**1 = x, x**2 = x*x, x**3 = x*(x*x))
__host__ __device__ __forceinline__ double powi_sq(double x, int m)
{
    double r = (m & 1) ? x : 1.0;
    while (m >>= 1) {
        x = x * x;
        if (m & 1) r = r * x;
    }
    return r;
}

// the rate of reaction r at electron temperature T and field magnitude
// efm, as getspecproduction takes it: getarrhrate (ion_elec_chem.f90:260-274)
// A T^alpha exp(-Ea/T), or getcustomrate: the fit of O2_chem.f90:413-448,
// C0 exp(c1/T + c2/T**2 + c3/T**3 + c4/T**4) (coefficients signed, the
// terms summed left to right), or alpha |E| exp(-E0/|E|) of
// H2_chem.f90:423-437. An Arrhenius rate with alpha = 0 or Ea = 0 skips
// the pow or the exp, which would return exactly 1 (T > 0).
__device__ __forceinline__ double reaction_rate(const ChemDev &C, int r, double T, double efm)
{
    const int kind = C.rkind[r];
    if (kind == 0) {
        const double a = C.alpha[r], ea = C.Ea[r];
        const double tp = a == 0.0 ? 1.0 : pow(T, a);
        const double ex = ea == 0.0 ? 1.0 : exp(-ea / T);
        return C.A[r] * tp * ex;
    }
    const double *q = C.rco[r];
    if (kind == 1) {
        // T**2, T**3, T**4 as the reference build forms them: a chain of
        // products ((T*T)*T)*T, not (T*T)*(T*T) (bit for bit on its rates)
        const double T2 = T * T, T3 = T2 * T, T4 = T3 * T;
        const double e = ((q[1] / T + q[2] / T2) + q[3] / T3) + q[4] / T4;
        return q[0] * exp(e);
    }
    return q[0] * fabs(efm) * exp(-q[1] / fabs(efm));
}

// product of the reactant densities of reaction r at cell c
__device__ __forceinline__ double reactant_product(const ChemDev &C, int r, const double *__restrict__ nd,
                                                   long long ng, long long c)
{
    double m = 1.0;
    for (int j = 0; j < C.ns; ++j) {
        const int k = C.react[r][j];
        if (k) m = m * powi_sq(nd[(long long)j * ng + c], k);
    }
    return m;
}

// findjheating_eachdir (eenergyeq.f90:461-555) along one axis with stride st:
// upwinded face fluxes mu n E - D grad n on both faces of the cell, and
// -e <E> <Gamma>, clipped at zero
__device__ __forceinline__ double joule_dir(const double *__restrict__ ne, const double *__restrict__ Eq,
                                            long long c, long long st, double mu, double D, double dx)
{
    const double nm = ne[c - st], n0 = ne[c], np = ne[c + st];
    const double Em = Eq[c - st], E0 = Eq[c], Ep = Eq[c + st];
    double grad = (n0 - nm) / dx;
    const double el = (E0 + Em) * 0.5;
    double den = (mu * el > 0.0) ? nm : n0;
    const double gl = mu * den * el - D * grad;
    grad = (np - n0) / dx;
    const double er = (E0 + Ep) * 0.5;
    den = (mu * er > 0.0) ? n0 : np;
    const double gr = mu * den * er - D * grad;
    double j = -(kEcharge * 0.5 * (el + er) * 0.5 * (gl + gr));
    if (j < 0.0) j = 0.0;
    return j;
}

// ghosted cell of this thread: (i, j) from the flattened x-y plane, k from
// blockIdx.y; interior = all three indices in 1..n
struct GCell {
    long long c;
    bool interior;
};

template <typename T>
__device__ __forceinline__ bool ghost_cell_of(const Lev<T> &L, GCell &g)
{
    const int px = L.nx + 2, py = L.ny + 2;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= px * py) return false;
    const int j = t / px, i = t - j * px, k = blockIdx.y;
    g.c = (long long)t + L.sz * k;
    g.interior = i >= 1 && i <= L.nx && j >= 1 && j <= L.ny && k >= 1 && k <= L.nz;
    return true;
}

struct TransportOut {
    double *vx, *vy, *vz, *dcoeff, *reac, *source, *prod;   // each nullable
};

struct TransportIn {
    const double *ex, *ey, *ez;   // cell-centre field
    const double *te;             // electron temperature, K
    const double *nd;             // ns ghosted density fields
    const double *soln;           // the equation's current solution (time term)
    double Tg, Pg;
    double dt;
    int timederiv;
    double vs, ds, rs, ss;        // vel, dcoeff, reac, source scalings
    double dxl[3];                // g_d * lscale per axis (Joule term)
};

// eden_update_transport (edeneq.f90:360-444) and ions_update_transport
// (ionseq.f90:478-558), KIND 0, for species sp; eenergy_update_transport
// (eenergyeq.f90:679-774) with jouleheatingterm, elastic_colterm and
// inelastic_colterm (:394-677), KIND 1.
template <int KIND>
__global__ void __launch_bounds__(256) k_transport(Lev<double> L, const ChemDev *__restrict__ Cp, int sp,
                                                   TransportIn in, TransportOut out)
{
    GCell g;
    if (!ghost_cell_of(L, g)) return;
    const ChemDev &C = *Cp;
    const long long c = g.c;
    const long long ng = L.sz * (L.nz + 2);
    double vx = 0.0, vy = 0.0, vz = 0.0, dc = 0.0, src = 0.0;
    if (g.interior) {
        const double Ex = in.ex[c], Ey = in.ey[c], Ez = in.ez[c];
        // efield_mag as the callers form it (edeneq.f90:405-409): only the
        // field-dependent custom rate reads it
        const double efm = sqrt((Ex * Ex + Ey * Ey) + Ez * Ez);
        const double Te = in.te[c];
        if (KIND == 0) {
            const double mu = spec_mobility(C, sp, in.nd, ng, c, Te);
            vx = mu * Ex;
            vy = mu * Ey;
            vz = mu * Ez;
            dc = spec_dcoeff(C, sp, mu, Te, in.Tg);
            // getspecproduction (ion_elec_chem.f90:111-156)
            double prod = 0.0;
            for (int r = 0; r < C.nr; ++r) {
                const int nm = C.nu[r][sp];
                if (nm == 0) continue;
                const double k = reaction_rate(C, r, Te, efm);
                prod = prod + (double)nm * k * reactant_product(C, r, in.nd, ng, c);
            }
            src = 0.0 + prod;
            if (out.prod) out.prod[c] = src;
        } else {
            const int e = C.e;
            const double mu = spec_mobility(C, e, in.nd, ng, c, Te), D = spec_dcoeff(C, e, mu, Te, in.Tg);
            const double five_by_three = 5.0 / 3.0;
            vx = five_by_three * mu * Ex;
            vy = five_by_three * mu * Ey;
            vz = five_by_three * mu * Ez;
            dc = five_by_three * D;
            const double *ne = in.nd + (long long)e * ng;
            const double jh = joule_dir(ne, in.ex, c, 1, mu, D, in.dxl[0]) +
                              joule_dir(ne, in.ey, c, L.sy, mu, D, in.dxl[1]) +
                              joule_dir(ne, in.ez, c, L.sz, mu, D, in.dxl[2]);
            // elastic_colterm: 3 k_B n_e (T_e - T_g) (m_e / m_bg) nu, with
            // getelectroncollisionfrequency nu = -e / m_e / mu_e
            const double me = C.mass[e], mbg = C.mass[C.bg];
            const double nu = -(kEcharge / me / mu);
            const double el = 3.0 * kBoltz * ne[c] * (Te - in.Tg) * (me / mbg) * nu;
            // inelastic_colterm with getelectroninelasticterm
            double inel = 0.0;
            for (int r = 0; r < C.nr; ++r) {
                if (C.eel[r] == 0.0) continue;
                const double k = reaction_rate(C, r, Te, efm);
                inel = inel + k * reactant_product(C, r, in.nd, ng, c) * C.eel[r];
            }
            src = jh - el + inel * kEcharge;
        }
    }
    if (out.vx) out.vx[c] = vx * in.vs;
    if (out.vy) out.vy[c] = vy * in.vs;
    if (out.vz) out.vz[c] = vz * in.vs;
    if (out.dcoeff) out.dcoeff[c] = dc * in.ds;
    if (out.reac) out.reac[c] = 0.0 * in.rs;
    if (out.source) {
        double s = src * in.ss;
        if (in.timederiv) s = s + in.soln[c] / in.dt;
        out.source[c] = s;
    }
}

// poisson_update_transport (poissoneq.f90:335-385): source = e/eps0
// (sum_i Z_i n_i - n_e) scaled, dcoeff = 1 scaled, vel = reac = 0
__global__ void __launch_bounds__(256) k_poisson_transport(Lev<double> L, const ChemDev *__restrict__ Cp,
                                                           const double *__restrict__ ne,
                                                           const double *__restrict__ ions,
                                                           TransportIn in, TransportOut out)
{
    GCell g;
    if (!ghost_cell_of(L, g)) return;
    const ChemDev &C = *Cp;
    const long long c = g.c;
    const long long ng = L.sz * (L.nz + 2);
    double src = 0.0;
    if (g.interior) {
        double tot = 0.0;
        for (int s = C.imin; s <= C.imax; ++s) tot = tot + C.charge[s] * ions[(long long)(s - C.imin) * ng + c];
        src = tot - ne[c];
    }
    src = src * kEcharge / kEps0;
    if (out.vx) out.vx[c] = 0.0;
    if (out.vy) out.vy[c] = 0.0;
    if (out.vz) out.vz[c] = 0.0;
    if (out.dcoeff) out.dcoeff[c] = 1.0 * in.ds;
    if (out.reac) out.reac[c] = 0.0;
    if (out.source) {
        double s = src * in.ss;
        if (in.timederiv) s = s + in.soln[c] / in.dt;
        out.source[c] = s;
    }
}

// The FLUX branch of ions_update_bcs (ionseq.f90:247-476): the face value
// of ion sp on the boundary cells of a FLUX face is the drift outflow
// mu n E_normal vscale where mu E_normal leaves the domain, else 0
// (n the scaled ion solution, mu from the densities); one face per
// blockIdx.y, the face arrays as the boundary conditions take them
// ((ny,nz), (nx,nz), (nx,ny), first index fastest). A null face is skipped.
struct IonFaces {
    double *f[6];
};

__global__ void __launch_bounds__(256) k_ion_drift_faces(Lev<double> L, const ChemDev *__restrict__ Cp, int sp,
                                                         const double *__restrict__ E,
                                                         const double *__restrict__ nd,
                                                         const double *__restrict__ ni, double vscale, IonFaces F)
{
    const int face = blockIdx.y;
    double *out = F.f[face];
    if (!out) return;
    const int axis = face / 2;
    const bool low = (face & 1) == 0;
    const int n0 = axis == 0 ? L.ny : L.nx;
    const int n1 = axis == 2 ? L.ny : L.nz;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n0 * n1) return;
    const int a = t % n0 + 1, b = t / n0 + 1;
    int i, j, k;
    if (axis == 0) {
        i = low ? 1 : L.nx;
        j = a;
        k = b;
    } else if (axis == 1) {
        i = a;
        j = low ? 1 : L.ny;
        k = b;
    } else {
        i = a;
        j = b;
        k = low ? 1 : L.nz;
    }
    const long long c = (long long)i + (long long)L.sy * j + L.sz * k;
    const long long ng = L.sz * (L.nz + 2);
    const ChemDev &C = *Cp;
    const double mu = spec_mobility(C, sp, nd, ng, c, 0.0);   // kinds 0 and 2: no T_e
    const double En = E[(long long)axis * ng + c];
    const double drift = mu * ni[c] * En * vscale;
    const double muE = mu * En;
    const bool leaves = low ? (muE < 0.0) : (muE > 0.0);
    out[t] = leaves ? (low ? -drift : drift) : 0.0;
}

// The face values of the thermal boundary updaters, one thread per boundary
// cell of each face with an output array, in the layout mps_set_bc takes:
//   kind 0  eden_update_bcs (edeneq.f90:203-358): FLUX faces
//           0.25 n cbar vscale, cbar = sqrt(c1 T) with T the boundary cell's
//           electron temperature
//   kind 1  neutrals_update_bcs (neutralseq.f90:206-358): FLUX faces
//           0.25 n cbar vscale with cbar = sqrt(c1 Tg)
//   kind 2  eenergy_update_bcs (eenergyeq.f90:208-392): DIRC faces
//           1.5 k_B n_e T_wall / escale; FLUX faces 4/3 n_eps cbar vscale,
//           cbar = sqrt(c1 T), T = n_eps escale / (1.5 n_e k_B)
// in the reference's operation order; faces of other codes keep their
// values. n is the scaled density (kinds 0, 1) or energy density (kind 2),
// T the electron temperature (kind 0) or the physical electron density
// (kind 2), all ghosted.
struct ThermalFaces {
    double *f[6];
    int code[6];
    double twall[6];
};

__device__ __forceinline__ long long face_cell(const Lev<double> &L, int face, int t, int n0)
{
    const int axis = face / 2;
    const bool low = (face & 1) == 0;
    const int a = t % n0 + 1, b = t / n0 + 1;
    int i, j, k;
    if (axis == 0) {
        i = low ? 1 : L.nx, j = a, k = b;
    } else if (axis == 1) {
        i = a, j = low ? 1 : L.ny, k = b;
    } else {
        i = a, j = b, k = low ? 1 : L.nz;
    }
    return (long long)i + (long long)L.sy * j + L.sz * k;
}

__global__ void __launch_bounds__(256) k_thermal_faces(Lev<double> L, int kind, const double *__restrict__ n,
                                                       const double *__restrict__ T, double c1, double Tg,
                                                       double escale, double vscale, ThermalFaces F)
{
    const int face = blockIdx.y;
    double *out = F.f[face];
    if (!out) return;
    const int code = F.code[face];
    if (code != BC_FLUX && !(kind == 2 && code == BC_DIRC)) return;
    const int axis = face / 2;
    const int n0 = axis == 0 ? L.ny : L.nx;
    const int n1 = axis == 2 ? L.ny : L.nz;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n0 * n1) return;
    const long long c = face_cell(L, face, t, n0);
    const double q = 0.25;   // a default-real literal, exact
    if (kind == 0) {
        const double cbar = sqrt(c1 * T[c]);
        out[t] = q * n[c] * cbar * vscale;
    } else if (kind == 1) {
        const double cbar = sqrt(c1 * Tg);
        out[t] = q * n[c] * cbar * vscale;
    } else {
        const double three_by_two = 1.5, four_by_three = 4.0 / 3.0;
        if (code == BC_DIRC) {
            out[t] = three_by_two * kBoltz * T[c] * F.twall[face] / escale;
        } else {
            const double etemp = n[c] * escale / (three_by_two * T[c] * kBoltz);
            const double cbar = sqrt(c1 * etemp);
            const double flux = four_by_three * n[c] * cbar;
            out[t] = flux * vscale;
        }
    }
}

// the density floor of eden_solve (edeneq.f90:509-518): interior values
// below vmin become vmin
__global__ void __launch_bounds__(256) k_floor(Lev<double> L, double *__restrict__ u, double vmin)
{
    int i, j, k;
    long long c;
    if (!cell_of(L, i, j, k, c)) return;
    if (u[c] < vmin) u[c] = vmin;
}

// update_electrontemp (eenergyeq.f90:842-877): T_e = n_eps scale /
// (3/2 n_e k_B) on the interior, clamped below at etempmin, where the
// energy density is reset to match the clamped temperature (in the cell
// itself: the reference writes it one cell down-left of it, an index slip)
__global__ void __launch_bounds__(256) k_electron_temp(Lev<double> L, double *__restrict__ eps,
                                                       const double *__restrict__ ne, double escale, double tmin,
                                                       double *__restrict__ te)
{
    int i, j, k;
    long long c;
    if (!cell_of(L, i, j, k, c)) return;
    const double three_by_two = 1.5;
    double T = eps[c] * escale / (three_by_two * ne[c] * kBoltz);
    if (T < tmin) {
        T = tmin;
        eps[c] = three_by_two * ne[c] * kBoltz * tmin / escale;
    }
    te[c] = T;
}

}  // namespace mps
