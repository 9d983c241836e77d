/*
 * orc.h -- the oracle: a plain-C, single-threaded CPU restatement of the APBS
 * PMG Poisson-Boltzmann path (src/pmgc/ *.c, src/mg/vpmg.c, src/mg/vpmgp.c,
 * src/generic/vpbe.c, src/generic/vacc.c).
 *
 * THIS IS TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker.  The
 * product (apbs_amd/, libapbs_pmg_hip.so) never links, loads or calls it.
 *
 * Every routine cites the reference file:line it restates.  Arrays are 0-based
 * with x fastest: IDX(i,j,k) = (k*ny + j)*nx + i, which is the same memory
 * layout as the reference's 1-based VAT3 (src/generic/vmatrix.h:76-87) and
 * IJK (src/mg/vpmg.h:1383).
 *
 * Parity pinning (tests/test_oracle_golden.py, tests/test_oracle_invariants.py):
 *  - golden energies of tests/test_cases.cfg (born mg-auto smol/mol at every
 *    focusing level, solv smol/mol, ion-protein npbe) to 1e-9 relative;
 *  - closed forms: Born's solvation energy, the Coulomb potential of a charge
 *    in a uniform dielectric, the Debye-Hueckel sphere with salt;
 *  - invariants: Galerkin = P^T A P (dense brute force, 7- and 27-point),
 *    restriction = P^T, interpolation = P, residual = f - (A + cc) x, one
 *    red-black sweep = the colour-by-colour update;
 *  - forces (qf + ib + db) = minus the central difference of the energy.
 *  - external maps: trilinear interpolation reproduces linear fields exactly;
 *    maps written from a solve's own coefficients, charges and potential
 *    reproduce that solve (tests/test_oracle_maps.py).
 *  - "parity unpinned": the exact iterate sequence of the lexicographic
 *    27-point Gauss-Seidel sweep and the banded LINPACK coarse solve beyond
 *    what the golden energies constrain; examples/born/polarforces is not a
 *    usable pin (the reference's own harness does not check it).
 */
#ifndef ORC_H
#define ORC_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXION 10

/* Enum values follow src/generic/vhal.h:103-123, 205-237. */
enum { ORC_SRFM_MOL = 0, ORC_SRFM_SMOL = 1, ORC_SRFM_SPL2 = 2, ORC_SRFM_SPL3 = 3, ORC_SRFM_SPL4 = 4 };
enum { ORC_CHGM_SPL0 = 0, ORC_CHGM_SPL2 = 1, ORC_CHGM_SPL4 = 2 };
enum { ORC_BCFL_ZERO = 0, ORC_BCFL_SDH = 1, ORC_BCFL_MDH = 2, ORC_BCFL_FOCUS = 4, ORC_BCFL_MAP = 6 };
/* External maps (usemap): Vpmg_fillco's dielXMap..chargeMap arguments. */
enum { ORC_MAP_DIELX = 0, ORC_MAP_DIELY = 1, ORC_MAP_DIELZ = 2, ORC_MAP_KAPPA = 3,
       ORC_MAP_POT = 4, ORC_MAP_CHARGE = 5 };
enum { ORC_LPBE = 0, ORC_NPBE = 1 };

typedef struct orc_params {
    int dime[3];
    double grid[3];        /* spacing h */
    double center[3];
    int nlev;              /* levels actually used (after MGparm_check) */
    int pbetype;           /* ORC_LPBE / ORC_NPBE */
    int nion;
    double ionq[ORC_MAXION], ionc[ORC_MAXION], ionr[ORC_MAXION];
    double pdie, sdie, srad, swin, sdens, temp;
    int srfm, chgm, bcfl;
    double etol;           /* <= 0: reference default 1e-6 (vpmgp.c:120-127) */
    /* Variant knobs, all 0 = reference behaviour.  1 = mirror the GPU schedule:
     * coarse_smoother 1: 8-colour Gauss-Seidel on 27-point levels;
     * deep_coarse 1: keep coarsening below nlev and solve the coarsest level
     * exactly with a dense Cholesky factorisation. */
    int coarse_smoother;
    int deep_coarse;
    /* > 0: stop Newton after this many steps (a timing sample, not a solve) */
    int newton_itmax;
} orc_params;

typedef struct orc_pmg orc_pmg;

/* Vpmg_ctor2 (vpmg.c:153-399) incl. Vpbe_ctor2 (vpbe.c:264-465).  When
 * old != NULL the new object focuses on it (focusFillBound + extEnergy) and
 * frees it, like vpmg.c:396. */
orc_pmg *orc_create(const orc_params *p, int natoms, const double *xyz,
                    const double *rad, const double *chg, orc_pmg *old,
                    int energy_flag);
void orc_destroy(orc_pmg *m);
int orc_fillco(orc_pmg *m);                         /* vpmg.c:5250 */
int orc_solve(orc_pmg *m);                          /* vpmg.c:401 */
double orc_energy(orc_pmg *m, int extFlag);         /* vpmg.c:1248 */
double orc_qf_energy(orc_pmg *m, int extFlag);      /* vpmg.c:1684 */
double orc_qm_energy(orc_pmg *m, int extFlag);      /* vpmg.c:1386 */
double orc_diel_energy(orc_pmg *m, int extFlag);    /* vpmg.c:1279 */
/* Vpmg_qfForce / ibForce / dbForce (vpmg.c:5440-6154); return 1/0. */
int orc_force(orc_pmg *m, int atom, double qf[3], double ib[3], double db[3]);
void orc_set_part_domain(orc_pmg *m);               /* setPartMG: routines.c:1676 */
/* Attaches a map (data copied, x fastest) for the next orc_fillco: n points
 * per axis, spacing h, first point origin.  Returns 1, or 0 on bad input. */
int orc_set_map(orc_pmg *m, int which, const int n[3], const double h[3],
                const double origin[3], const double *data);
/* Vgrid_value (vgrid.c:179-260) on such a map: 1 on the map, 0 off it. */
int orc_map_value(const int n[3], const double h[3], const double origin[3],
                  const double *data, const double pt[3], double *value);

/* Accessors: copies of the named n=nx*ny*nz array into out.
 * which: 0 u, 1 epsx, 2 epsy, 3 epsz, 4 kappa, 5 charge, 6 pvec. */
int orc_get_array(orc_pmg *m, int which, double *out);
/* Boundary arrays gxcf (2*ny*nz), gycf (2*nx*nz), gzcf (2*nx*ny). */
int orc_get_boundary(orc_pmg *m, double *gx, double *gy, double *gz);
/* Statistics: [0] solver iterations (V-cycles or Newton steps), [1] total
 * inner V-cycles (Newton), [2] nlev used, [3] final relative residual,
 * [4] kappa^2-bar (zkappa2), [5] zmagic, [6] xkappa, [7] ionic strength. */
int orc_stats(orc_pmg *m, double *out);

/* ---- low-level kernels (for kernel-by-kernel parity of the GPU path) ---- */
void orc_buildA_fv(int nx, int ny, int nz, const double *xf, const double *yf,
                   const double *zf, const double *gxcf, const double *gycf,
                   const double *gzcf, const double *a1cf, const double *a2cf,
                   const double *a3cf, const double *ccf, const double *fcf,
                   double *oC, double *cc, double *fc, double *oE, double *oN,
                   double *uC);
/* stencil 7: a holds 4 arrays; 27: 14 arrays (oC,oE,oN,uC,oNE,oNW,uE,uW,uN,
 * uS,uNE,uNW,uSE,uSW), each n long, packed back to back. */
void orc_gsrb(int nx, int ny, int nz, int stencil, const double *a,
              const double *cc, const double *fc, double *x, int itmax,
              int iadjoint, int colour8);
void orc_resid(int nx, int ny, int nz, int stencil, const double *a,
               const double *cc, const double *fc, const double *x, double *r);
void orc_matvec(int nx, int ny, int nz, int stencil, const double *a,
                const double *cc, const double *x, double *y);
void orc_restrict(int nxf, int nyf, int nzf, int nxc, int nyc, int nzc,
                  double *xin, double *xout);
void orc_interp(int nxc, int nyc, int nzc, int nxf, int nyf, int nzf,
                double *xin, double *xout);
void orc_galerkin(int nxf, int nyf, int nzf, int stencil_f, const double *af,
                  int nxc, int nyc, int nzc, double *ac);
/* Nonlinear term (mypdec.c:145-223) and its derivative (mypdec.c:366-433). */
void orc_cvec(int n, int nion, const double *q, const double *sconc,
              const double *coef, const double *u, double *out);
void orc_dcvec(int n, int nion, const double *q, const double *sconc,
               const double *coef, const double *u, double *out);
/* Reference sphere for SAS points (vacc.c:924-978). */
int orc_ref_sphere(int npts, double *xyz_out, int cap);

#ifdef __cplusplus
}
#endif
#endif
