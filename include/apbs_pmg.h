/*
 * apbs_pmg.h -- C ABI of the MI355X PMG Poisson-Boltzmann solver
 * (libapbs_pmg.so).
 *
 * The entry points keep the names, argument order, argument meaning and
 * return conventions of APBS's Vpmg / Vpmgp API (src/mg/vpmg.h,
 * src/mg/vpmgp.h of Electrostatics/apbs 3.4.1), so the pmgc path of an APBS
 * caller (src/routines.c initMG/solveMG/energyMG/forceMG) binds to this
 * library instead.  Each declaration cites the reference interface it
 * replaces.  Plain pointers and sizes only; all device memory is owned by the
 * library.  Return codes: 1 = success, 0 = failure, as in the reference.
 *
 * The objects the Vpmg constructor borrows (Valist, Vpbe, MGparm) have the
 * reference's memory layout, so objects built by APBS's own src/generic code
 * can be passed in.  The library also supplies constructors for them with
 * the reference signatures, for callers without the rest of APBS.
 */
#ifndef APBS_PMG_H
#define APBS_PMG_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ enums
 * Values follow src/generic/vhal.h:83-295, src/generic/mgparm.h:77-83 and
 * src/generic/pbeparm.h:82-100. */
typedef enum { VSOL_CGMG = 0, VSOL_Newton = 1, VSOL_MG = 2, VSOL_CG = 3, VSOL_SOR = 4,
               VSOL_RBGS = 5, VSOL_WJ = 6, VSOL_Richardson = 7, VSOL_CGMGAqua = 8,
               VSOL_NewtonAqua = 9 } Vsol_Meth;
typedef enum { VSM_MOL = 0, VSM_MOLSMOOTH = 1, VSM_SPLINE = 2, VSM_SPLINE3 = 3,
               VSM_SPLINE4 = 4 } Vsurf_Meth;
typedef enum { PBE_LPBE = 0, PBE_NPBE = 1, PBE_LRPBE = 2, PBE_NRPBE = 3,
               PBE_SMPBE = 4 } Vhal_PBEType;
typedef enum { NONLIN_LPBE = 0, NONLIN_NPBE = 1, NONLIN_SMPBE = 2, NONLIN_LPBEAQUA = 3,
               NONLIN_NPBEAQUA = 4 } Vhal_NonlinType;
typedef enum { IPKEY_SMPBE = -2, IPKEY_LPBE = -1, IPKEY_NPBE = 0 } Vhal_IPKEYType;
typedef enum { BCFL_ZERO = 0, BCFL_SDH = 1, BCFL_MDH = 2, BCFL_UNUSED = 3, BCFL_FOCUS = 4,
               BCFL_MEM = 5, BCFL_MAP = 6 } Vbcfl;
typedef enum { VCM_TRIL = 0, VCM_BSPL2 = 1, VCM_BSPL4 = 2 } Vchrg_Meth;
typedef enum { VCM_CHARGE = 0, VCM_PERMANENT = 1, VCM_INDUCED = 2,
               VCM_NLINDUCED = 3 } Vchrg_Src;
typedef enum { VDT_CHARGE = 0, VDT_POT = 1, VDT_ATOMPOT = 2, VDT_SMOL = 3, VDT_SSPL = 4,
               VDT_VDW = 5, VDT_IVDW = 6, VDT_LAP = 7, VDT_EDENS = 8, VDT_NDENS = 9,
               VDT_QDENS = 10, VDT_DIELX = 11, VDT_DIELY = 12, VDT_DIELZ = 13,
               VDT_KAPPA = 14 } Vdata_Type;
typedef enum { MCT_MANUAL = 0, MCT_AUTO = 1, MCT_PARALLEL = 2, MCT_DUMMY = 3,
               MCT_NONE = 4 } MGparm_CalcType;
typedef enum { MCM_POINT = 0, MCM_MOLECULE = 1, MCM_FOCUS = 2 } MGparm_CentMeth;
typedef enum { PCE_NO = 0, PCE_TOTAL = 1, PCE_COMPS = 2 } PBEparm_calcEnergy;

#define APBS_PMG_MAXION 10   /* src/generic/vhal.h:377 MAXION */
/* Calculations one batched launch holds (Vpmg_solveBatch, PmgServer). */
#define APBS_PMG_MAXBATCH 8

typedef struct sVmem Vmem;   /* opaque; unused by this library */
/* struct sVgrid, field for field (src/mg/vgrid.h:85-104): a scalar map on a
 * regular grid, data x fastest.  Used by Vpmg_fillco for the coefficient,
 * charge and potential maps (usemap, bcfl map). */
typedef struct sVgrid {
    int nx, ny, nz;
    double hx, hy, hzed;
    double xmin, ymin, zmin;
    double xmax, ymax, zmax;
    double *data;
    int readdata;
    int ctordata;
    void *mem;
} Vgrid;

/* ------------------------------------------------------------------ atoms
 * struct sVatom (src/generic/vatom.h:84-100) and struct sValist
 * (src/generic/valist.h:78-89), byte for byte: a Valist built by APBS's own
 * src/generic code can be handed to this library. */
#define APBS_PMG_RECLEN 64   /* src/generic/vatom.h:77 VMAX_RECLEN */
typedef struct sVatom {
    double position[3];
    double radius;
    double charge;
    double partID;
    double epsilon;
    int id;
    char resName[APBS_PMG_RECLEN];
    char atomName[APBS_PMG_RECLEN];
} Vatom;

typedef struct sValist {
    int number;
    double center[3];
    double mincrd[3];
    double maxcrd[3];
    double maxrad;
    double charge;
    Vatom *atoms;
    Vmem *vmem;
} Valist;

/* Valist_ctor / Valist_dtor (src/generic/valist.c:122-186) */
Valist *Valist_ctor(void);
void Valist_dtor(Valist **thee);
/* Replaces Valist_readPQR(thee, param, sock) (src/generic/valist.c:606-723),
 * which reads through a MALOC socket; reads the whitespace PQR records of a
 * file path.  Returns 1/0. */
int Valist_readPQRFile(Valist *thee, const char *path);
/* No reference counterpart: fills the list from plain arrays (xyz: 3n). */
int Valist_setAtoms(Valist *thee, int n, const double *xyz, const double *radius,
                    const double *charge);
/* Valist_getStatistics (src/generic/valist.c:869-921) */
int Valist_getStatistics(Valist *thee);
int Valist_getNumberAtoms(Valist *thee);        /* valist.h:141 */
Vatom *Valist_getAtom(Valist *thee, int i);      /* valist.h:150 */

/* ------------------------------------------------------------------- Vpbe
 * struct sVpbe (src/generic/vpbe.h:84-137) byte for byte, and the leading
 * fields of struct sVclist (src/generic/vclist.h:118-130), struct sVacc
 * (src/generic/vacc.h:109-125) and struct sVaccSurf (vacc.h:84-95) up to the
 * last field this library reads (Vclist max_radius, the Vacc reference
 * sphere).  A Vpbe built by APBS's own Vpbe_ctor can therefore be passed to
 * Vpmg_ctor; the library never reads past these fields. */
typedef struct sVaccSurf {
    Vmem *mem;
    double *xpts, *ypts, *zpts;
    char *bpts;
    double area;
    int npts;
    double probe_radius;
} VaccSurf;

typedef struct sVclist {
    Vmem *vmem;
    Valist *alist;
    int mode;               /* Vclist_DomainMode */
    int npts[3];
    int n;
    double max_radius;
    /* ... further fields of the reference are not read */
} Vclist;

typedef struct sVacc {
    Vmem *mem;
    Valist *alist;
    Vclist *clist;
    int *atomFlags;
    VaccSurf *refSphere;
    /* ... further fields of the reference are not read */
} Vacc;

typedef struct sVpbe {
    Vmem *vmem;
    Valist *alist;
    Vclist *clist;
    Vacc *acc;
    double T;
    double soluteDiel;
    double solventDiel;
    double solventRadius;
    double bulkIonicStrength;
    double maxIonRadius;
    int numIon;
    double ionConc[APBS_PMG_MAXION];
    double ionRadii[APBS_PMG_MAXION];
    double ionQ[APBS_PMG_MAXION];
    double xkappa;
    double deblen;
    double zkappa2;
    double zmagic;
    double soluteCenter[3];
    double soluteRadius;
    double soluteXlen;
    double soluteYlen;
    double soluteZlen;
    double soluteCharge;
    double smvolume;
    double smsize;
    int ipkey;
    int paramFlag;
    double z_mem;
    double L;
    double membraneDiel;
    double V;
    int param2Flag;
} Vpbe;

/* Vpbe_ctor (src/generic/vpbe.c:246-262, body :264-465): also builds the
 * cell-list radius and the reference sphere of vpbe.c:415-439 and
 * vacc.c:146-189. */
Vpbe *Vpbe_ctor(Valist *alist, int ionNum, double *ionConc, double *ionRadii, double *ionQ,
                double T, double soluteDiel, double solventDiel, double solventRadius,
                int focusFlag, double sdens, double z_mem, double L, double membraneDiel,
                double V);
void Vpbe_dtor(Vpbe **thee);
double Vpbe_getBulkIonicStrength(Vpbe *thee);   /* vpbe.h:174 */
double Vpbe_getMaxIonRadius(Vpbe *thee);        /* vpbe.h:182 */
double Vpbe_getTemperature(Vpbe *thee);         /* vpbe.h:190 */
double Vpbe_getSoluteDiel(Vpbe *thee);          /* vpbe.h:198 */
double Vpbe_getSolventDiel(Vpbe *thee);         /* vpbe.h:262 */
double Vpbe_getSolventRadius(Vpbe *thee);       /* vpbe.h:270 */
double Vpbe_getXkappa(Vpbe *thee);              /* vpbe.h:278 */
double Vpbe_getZkappa2(Vpbe *thee);             /* vpbe.h:294 */
double Vpbe_getZmagic(Vpbe *thee);              /* vpbe.h:302 */
double Vpbe_getSoluteRadius(Vpbe *thee);        /* vpbe.h:214 */
double Vpbe_getSoluteCharge(Vpbe *thee);        /* vpbe.h:254 */
double *Vpbe_getSoluteCenter(Vpbe *thee);       /* vpbe.h:246 */
Valist *Vpbe_getValist(Vpbe *thee);             /* vpbe.h:158 */
int Vpbe_getIons(Vpbe *thee, int *nion, double ionConc[APBS_PMG_MAXION],
                 double ionRadii[APBS_PMG_MAXION], double ionQ[APBS_PMG_MAXION]);

/* ----------------------------------------------------------------- MGparm
 * struct sMGparm (src/generic/mgparm.h:112-183) byte for byte. */
typedef struct sMGparm {
    MGparm_CalcType type;
    int parsed;
    int dime[3];      int setdime;
    Vchrg_Meth chgm;  int setchgm;
    Vchrg_Src chgs;
    int nlev;         int setnlev;
    double etol;      int setetol;
    double grid[3];   int setgrid;
    double glen[3];   int setglen;
    MGparm_CentMeth cmeth;   double center[3];  int centmol;  int setgcent;
    double cglen[3];  int setcglen;
    double fglen[3];  int setfglen;
    MGparm_CentMeth ccmeth;  double ccenter[3]; int ccentmol; int setcgcent;
    MGparm_CentMeth fcmeth;  double fcenter[3]; int fcentmol; int setfgcent;
    double partDisjCenter[3];
    double partDisjLength[3];
    int partDisjOwnSide[6];
    int pdime[3];     int setpdime;
    int proc_rank;    int setrank;
    int proc_size;    int setsize;
    double ofrac;     int setofrac;
    int async;        int setasync;
    int nonlintype;   int setnonlintype;
    int method;       int setmethod;
    int useAqua;      int setUseAqua;
} MGparm;

MGparm *MGparm_ctor(MGparm_CalcType type);      /* src/generic/mgparm.c:114 */
void MGparm_dtor(MGparm **thee);                /* src/generic/mgparm.c:172 */
/* MGparm_check (src/generic/mgparm.c:182-330): validates and derives nlev
 * and dime; returns 1 (success) / 0. */
int MGparm_check(MGparm *thee);
/* NOsh_setupCalcMGAUTO (src/generic/nosh.c:1828-2101): expands an mg-auto
 * ELEC block into its focusing chain; writes up to maxcalc MGparm
 * copies to out (the caller allocates them with MGparm_ctor) and returns
 * the number of levels, 0 on failure.  No reference counterpart as a
 * separate call: the reference appends to the NOsh calculation list. */
int MGparm_setupAuto(MGparm *elec, MGparm **out, int maxcalc);
/* NOsh_setupCalcMGPARA (src/generic/nosh.c:2103-2394) for processor `rank`
 * (the `async` rank) of an mg-para block (type MCT_PARALLEL, pdime and ofrac
 * set): the processor's overlapping fine grid and disjoint partition, then
 * that grid's focusing chain as MGparm_setupAuto writes it, every level
 * MCT_PARALLEL with the partition kept.  Returns the number of levels, 0 on
 * failure.  No reference counterpart as a separate call (see above). */
int MGparm_setupPara(MGparm *elec, int rank, MGparm **out, int maxcalc);

/* ------------------------------------------------------------------ Vgrid */
/* Vgrid_ctor (vgrid.c:86-110): a map over data (borrowed, not copied); xmax
 * = xmin + (nx-1) hx and likewise.  Returns NULL on failure. */
Vgrid *Vgrid_ctor(int nx, int ny, int nz, double hx, double hy, double hzed, double xmin,
                  double ymin, double zmin, double *data);
void Vgrid_dtor(Vgrid **thee);                                   /* vgrid.c:152 */
/* Vgrid_value (vgrid.c:179-260): trilinear value at pt; 1 on the map, 0 off
 * it.  Coordinates within 1e-4 of a face snap to it, as in the reference. */
int Vgrid_value(Vgrid *thee, double pt[3], double *value);

/* ------------------------------------------------------------------ Vpmgp
 * struct sVpmgp, field for field (src/mg/vpmgp.h:80-198). */
typedef struct sVpmgp {
    int nx, ny, nz, nlev;
    double hx, hy, hzed;
    int nonlin;
    int nxc, nyc, nzc, nf, nc, narrc, n_rpc, n_iz, n_ipc;
    size_t nrwk;
    int niwk, narr, ipkey;
    double xcent, ycent, zcent, errtol;
    int itmax, istop, iinfo;
    Vbcfl bcfl;
    int key, iperf, meth, mgkey, nu1, nu2, mgsmoo, mgprol, mgcoar, mgsolv, mgdisc;
    double omegal, omegan;
    int irite, ipcon;
    double xlen, ylen, zlen, xmin, ymin, zmin, xmax, ymax, zmax;
} Vpmgp;

Vpmgp *Vpmgp_ctor(MGparm *mgparm);              /* src/mg/vpmgp.c:76-170 */
void Vpmgp_dtor(Vpmgp **thee);                  /* src/mg/vpmgp.c:178 */
void Vpmgp_size(Vpmgp *thee);                   /* src/mg/vpmgp.c:196-291 */
void Vpmgp_makeCoarse(int numLevel, int nxOld, int nyOld, int nzOld, int *nxNew,
                      int *nyNew, int *nzNew);  /* src/mg/vpmgp.c:312-334 */

/* ------------------------------------------------------------------- Vpmg
 * struct sVpmg: the reference field order up to chargeMap
 * (src/mg/vpmg.h:116-193), which callers dereference directly (pmgp, pbe,
 * u, pvec, rwork, vmem).  u, pvec and rwork are host arrays of nx*ny*nz
 * doubles; u is mirrored from the device after Vpmg_solve and pvec after
 * Vpmg_setPart.  epsx..charge are mirrored on demand by Vpmg_fillArray and
 * are NULL until then.  The arrays the reference uses only as PMG
 * workspace (iparm..tcf, xf..gzcf) are NULL.  Device state hangs off the
 * trailing private field. */
typedef struct sVpmg {
    Vmem *vmem;
    Vpmgp *pmgp;
    Vpbe *pbe;
    double *epsx, *epsy, *epsz, *kappa, *pot, *charge;
    int *iparm;
    double *rparm;
    int *iwork;
    double *rwork;
    double *a1cf, *a2cf, *a3cf, *ccf, *fcf, *tcf;
    double *u;
    double *xf, *yf, *zf;
    double *gxcf, *gycf, *gzcf;
    double *pvec;
    double extDiEnergy, extQmEnergy, extQfEnergy, extNpEnergy;
    Vsurf_Meth surfMeth;
    double splineWin;
    Vchrg_Meth chargeMeth;
    Vchrg_Src chargeSrc;
    int filled;
    int useDielXMap;  Vgrid *dielXMap;
    int useDielYMap;  Vgrid *dielYMap;
    int useDielZMap;  Vgrid *dielZMap;
    int useKappaMap;  Vgrid *kappaMap;
    int usePotMap;    Vgrid *potMap;
    int useChargeMap; Vgrid *chargeMap;
    void *dev;   /* private: device state */
} Vpmg;

/* Vpmg_ctor / Vpmg_ctor2 (src/mg/vpmg.h:226-242, vpmg.c:141-399).  Borrows
 * pmgp and pbe; consumes and frees pmgOLD (focusing), like vpmg.c:396. */
Vpmg *Vpmg_ctor(Vpmgp *parms, Vpbe *pbe, int focusFlag, Vpmg *pmgOLD, MGparm *mgparm,
                PBEparm_calcEnergy energyFlag);
int Vpmg_ctor2(Vpmg *thee, Vpmgp *parms, Vpbe *pbe, int focusFlag, Vpmg *pmgOLD,
               MGparm *mgparm, PBEparm_calcEnergy energyFlag);
void Vpmg_dtor(Vpmg **thee);                    /* src/mg/vpmg.h:260 */
void Vpmg_dtor2(Vpmg *thee);                    /* src/mg/vpmg.h:269 */
/* Vpmg_fillco (src/mg/vpmg.h:281-298, vpmg.c:5250-5415).  External maps are
 * supported as in the reference: dielectric x/y/z (sampled at the shifted
 * face points), kappa (scaled by its maximum), charge (times zmagic) and the
 * potential map of bcfl map.  Returns 0 (with the reference's message) when a
 * grid point falls off a map in use, when a kappa map holds a negative value,
 * when coefficient maps are used without all three dielectric maps or (with
 * salt) without a kappa map, or when a map in use is NULL or has no data. */
int Vpmg_fillco(Vpmg *thee, Vsurf_Meth surfMeth, double splineWin, Vchrg_Meth chargeMeth,
                int useDielXMap, Vgrid *dielXMap, int useDielYMap, Vgrid *dielYMap,
                int useDielZMap, Vgrid *dielZMap, int useKappaMap, Vgrid *kappaMap,
                int usePotMap, Vgrid *potMap, int useChargeMap, Vgrid *chargeMap);
/* Vpmg_solve (src/mg/vpmg.h:306, vpmg.c:401-570): 0 if not filled. */
int Vpmg_solve(Vpmg *thee);
/* Energies in kT (src/mg/vpmg.h:334-445, vpmg.c:1248-1891) */
double Vpmg_energy(Vpmg *thee, int extFlag);
double Vpmg_qfEnergy(Vpmg *thee, int extFlag);
double Vpmg_qmEnergy(Vpmg *thee, int extFlag);
double Vpmg_dielEnergy(Vpmg *thee, int extFlag);
double Vpmg_qfAtomEnergy(Vpmg *thee, Vatom *atom);   /* vpmg.h:386 */
/* Forces in kT/A (src/mg/vpmg.h:482-546, vpmg.c:5417-6154) */
int Vpmg_force(Vpmg *thee, double *force, int atomID, Vsurf_Meth srfm, Vchrg_Meth chgm);
int Vpmg_qfForce(Vpmg *thee, double *force, int atomID, Vchrg_Meth chgm);
int Vpmg_ibForce(Vpmg *thee, double *force, int atomID, Vsurf_Meth srfm);
int Vpmg_dbForce(Vpmg *thee, double *dbForce, int atomID, Vsurf_Meth srfm);
/* Partitions (src/mg/vpmg.h:553-569, vpmg.c:627-890) */
void Vpmg_setPart(Vpmg *thee, double lowerCorner[3], double upperCorner[3], int bflags[6]);
void Vpmg_unsetPart(Vpmg *thee);
/* Vpmg_fillArray (src/mg/vpmg.h:576, vpmg.c:892): VDT_POT, VDT_CHARGE,
 * VDT_DIELX/Y/Z, VDT_KAPPA; pbeparm is unused and may be NULL. */
int Vpmg_fillArray(Vpmg *thee, double *vec, Vdata_Type type, double parm,
                   Vhal_PBEType pbetype, void *pbeparm);
unsigned long int Vpmg_memChk(Vpmg *thee);     /* src/mg/vpmg.h:208 */

/* ------------------------------------------------------ extensions
 * No reference counterpart. */

/* Precision of the multigrid solve: 0 = FP64 throughout (default, the
 * reference arithmetic); 1 = FP32 V-cycles (smoother, transfers, coarse
 * operators) inside an FP64 defect-correction loop whose fine residual,
 * coarsest solve, Newton update and energy/force sums stay FP64; 2 = mode 1,
 * and a later Vpmg_fillco evaluates each atom's term of the MDH boundary
 * (bcfl mdh) in FP32 and sums the terms in FP64 (boundary values within
 * about 1e-6 relative of the FP64 ones; the problem solved changes by that
 * much, so the result no longer equals the FP64 solve to the solver
 * tolerance).  tol is the relative FP64 residual the mixed loop stops at
 * (<= 0: pmgp->errtol).  Returns 0 for another mode. */
int Vpmg_setPrecision(Vpmg *thee, int mode, double tol);
/* Vpmg_solve for n calculations at once: consecutive runs of up to 8
 * calculations with the same grid, level count, method (lpbe or npbe) and
 * precision share every V-cycle launch (each stage is one kernel launch over
 * the run) and the coarsest-level inverses; each calculation keeps its own
 * Newton path, stopping test and results, bitwise those of Vpmg_solve.  The
 * run uses the first calculation's stream.  ok (optional, n entries)
 * receives each Vpmg_solve return value; returns the number solved. */
int Vpmg_solveBatch(Vpmg **pmgs, int n, int *ok);
/* Whether the host arrays callers dereference are kept current (default 1):
 * Vpmg_solve copies u back to the host, Vpmg_setPart and Vpmg_unsetPart
 * write pvec there.  With 0 both stay on the device only (Vpmg_copyArray
 * reads them on demand). */
int Vpmg_setHostMirror(Vpmg *thee, int on);
/* Solver statistics of the last solve: [0] V-cycles (linear) or Newton
 * steps, [1] total inner V-cycles, [2] levels used, [3] final relative
 * residual, [4] solve time (s), [5] fillco time (s), [6] operator-build time
 * (s), [7] mixed-precision outer iterations.  Returns the count written. */
int Vpmg_getStats(Vpmg *thee, double *out, int n);
/* Copy a fine-grid device array to host: 0 u, 1 epsx, 2 epsy, 3 epsz,
 * 4 kappa, 5 charge, 6 pvec, 7 gxcf (2*ny*nz), 8 gycf, 9 gzcf. */
int Vpmg_copyArray(Vpmg *thee, int which, double *out, size_t n);
/* Kernel timing with HIP events on the object's stream: on != 0 starts (and
 * zeroes) the accumulators, 0 stops.  Vpmg_getKernelTiming writes up to 8
 * values: [0] fine-level smoother time (ms), [1] its kernel launches,
 * [2] V-cycle time (ms), [3] V-cycles, [4] word size of the last cycle (4 or
 * 8 bytes), [5] fine-level points, [6] algorithmic words moved per fine
 * point by one smoother launch, [7] 1 when the fused sweep ran in its vector
 * form (16-byte loads), else 0. */
int Vpmg_setKernelTiming(Vpmg *thee, int on);
int Vpmg_getKernelTiming(Vpmg *thee, double *out, int n);
/* Vpmg_energy(thee, extFlag) computed without allocating any memory, for
 * callers that cannot allocate (tight serving loops, code that must not call
 * hipMalloc).  No reference counterpart.  Writes the energy (kT) to *energy and
 * returns 1; returns 0 with a message on a null argument, when the energy is
 * not finite (the exponential of the nonlinear term was chopped, as
 * Vpmg_energy reports with NAN), or for the point-charge energy (chgm spl0
 * without a charge map) of more than 131072 atoms, the per-atom partials the
 * library's static device buffer holds.  Its reductions use that buffer
 * (about 1 MB per device, held until the process ends) instead of the
 * object's buffers: calls that overlap on one device race. */
int Vpmg_energyNoAlloc(Vpmg *thee, int extFlag, double *energy);
/* hipDeviceSynchronize(); 1 on success. */
int apbs_pmg_synchronize(void);
/* Library version string. */
const char *apbs_pmg_version(void);

#ifdef __cplusplus
}
#endif
#endif
