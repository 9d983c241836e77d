/* A C caller of libapbs_pmg.so through include/apbs_pmg.h only, following
 * the call sequence of APBS's src/routines.c for one mg-manual ELEC block:
 * initMG (:1314-1600), solveMG (:1635), setPartMG (:1676), energyMG (:1726)
 * and forceMG (:1803).
 *
 *   c_caller layout            print struct sizes and field offsets
 *   c_caller solve PQR DIME H  Born-type calculation: lpbe, srfm spl2,
 *                              chgm spl2, bcfl mdh, 0.15 M 1:1 salt; prints
 *                              the energy (kJ/mol) and the force on atom 0
 */
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "apbs_pmg.h"

#define OFF(T, f) printf(#T "." #f " %zu\n", offsetof(T, f))

static int layout(void) {
    printf("sizeof.Vatom %zu\nsizeof.Valist %zu\nsizeof.MGparm %zu\nsizeof.Vpmgp %zu\n"
           "sizeof.Vpbe %zu\n",
           sizeof(Vatom), sizeof(Valist), sizeof(MGparm), sizeof(Vpmgp), sizeof(Vpbe));
    OFF(Vatom, epsilon); OFF(Vatom, id); OFF(Vatom, atomName);
    OFF(Valist, atoms); OFF(Valist, vmem);
    OFF(MGparm, chgm); OFF(MGparm, nlev); OFF(MGparm, center); OFF(MGparm, partDisjOwnSide);
    OFF(MGparm, ofrac); OFF(MGparm, nonlintype); OFF(MGparm, useAqua);
    OFF(Vpmgp, nrwk); OFF(Vpmgp, errtol); OFF(Vpmgp, bcfl); OFF(Vpmgp, omegal); OFF(Vpmgp, zmax);
    OFF(Vpbe, T); OFF(Vpbe, numIon); OFF(Vpbe, ionConc); OFF(Vpbe, zmagic); OFF(Vpbe, smvolume);
    OFF(Vpbe, param2Flag);
    OFF(Vpmg, pbe); OFF(Vpmg, epsx); OFF(Vpmg, u); OFF(Vpmg, pvec); OFF(Vpmg, filled);
    OFF(Vpmg, chargeMap); OFF(Vpmg, dev);
    return 0;
}

static int solve(const char *pqr, int dime, double h) {
    Valist *alist = Valist_ctor();
    if (!Valist_readPQRFile(alist, pqr)) return 1;
    double conc[2] = {0.15, 0.15}, radii[2] = {2.0, 2.0}, q[2] = {1.0, -1.0};
    /* initMG: Vpbe_ctor (routines.c:1402) with sparm = swin for spl2 */
    Vpbe *pbe = Vpbe_ctor(alist, 2, conc, radii, q, 298.15, 1.0, 78.54, 0.3, 0, 10.0, 0, 0, 0,
                          0);
    MGparm *mg = MGparm_ctor(MCT_MANUAL);
    for (int j = 0; j < 3; j++) {
        mg->dime[j] = dime;
        mg->grid[j] = h;
        mg->center[j] = alist->center[j];
    }
    mg->setdime = mg->setgrid = mg->setgcent = 1;
    mg->chgm = VCM_BSPL2;
    mg->setchgm = 1;
    mg->nonlintype = NONLIN_LPBE;
    mg->method = VSOL_MG;
    if (!MGparm_check(mg)) return 2;
    Vpmgp *pmgp = Vpmgp_ctor(mg);
    pmgp->bcfl = BCFL_MDH;
    pmgp->xcent = mg->center[0];
    pmgp->ycent = mg->center[1];
    pmgp->zcent = mg->center[2];
    Vpmg *pmg = Vpmg_ctor(pmgp, pbe, 0, NULL, mg, PCE_TOTAL);
    if (!pmg) return 3;
    if (!Vpmg_fillco(pmg, VSM_SPLINE, 0.3, VCM_BSPL2, 0, NULL, 0, NULL, 0, NULL, 0, NULL, 0, NULL,
                     0, NULL))
        return 4;
    /* solveMG */
    if (!Vpmg_solve(pmg)) return 5;
    /* setPartMG: the whole grid */
    double lo[3], hi[3];
    for (int j = 0; j < 3; j++) {
        lo[j] = mg->center[j] - 0.5 * mg->glen[j];
        hi[j] = mg->center[j] + 0.5 * mg->glen[j];
    }
    Vpmg_setPart(pmg, lo, hi, mg->partDisjOwnSide);
    /* energyMG (routines.c:1756-1759): kT -> kJ/mol, in that order */
    double e = 1.3806581e-23 * 298.15 * 1e-3 * 6.0221367e+23 * Vpmg_energy(pmg, 1);
    /* forceMG: total = qf + ib + db (Vpmg_force) */
    double f[3];
    if (!Vpmg_force(pmg, f, 0, VSM_SPLINE, VCM_BSPL2)) return 6;
    /* the potential is mirrored to the host array callers read */
    double umax = 0;
    for (int i = 0; i < pmgp->nx * pmgp->ny * pmgp->nz; i++)
        if (pmg->u[i] > umax) umax = pmg->u[i];
    printf("energy %.17g\nforce %.17g %.17g %.17g\numax %.17g\n", e, f[0], f[1], f[2], umax);
    Vpmg_dtor(&pmg);
    Vpmgp_dtor(&pmgp);
    MGparm_dtor(&mg);
    Vpbe_dtor(&pbe);
    Valist_dtor(&alist);
    return 0;
}

int main(int argc, char **argv) {
    if (argc >= 2 && strcmp(argv[1], "layout") == 0) return layout();
    if (argc >= 5 && strcmp(argv[1], "solve") == 0)
        return solve(argv[2], atoi(argv[3]), atof(argv[4]));
    fprintf(stderr, "usage: c_caller layout | solve PQR DIME H\n");
    return 64;
}
