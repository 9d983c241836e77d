/*
 * orc_pb.c -- oracle restatement of the Vpmg object around the multigrid core:
 * Vpbe constants (src/generic/vpbe.c), coefficient and charge maps
 * (src/mg/vpmg.c fillco*), SAS points (src/generic/vacc.c), boundary
 * conditions, focusing, energies and forces.  Test infrastructure only.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "orc_int.h"

struct orc_pmg {
    orc_params p;
    int natoms;
    double *apos, *arad, *achg, *partID;
    /* Vpbe (vpbe.c:264-465) */
    double center[3], soluteRadius, soluteCharge;
    double ionstr, maxIonRadius, xkappa, zkappa2, zmagic, solventRadius;
    double clist_maxrad;
    int nsphere;
    /* Vpmgp */
    int nx, ny, nz, nlev;
    double hx, hy, hz, xlen, ylen, zlen, xmin, ymin, zmin, xmax, ymax, zmax;
    size_t n;
    double *epsx, *epsy, *epsz, *kappa, *charge, *u, *pvec;
    double *gxcf, *gycf, *gzcf;
    double *xf, *yf, *zf;
    double extQf, extQm, extDi;
    int filled, solved;
    omg g;
    /* external maps (Vpmg_fillco's useXMap / XMap, vpmg.c:5299-5310) */
    int usemap[6];
    struct ogrid *map[6];
};

/* A scalar field on a regular grid: a Vgrid (vgrid.h:85-104), or a coarser
 * solution seen by focusing. */
typedef struct ogrid {
    int nx, ny, nz;
    double hx, hy, hz, xmin, ymin, zmin, xmax, ymax, zmax;
    double *data;
} ogrid;

void orc_fbound_apply(int nx, int ny, int nz, double *u, const double *gx,
                      const double *gy, const double *gz);

#define IJK(i, j, k) OIDX(i, j, k, nx, ny)

/* ------------------------------------------------------------------ Vpbe */
static void setup_pbe(orc_pmg *m)
{
    const double N_A = 6.022045000e+23, e_c = 4.803242384e-10, k_B = 1.380662000e-16;
    const double pi = 4. * atan(1.);
    int na = m->natoms;
    /* Valist_getStatistics (valist.c:869-912): bounding-box midpoint */
    double mn[3], mx[3];
    for (int d = 0; d < 3; d++) mn[d] = mx[d] = m->apos[d];
    for (int a = 0; a < na; a++)
        for (int d = 0; d < 3; d++) {
            double v = m->apos[3 * a + d];
            if (v < mn[d]) mn[d] = v;
            if (v > mx[d]) mx[d] = v;
        }
    for (int d = 0; d < 3; d++) m->center[d] = 0.5 * (mx[d] + mn[d]);
    /* solute radius and charge (vpbe.c:305-345) */
    double radius = 0, charge = 0;
    for (int a = 0; a < na; a++) {
        double dx = m->apos[3 * a] - m->center[0], dy = m->apos[3 * a + 1] - m->center[1],
               dz = m->apos[3 * a + 2] - m->center[2];
        double dist = sqrt(dx * dx + dy * dy + dz * dz) + m->arad[a];
        if (dist > radius) radius = dist;
        charge += m->achg[a];
    }
    m->soluteRadius = radius;
    m->soluteCharge = charge;
    /* ions (vpbe.c:352-374) */
    m->ionstr = 0.0; m->maxIonRadius = 0.0;
    for (int i = 0; i < m->p.nion; i++) {
        if (m->p.ionr[i] > m->maxIonRadius) m->maxIonRadius = m->p.ionr[i];
        m->ionstr += 0.5 * m->p.ionc[i] * OSQR(m->p.ionq[i]);
    }
    double T = m->p.temp;
    if (m->ionstr == 0.) { m->xkappa = 0.; m->zkappa2 = 0.; }
    else {
        m->xkappa = sqrt(m->ionstr * 1.0e-16 * ((8.0 * pi * N_A * e_c * e_c) /
                                                  (1000.0 * m->p.sdie * k_B * T)));
        m->zkappa2 = m->p.sdie * OSQR(m->xkappa);
    }
    m->zmagic = ((4.0 * pi * e_c * e_c) / (k_B * T)) * 1.0e+8;
    /* solvent radius handed to Vpbe (routines.c:1371-1378) */
    m->solventRadius = (m->p.srfm == ORC_SRFM_SPL2) ? m->p.swin : m->p.srad;
    /* Vclist max radius (vpbe.c:427-431) and refSphere size (vacc.c:170-184) */
    m->clist_maxrad = (m->maxIonRadius > m->solventRadius ? m->maxIonRadius : m->solventRadius) + 0.5;
    double maxrad = 0.0;
    for (int a = 0; a < na; a++) if (m->arad[a] > maxrad) maxrad = m->arad[a];
    maxrad += m->clist_maxrad;
    m->nsphere = (int)ceil(4.0 * O_VPI * maxrad * maxrad * m->p.sdens);
}

/* ----------------------------------------------------------- SAS surface */

/* VaccSurf_refSphere (vacc.c:924-978); VRINT = floor(x+0.5) (MALOC) */
int orc_ref_sphere(int npts, double *xyz, int cap)
{
    double frac = ((double)npts) / 4.0;
    int ntheta = (int)floor(sqrt(O_VPI * frac) + 0.5);
    double dtheta = O_VPI / ((double)ntheta);
    int nphimax = 2 * ntheta, nact = 0;
    for (int it = 0; it < ntheta; it++) {
        double theta = dtheta * ((double)it);
        double sintheta = sin(theta), costheta = cos(theta);
        int nphi = (int)floor(sintheta * nphimax + 0.5);
        if (nphi != 0) {
            double dphi = 2 * O_VPI / ((double)nphi);
            for (int ip = 0; ip < nphi; ip++) {
                double phi = dphi * ((double)ip);
                if (xyz && nact < cap) {
                    xyz[3 * nact] = cos(phi) * sintheta;
                    xyz[3 * nact + 1] = sin(phi) * sintheta;
                    xyz[3 * nact + 2] = costheta;
                }
                nact++;
            }
        }
    }
    return nact;
}

/* Simple uniform cell hash for the exclusion test; the reference's Vclist
 * (vclist.c) registers each atom in every cell its inflated sphere touches,
 * so its test is equivalent to checking every atom. */
typedef struct { double lo[3], cs; int dim[3]; int *start, *list; } ocell;

static void cell_build(ocell *c, const double *pos, const double *rad, int na, double reach)
{
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    double rmax = 0;
    for (int a = 0; a < na; a++) {
        for (int d = 0; d < 3; d++) {
            if (pos[3 * a + d] < mn[d]) mn[d] = pos[3 * a + d];
            if (pos[3 * a + d] > mx[d]) mx[d] = pos[3 * a + d];
        }
        if (rad[a] > rmax) rmax = rad[a];
    }
    c->cs = rmax + reach;
    if (c->cs < 1.0) c->cs = 1.0;
    for (int d = 0; d < 3; d++) {
        c->lo[d] = mn[d] - c->cs;
        c->dim[d] = (int)((mx[d] - c->lo[d]) / c->cs) + 2;
    }
    int nc = c->dim[0] * c->dim[1] * c->dim[2];
    c->start = (int *)calloc((size_t)nc + 1, sizeof(int));
    c->list = (int *)malloc((size_t)(na > 0 ? na : 1) * sizeof(int));
    int *cid = (int *)malloc((size_t)(na > 0 ? na : 1) * sizeof(int));
    for (int a = 0; a < na; a++) {
        int ic[3];
        for (int d = 0; d < 3; d++) ic[d] = (int)((pos[3 * a + d] - c->lo[d]) / c->cs);
        cid[a] = ic[0] + c->dim[0] * (ic[1] + c->dim[1] * ic[2]);
        c->start[cid[a] + 1]++;
    }
    for (int i = 0; i < nc; i++) c->start[i + 1] += c->start[i];
    int *fill = (int *)calloc((size_t)nc, sizeof(int));
    for (int a = 0; a < na; a++) c->list[c->start[cid[a]] + fill[cid[a]]++] = a;
    free(fill); free(cid);
}

static void cell_free(ocell *c) { free(c->start); free(c->list); }

/* ivdwAccExclus (vacc.c:78-128): 1 if no atom other than self lies within
 * (its radius + probe) of the point. */
static int acc_exclus(const ocell *c, const double *pos, const double *rad,
                      const double *pt, double probe, int self)
{
    int ic[3];
    for (int d = 0; d < 3; d++) ic[d] = (int)floor((pt[d] - c->lo[d]) / c->cs);
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                int x = ic[0] + dx, y = ic[1] + dy, z = ic[2] + dz;
                if (x < 0 || y < 0 || z < 0 || x >= c->dim[0] || y >= c->dim[1] || z >= c->dim[2]) continue;
                int id = x + c->dim[0] * (y + c->dim[1] * z);
                for (int q = c->start[id]; q < c->start[id + 1]; q++) {
                    int a = c->list[q];
                    if (a == self) continue;
                    double d2 = OSQR(pt[0] - pos[3 * a]) + OSQR(pt[1] - pos[3 * a + 1]) +
                                OSQR(pt[2] - pos[3 * a + 2]);
                    if (d2 < OSQR(rad[a] + probe)) return 0;
                }
            }
    return 1;
}

/* ------------------------------------------------------------- markSphere
 * vpmg.c:6444-6491 */
static void mark_sphere(double rtot, const double *tpos, int nx, int ny, int nz,
                        double hx, double hy, double hz, double xmin, double ymin,
                        double zmin, double *array, double markVal)
{
    double posx = tpos[0] - xmin, posy = tpos[1] - ymin, posz = tpos[2] - zmin;
    double rtot2 = OSQR(rtot);
    double xr = rtot + 0.5 * hx, yr = rtot + 0.5 * hy, zr = rtot + 0.5 * hz;
    int imin = OMAX(0, (int)ceil((posx - xr) / hx));
    int jmin = OMAX(0, (int)ceil((posy - yr) / hy));
    int kmin = OMAX(0, (int)ceil((posz - zr) / hz));
    int imax = OMIN(nx - 1, (int)floor((posx + xr) / hx));
    int jmax = OMIN(ny - 1, (int)floor((posy + yr) / hy));
    int kmax = OMIN(nz - 1, (int)floor((posz + zr) / hz));
    double fi, fj, fk;
    int i, j, k;
    for (i = imin, fi = imin; i <= imax; i++, fi += 1.) {
        double dx2 = OSQR(posx - hx * fi);
        for (j = jmin, fj = jmin; j <= jmax; j++, fj += 1.) {
            double dy2 = OSQR(posy - hy * fj);
            if ((dx2 + dy2) > rtot2) continue;
            for (k = kmin, fk = kmin; k <= kmax; k++, fk += 1.) {
                double dz2 = OSQR(posz - hz * fk);
                if ((dz2 + dy2 + dx2) <= rtot2) array[IJK(i, j, k)] = markVal;
            }
        }
    }
}

static int off_mesh_strict(const orc_pmg *m, const double *a)
{
    return (a[0] <= m->xmin) || (a[0] >= m->xmax) || (a[1] <= m->ymin) || (a[1] >= m->ymax) ||
           (a[2] <= m->zmin) || (a[2] >= m->zmax);
}

/* fillcoCoefMolIon (vpmg.c:4237-4333) */
static void fill_mol_ion(orc_pmg *m)
{
    int nx = m->nx, ny = m->ny, nz = m->nz;
    double irad = m->maxIonRadius;
    double ionmask = (m->ionstr > O_VPMGSMALL) ? 1.0 : 0.0;
    for (size_t i = 0; i < m->n; i++) m->kappa[i] = ionmask;
    if (m->ionstr < O_VPMGSMALL) return;
    for (int a = 0; a < m->natoms; a++) {
        const double *ap = m->apos + 3 * a;
        double ar = m->arad[a];
        if (ar > O_VSMALL) {
            if ((ap[0] < (m->xmin - irad - ar)) || (ap[0] > (m->xmax + irad + ar)) ||
                (ap[1] < (m->ymin - irad - ar)) || (ap[1] > (m->ymax + irad + ar)) ||
                (ap[2] < (m->zmin - irad - ar)) || (ap[2] > (m->zmax + irad + ar)))
                continue;
            mark_sphere(irad + ar, ap, nx, ny, nz, m->hx, m->hy, m->hz, m->xmin, m->ymin, m->zmin,
                        m->kappa, 0.0);
        }
    }
}

/* fillcoCoefMolDielNoSmooth (vpmg.c:4346-4498) */
static void fill_mol_diel(orc_pmg *m)
{
    int nx = m->nx, ny = m->ny, nz = m->nz;
    double srad = m->solventRadius, epsw = m->p.sdie, epsp = m->p.pdie;
    for (size_t i = 0; i < m->n; i++) m->epsx[i] = m->epsy[i] = m->epsz[i] = epsw;
    for (int a = 0; a < m->natoms; a++) {
        const double *ap = m->apos + 3 * a;
        double ar = m->arad[a];
        if (off_mesh_strict(m, ap)) continue;
        if (ar > O_VSMALL) {
            mark_sphere(ar + srad, ap, nx, ny, nz, m->hx, m->hy, m->hz, m->xmin + 0.5 * m->hx,
                        m->ymin, m->zmin, m->epsx, epsp);
            mark_sphere(ar + srad, ap, nx, ny, nz, m->hx, m->hy, m->hz, m->xmin,
                        m->ymin + 0.5 * m->hy, m->zmin, m->epsy, epsp);
            mark_sphere(ar + srad, ap, nx, ny, nz, m->hx, m->hy, m->hz, m->xmin, m->ymin,
                        m->zmin + 0.5 * m->hz, m->epsz, epsp);
        }
    }
    if (srad > O_VSMALL) {
        int nref = orc_ref_sphere(m->nsphere, NULL, 0);
        double *ref = (double *)malloc((size_t)3 * nref * sizeof(double));
        orc_ref_sphere(m->nsphere, ref, nref);
        ocell c;
        cell_build(&c, m->apos, m->arad, m->natoms, srad);
        for (int a = 0; a < m->natoms; a++) {
            double ar = m->arad[a];
            if (ar < O_VSMALL) continue;             /* Vacc_atomSurf: empty */
            double rad = ar + srad;
            const double *ap = m->apos + 3 * a;
            for (int q = 0; q < nref; q++) {
                double pt[3];
                pt[0] = rad * ref[3 * q] + ap[0];
                pt[1] = rad * ref[3 * q + 1] + ap[1];
                pt[2] = rad * ref[3 * q + 2] + ap[2];
                if (!acc_exclus(&c, m->apos, m->arad, pt, srad, a)) continue;
                mark_sphere(srad, pt, nx, ny, nz, m->hx, m->hy, m->hz, m->xmin + 0.5 * m->hx,
                            m->ymin, m->zmin, m->epsx, epsw);
                mark_sphere(srad, pt, nx, ny, nz, m->hx, m->hy, m->hz, m->xmin,
                            m->ymin + 0.5 * m->hy, m->zmin, m->epsy, epsw);
                mark_sphere(srad, pt, nx, ny, nz, m->hx, m->hy, m->hz, m->xmin, m->ymin,
                            m->zmin + 0.5 * m->hz, m->epsz, epsw);
            }
        }
        cell_free(&c);
        free(ref);
    }
}

/* fillcoCoefMolDielSmooth (vpmg.c:4500-4628): 9-point harmonic average */
static void fill_mol_smooth(orc_pmg *m)
{
    int nx = m->nx, ny = m->ny, nz = m->nz;
    double epsw = m->p.sdie;
    double *a1 = (double *)malloc(m->n * sizeof(double));
    double *a2 = (double *)malloc(m->n * sizeof(double));
    double *a3 = (double *)malloc(m->n * sizeof(double));
    for (size_t i = 0; i < m->n; i++) {
        a1[i] = m->epsx[i]; a2[i] = m->epsy[i]; a3[i] = m->epsz[i];
        m->epsx[i] = m->epsy[i] = m->epsz[i] = epsw;
    }
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                double frac;
                int np;
                frac = 1.0 / a1[IJK(i, j, k)];
                frac += 1.0 / a2[IJK(i, j, k)];
                frac += 1.0 / a3[IJK(i, j, k)];
                np = 3;
                if (j > 0) { frac += 1.0 / a2[IJK(i, j - 1, k)]; np += 1; }
                if (k > 0) { frac += 1.0 / a3[IJK(i, j, k - 1)]; np += 1; }
                if (i < nx - 1) {
                    frac += 1.0 / a2[IJK(i + 1, j, k)];
                    frac += 1.0 / a3[IJK(i + 1, j, k)];
                    np += 2;
                    if (j > 0) { frac += 1.0 / a2[IJK(i + 1, j - 1, k)]; np += 1; }
                    if (k > 0) { frac += 1.0 / a3[IJK(i + 1, j, k - 1)]; np += 1; }
                }
                m->epsx[IJK(i, j, k)] = np / frac;

                frac = 1.0 / a2[IJK(i, j, k)];
                frac += 1.0 / a1[IJK(i, j, k)];
                frac += 1.0 / a3[IJK(i, j, k)];
                np = 3;
                if (i > 0) { frac += 1.0 / a1[IJK(i - 1, j, k)]; np += 1; }
                if (k > 0) { frac += 1.0 / a3[IJK(i, j, k - 1)]; np += 1; }
                if (j < ny - 1) {
                    frac += 1.0 / a1[IJK(i, j + 1, k)];
                    frac += 1.0 / a3[IJK(i, j + 1, k)];
                    np += 2;
                    if (i > 0) { frac += 1.0 / a1[IJK(i - 1, j + 1, k)]; np += 1; }
                    if (k > 0) { frac += 1.0 / a3[IJK(i, j + 1, k - 1)]; np += 1; }
                }
                m->epsy[IJK(i, j, k)] = np / frac;

                frac = 1.0 / a3[IJK(i, j, k)];
                frac += 1.0 / a1[IJK(i, j, k)];
                frac += 1.0 / a2[IJK(i, j, k)];
                np = 3;
                if (i > 0) { frac += 1.0 / a1[IJK(i - 1, j, k)]; np += 1; }
                if (j > 0) { frac += 1.0 / a2[IJK(i, j - 1, k)]; np += 1; }
                if (k < nz - 1) {
                    frac += 1.0 / a1[IJK(i, j, k + 1)];
                    frac += 1.0 / a2[IJK(i, j, k + 1)];
                    np += 2;
                    if (i > 0) { frac += 1.0 / a1[IJK(i - 1, j, k + 1)]; np += 1; }
                    if (j > 0) { frac += 1.0 / a2[IJK(i, j - 1, k + 1)]; np += 1; }
                }
                m->epsz[IJK(i, j, k)] = np / frac;
            }
    free(a1); free(a2); free(a3);
}

/* Coefficients of the polynomial windows of srfm spl3 and spl4 over
 * [b, e] = [r - win, r + win]: fillcoCoefSpline3 (vpmg.c:7597-7619) and
 * fillcoCoefSpline4 (vpmg.c:7124-7147); Vacc_splineAccGradAtomNorm3 / 4
 * (vacc.c:1097, 1004) use the same formulas. */
static void win_coef(int order, double b, double e, double *c)
{
    for (int n = 0; n < 8; n++) c[n] = 0.0;
    double e2 = e * e, e3 = e2 * e, e4 = e3 * e, e5 = e4 * e;
    double b2 = b * b, b3 = b2 * b, b4 = b3 * b, b5 = b4 * b;
    if (order == 3) {
        double denom = pow((e - b), 5.0);
        c[0] = -10.0 * e2 * b3 + 5.0 * e * b4 - b5;
        c[1] = 30.0 * e2 * b2;
        c[2] = -30.0 * (e2 * b + e * b2);
        c[3] = 10.0 * (e2 + 4.0 * e * b + b2);
        c[4] = -15.0 * (e + b);
        c[5] = 6;
        for (int n = 0; n < 6; n++) c[n] = c[n] / denom;
    } else {
        double e6 = e5 * e, e7 = e6 * e, b6 = b5 * b, b7 = b6 * b;
        double denom = e7 - 7.0 * b * e6 + 21.0 * b2 * e5 - 35.0 * e4 * b3 + 35.0 * e3 * b4 -
                       21.0 * b5 * e2 + 7.0 * e * b6 - b7;
        c[0] = b4 * (35.0 * e3 - 21.0 * b * e2 + 7 * e * b2 - b3) / denom;
        c[1] = -140.0 * b3 * e3 / denom;
        c[2] = 210.0 * e2 * b2 * (e + b) / denom;
        c[3] = -140.0 * e * b * (e2 + 3.0 * b * e + b2) / denom;
        c[4] = 35.0 * (e3 + 9.0 * b * e2 + 9.0 * e * b2 + b3) / denom;
        c[5] = -84.0 * (e2 + 3.0 * b * e + b2) / denom;
        c[6] = 70.0 * (e + b) / denom;
        c[7] = -20.0 / denom;
    }
}

/* the clamped window value of spl3 / spl4 at distance sm (with sm2 as the
 * branch of the reference forms it) */
static double win_chi(int order, const double *c, double sm, double sm2)
{
    double sm3 = sm2 * sm, sm4 = sm3 * sm, sm5 = sm4 * sm, v;
    if (order == 3) {
        v = c[0] + c[1] * sm + c[2] * sm2 + c[3] * sm3 + c[4] * sm4 + c[5] * sm5;
    } else {
        double sm6 = sm5 * sm, sm7 = sm6 * sm;
        v = c[0] + c[1] * sm + c[2] * sm2 + c[3] * sm3 + c[4] * sm4 + c[5] * sm5 + c[6] * sm6 +
            c[7] * sm7;
    }
    if (v > 1.0) v = 1.0;
    else if (v < 0.0) v = 0.0;
    return v;
}

/* fillcoCoefSpline (vpmg.c:4631-4854), order 2: multiplicative cubic window
 * per atom; order 3 / 4: fillcoCoefSpline3 (vpmg.c:7525-7900) /
 * fillcoCoefSpline4 (vpmg.c:7052-7372), the same loops with polynomial
 * windows of the distance.  spl3's kappa window keeps the reference's
 * coefficients: the dielectric ones divided again by the ion denominator
 * (vpmg.c:7632-7643). */
static void fill_spline(orc_pmg *m, int order)
{
    int nx = m->nx, ny = m->ny, nz = m->nz;
    double win = m->p.swin;
    double w2i = 1.0 / (win * win), w3i = 1.0 / (win * win * win);
    double irad = m->maxIonRadius, epsw = m->p.sdie, epsp = m->p.pdie;
    double hx = m->hx, hy = m->hy, hzed = m->hz;
    double ionmask = (m->ionstr > O_VPMGSMALL) ? 1.0 : 0.0;
    for (size_t i = 0; i < m->n; i++) m->kappa[i] = m->epsx[i] = m->epsy[i] = m->epsz[i] = 1.0;
    for (int a = 0; a < m->natoms; a++) {
        const double *ap = m->apos + 3 * a;
        double arad = m->arad[a];
        if (off_mesh_strict(m, ap)) continue;
        if (!(arad > O_VPMGSMALL)) continue;
        double cd[8], ci[8];
        if (order > 2) {
            win_coef(order, arad - win, arad + win, cd);
            if (order == 3) {
                double b = irad + arad - win, e = irad + arad + win;
                double denom = pow((e - b), 5.0);
                for (int n = 0; n < 8; n++) ci[n] = n < 6 ? cd[n] / denom : 0.0;
            } else {
                win_coef(order, irad + arad - win, irad + arad + win, ci);
            }
        }
        double pos[3] = {ap[0] - m->xmin, ap[1] - m->ymin, ap[2] - m->zmin};
        double itot = irad + arad + win, itot2 = OSQR(itot);
        double ictot = OMAX(0, (irad + arad - win)), ictot2 = OSQR(ictot);
        double stot = arad + win, stot2 = OSQR(stot);
        double sctot = OMAX(0, (arad - win)), sctot2 = OSQR(sctot);
        double rtot = OMAX(itot, stot);
        double dx = rtot + 0.5 * hx, dy = rtot + 0.5 * hy, dz = rtot + 0.5 * hzed;
        int imin = OMAX(0, (int)floor((pos[0] - dx) / hx));
        int imax = OMIN(nx - 1, (int)ceil((pos[0] + dx) / hx));
        int jmin = OMAX(0, (int)floor((pos[1] - dy) / hy));
        int jmax = OMIN(ny - 1, (int)ceil((pos[1] + dy) / hy));
        int kmin = OMAX(0, (int)floor((pos[2] - dz) / hzed));
        int kmax = OMIN(nz - 1, (int)ceil((pos[2] + dz) / hzed));
        for (int i = imin; i <= imax; i++) {
            double dx2 = OSQR(pos[0] - hx * i);
            for (int j = jmin; j <= jmax; j++) {
                double dy2 = OSQR(pos[1] - hy * j);
                for (int k = kmin; k <= kmax; k++) {
                    double dz2 = OSQR(pos[2] - k * hzed);
                    size_t p = IJK(i, j, k);
                    if (m->kappa[p] > O_VPMGSMALL) {
                        double d2 = dz2 + dy2 + dx2;
                        if (d2 <= ictot2) m->kappa[p] = 0.0;
                        if ((d2 < itot2) && (d2 > ictot2)) {
                            if (order == 2) {
                                double dist = sqrt(d2), sm = dist - (arad + irad) + win, sm2 = OSQR(sm);
                                m->kappa[p] *= 0.75 * sm2 * w2i - 0.25 * sm * sm2 * w3i;
                            } else {
                                m->kappa[p] *= win_chi(order, ci, sqrt(d2), d2);
                            }
                        }
                    }
                    if (m->epsx[p] > O_VPMGSMALL) {
                        double d2 = dz2 + dy2 + OSQR(pos[0] - (i + 0.5) * hx);
                        if (d2 >= stot2) m->epsx[p] *= 1.0;
                        if (d2 <= sctot2) m->epsx[p] = 0.0;
                        if ((d2 > sctot2) && (d2 < stot2)) {
                            if (order == 2) {
                                double dist = sqrt(d2), sm = dist - arad + win, sm2 = OSQR(sm);
                                m->epsx[p] *= 0.75 * sm2 * w2i - 0.25 * sm * sm2 * w3i;
                            } else {
                                double sm = sqrt(d2);
                                m->epsx[p] *= win_chi(order, cd, sm, OSQR(sm));
                            }
                        }
                    }
                    if (m->epsy[p] > O_VPMGSMALL) {
                        double d2 = dz2 + dx2 + OSQR(pos[1] - (j + 0.5) * hy);
                        if (d2 >= stot2) m->epsy[p] *= 1.0;
                        if (d2 <= sctot2) m->epsy[p] = 0.0;
                        if ((d2 > sctot2) && (d2 < stot2)) {
                            if (order == 2) {
                                double dist = sqrt(d2), sm = dist - arad + win, sm2 = OSQR(sm);
                                m->epsy[p] *= 0.75 * sm2 * w2i - 0.25 * sm * sm2 * w3i;
                            } else {
                                double sm = sqrt(d2);
                                m->epsy[p] *= win_chi(order, cd, sm, OSQR(sm));
                            }
                        }
                    }
                    if (m->epsz[p] > O_VPMGSMALL) {
                        double d2 = dy2 + dx2 + OSQR(pos[2] - (k + 0.5) * hzed);
                        if (d2 >= stot2) m->epsz[p] *= 1.0;
                        if (d2 <= sctot2) m->epsz[p] = 0.0;
                        if ((d2 > sctot2) && (d2 < stot2)) {
                            if (order == 2) {
                                double dist = sqrt(d2), sm = dist - arad + win, sm2 = OSQR(sm);
                                m->epsz[p] *= 0.75 * sm2 * w2i - 0.25 * sm * sm2 * w3i;
                            } else {
                                double sm = sqrt(d2);
                                m->epsz[p] *= win_chi(order, cd, sm, d2);
                            }
                        }
                    }
                }
            }
        }
    }
    for (size_t p = 0; p < m->n; p++) {
        m->kappa[p] = ionmask * m->kappa[p];
        m->epsx[p] = (epsw - epsp) * m->epsx[p] + epsp;
        m->epsy[p] = (epsw - epsp) * m->epsy[p] + epsp;
        m->epsz[p] = (epsw - epsp) * m->epsz[p] + epsp;
    }
}

/* ---------------------------------------------------------------- charges */

/* bspline2 (vpmg.c:5091-5104) and its derivative (:5106-5121) */
static double bspline2(double x)
{
    double m2m, m2, m3;
    if ((x >= 0.0) && (x <= 2.0)) m2m = 1.0 - fabs(x - 1.0); else m2m = 0.0;
    if ((x >= 1.0) && (x <= 3.0)) m2 = 1.0 - fabs(x - 2.0); else m2 = 0.0;
    if ((x >= 0.0) && (x <= 3.0)) m3 = 0.5 * x * m2m + 0.5 * (3.0 - x) * m2; else m3 = 0.0;
    return m3;
}

static double dbspline2(double x)
{
    double m2m, m2;
    if ((x >= 0.0) && (x <= 2.0)) m2m = 1.0 - fabs(x - 1.0); else m2m = 0.0;
    if ((x >= 1.0) && (x <= 3.0)) m2 = 1.0 - fabs(x - 2.0); else m2 = 0.0;
    return m2m - m2;
}

#define VFCHI(iint, iflt) (1.5 + ((double)(iint) - (iflt)))

/* fillcoChargeSpline2 (vpmg.c:5123-5248) */
static void fill_charge_spl2(orc_pmg *m)
{
    int nx = m->nx, ny = m->ny, nz = m->nz;
    double hx = m->hx, hy = m->hy, hzed = m->hz;
    for (size_t i = 0; i < m->n; i++) m->charge[i] = 0.0;
    for (int a = 0; a < m->natoms; a++) {
        const double *ap = m->apos + 3 * a;
        if ((ap[0] <= (m->xmin - hx)) || (ap[0] >= (m->xmax + hx)) || (ap[1] <= (m->ymin - hy)) ||
            (ap[1] >= (m->ymax + hy)) || (ap[2] <= (m->zmin - hzed)) || (ap[2] >= (m->zmax + hzed)))
            continue;
        double pos[3] = {ap[0] - m->xmin, ap[1] - m->ymin, ap[2] - m->zmin};
        double charge = m->achg[a] * m->zmagic / (hx * hy * hzed);
        double fi = pos[0] / hx, fj = pos[1] / hy, fk = pos[2] / hzed;
        int ip1 = (int)ceil(fi), ip2 = ip1 + 1, im1 = (int)floor(fi), im2 = im1 - 1;
        int jp1 = (int)ceil(fj), jp2 = jp1 + 1, jm1 = (int)floor(fj), jm2 = jm1 - 1;
        int kp1 = (int)ceil(fk), kp2 = kp1 + 1, km1 = (int)floor(fk), km2 = km1 - 1;
        ip2 = OMIN(ip2, nx - 1); ip1 = OMIN(ip1, nx - 1); im1 = OMAX(im1, 0); im2 = OMAX(im2, 0);
        jp2 = OMIN(jp2, ny - 1); jp1 = OMIN(jp1, ny - 1); jm1 = OMAX(jm1, 0); jm2 = OMAX(jm2, 0);
        kp2 = OMIN(kp2, nz - 1); kp1 = OMIN(kp1, nz - 1); km1 = OMAX(km1, 0); km2 = OMAX(km2, 0);
        (void)ip1; (void)jp1; (void)kp1; (void)im1; (void)jm1; (void)km1;
        for (int ii = im2; ii <= ip2; ii++) {
            double mx = bspline2(VFCHI(ii, fi));
            for (int jj = jm2; jj <= jp2; jj++) {
                double my = bspline2(VFCHI(jj, fj));
                for (int kk = km2; kk <= kp2; kk++) {
                    double mz = bspline2(VFCHI(kk, fk));
                    m->charge[IJK(ii, jj, kk)] += (charge * mx * my * mz);
                }
            }
        }
    }
}

/* bspline4 / dbspline4 / d2bspline4 (vpmg.c:6731-6822): the 5th-order
 * B-spline of chgm spl4 and its derivatives, at VFCHI4(i, f) = 2.5 + (i - f)
 * (vpmg.h:939). */
static double bspline4(double x)
{
    double m, m2;
    double one6 = 1.0 / 6.0, one8 = 1.0 / 8.0, one24 = 1.0 / 24.0;
    double thirteen24 = 13.0 / 24.0, fourtyseven24 = 47.0 / 24.0, seventeen24 = 17.0 / 24.0;
    if ((x > 0.0) && (x <= 1.0)) { m = x * x; return one24 * m * m; }
    if ((x > 1.0) && (x <= 2.0)) {
        m = x - 1.0; m2 = m * m;
        return -one8 + one6 * x + m2 * (0.25 + one6 * m - one6 * m2);
    }
    if ((x > 2.0) && (x <= 3.0)) {
        m = x - 2.0; m2 = m * m;
        return -thirteen24 + 0.5 * x + m2 * (-0.25 - 0.5 * m + 0.25 * m2);
    }
    if ((x > 3.0) && (x <= 4.0)) {
        m = x - 3.0; m2 = m * m;
        return fourtyseven24 - 0.5 * x + m2 * (-0.25 + 0.5 * m - one6 * m2);
    }
    if ((x > 4.0) && (x <= 5.0)) {
        m = x - 4.0; m2 = m * m;
        return seventeen24 - one6 * x + m2 * (0.25 - one6 * m + one24 * m2);
    }
    return 0.0;
}

static double dbspline4(double x)
{
    double m, m2;
    double one6 = 1.0 / 6.0, one3 = 1.0 / 3.0, two3 = 2.0 / 3.0, thirteen6 = 13.0 / 6.0;
    if ((x > 0.0) && (x <= 1.0)) { m2 = x * x; return one6 * x * m2; }
    if ((x > 1.0) && (x <= 2.0)) { m = x - 1.0; m2 = m * m; return -one3 + 0.5 * x + m2 * (0.5 - two3 * m); }
    if ((x > 2.0) && (x <= 3.0)) { m = x - 2.0; m2 = m * m; return 1.5 - 0.5 * x + m2 * (-1.5 + m); }
    if ((x > 3.0) && (x <= 4.0)) { m = x - 3.0; m2 = m * m; return 1.0 - 0.5 * x + m2 * (1.5 - two3 * m); }
    if ((x > 4.0) && (x <= 5.0)) { m = x - 4.0; m2 = m * m; return -thirteen6 + 0.5 * x + m2 * (-0.5 + one6 * m); }
    return 0.0;
}

static double d2bspline4(double x)
{
    double m, m2;
    if ((x > 0.0) && (x <= 1.0)) return 0.5 * x * x;
    if ((x > 1.0) && (x <= 2.0)) { m = x - 1.0; m2 = m * m; return -0.5 + x - 2.0 * m2; }
    if ((x > 2.0) && (x <= 3.0)) { m = x - 2.0; m2 = m * m; return 5.5 - 3.0 * x + 3.0 * m2; }
    if ((x > 3.0) && (x <= 4.0)) { m = x - 3.0; m2 = m * m; return -9.5 + 3.0 * x - 2.0 * m2; }
    if ((x > 4.0) && (x <= 5.0)) { m = x - 4.0; m2 = m * m; return 4.5 - x + 0.5 * m2; }
    return 0.0;
}

#define VFCHI4(iint, iflt) (2.5 + ((double)(iint) - (iflt)))

/* chgm spl4: fillcoPermanentMultipole (vpmg.c:6835-7050) with the atom
 * charge only; the dipole and quadrupole terms are zero and stay in the sum */
static void fill_charge_spl4(orc_pmg *m)
{
    int nx = m->nx, ny = m->ny, nz = m->nz;
    double hx = m->hx, hy = m->hy, hzed = m->hz;
    double f = m->zmagic / (hx * hy * hzed);
    for (size_t i = 0; i < m->n; i++) m->charge[i] = 0.0;
    for (int a = 0; a < m->natoms; a++) {
        const double *ap = m->apos + 3 * a;
        double c = m->achg[a] * f;
        double ux = 0.0, uy = 0.0, uz = 0.0, qxx = 0.0, qyx = 0.0, qyy = 0.0, qzx = 0.0,
               qzy = 0.0, qzz = 0.0;
        if ((ap[0] <= (m->xmin - 2 * hx)) || (ap[0] >= (m->xmax + 2 * hx)) ||
            (ap[1] <= (m->ymin - 2 * hy)) || (ap[1] >= (m->ymax + 2 * hy)) ||
            (ap[2] <= (m->zmin - 2 * hzed)) || (ap[2] >= (m->zmax + 2 * hzed)))
            continue;
        double pos[3] = {ap[0] - m->xmin, ap[1] - m->ymin, ap[2] - m->zmin};
        double fi = pos[0] / hx, fj = pos[1] / hy, fk = pos[2] / hzed;
        int ip1 = (int)ceil(fi), ip2 = ip1 + 2, im1 = (int)floor(fi), im2 = im1 - 2;
        int jp1 = (int)ceil(fj), jp2 = jp1 + 2, jm1 = (int)floor(fj), jm2 = jm1 - 2;
        int kp1 = (int)ceil(fk), kp2 = kp1 + 2, km1 = (int)floor(fk), km2 = km1 - 2;
        ip2 = OMIN(ip2, nx - 1); im2 = OMAX(im2, 0);
        jp2 = OMIN(jp2, ny - 1); jm2 = OMAX(jm2, 0);
        kp2 = OMIN(kp2, nz - 1); km2 = OMAX(km2, 0);
        for (int ii = im2; ii <= ip2; ii++) {
            double mi = VFCHI4(ii, fi);
            double mx = bspline4(mi), dmx = dbspline4(mi), d2mx = d2bspline4(mi);
            for (int jj = jm2; jj <= jp2; jj++) {
                double mj = VFCHI4(jj, fj);
                double my = bspline4(mj), dmy = dbspline4(mj), d2my = d2bspline4(mj);
                for (int kk = km2; kk <= kp2; kk++) {
                    double mk = VFCHI4(kk, fk);
                    double mz = bspline4(mk), dmz = dbspline4(mk), d2mz = d2bspline4(mk);
                    double q = mx * my * mz * c - dmx * my * mz * ux - mx * dmy * mz * uy -
                               mx * my * dmz * uz + d2mx * my * mz * qxx + dmx * dmy * mz * qyx +
                               mx * d2my * mz * qyy + dmx * my * dmz * qzx + mx * dmy * dmz * qzy +
                               mx * my * d2mz * qzz;
                    m->charge[IJK(ii, jj, kk)] += q;
                }
            }
        }
    }
}

/* fillcoChargeSpline1 (vpmg.c:4986-5089): trilinear */
static void fill_charge_spl0(orc_pmg *m)
{
    int nx = m->nx, ny = m->ny;
    double hx = m->hx, hy = m->hy, hzed = m->hz;
    for (size_t i = 0; i < m->n; i++) m->charge[i] = 0.0;
    for (int a = 0; a < m->natoms; a++) {
        const double *ap = m->apos + 3 * a;
        if (off_mesh_strict(m, ap)) continue;
        double pos[3] = {ap[0] - m->xmin, ap[1] - m->ymin, ap[2] - m->zmin};
        double charge = m->achg[a] * m->zmagic / (hx * hy * hzed);
        double fi = pos[0] / hx, fj = pos[1] / hy, fk = pos[2] / hzed;
        int ihi = (int)ceil(fi), ilo = (int)floor(fi);
        int jhi = (int)ceil(fj), jlo = (int)floor(fj);
        int khi = (int)ceil(fk), klo = (int)floor(fk);
        double dx = fi - (double)(ilo), dy = fj - (double)(jlo), dz = fk - (double)(klo);
        m->charge[IJK(ihi, jhi, khi)] += (dx * dy * dz * charge);
        m->charge[IJK(ihi, jlo, khi)] += (dx * (1.0 - dy) * dz * charge);
        m->charge[IJK(ihi, jhi, klo)] += (dx * dy * (1.0 - dz) * charge);
        m->charge[IJK(ihi, jlo, klo)] += (dx * (1.0 - dy) * (1.0 - dz) * charge);
        m->charge[IJK(ilo, jhi, khi)] += ((1.0 - dx) * dy * dz * charge);
        m->charge[IJK(ilo, jlo, khi)] += ((1.0 - dx) * (1.0 - dy) * dz * charge);
        m->charge[IJK(ilo, jhi, klo)] += ((1.0 - dx) * dy * (1.0 - dz) * charge);
        m->charge[IJK(ilo, jlo, klo)] += ((1.0 - dx) * (1.0 - dy) * (1.0 - dz) * charge);
    }
}

/* ------------------------------------------------------ boundary conditions */

/* bcflnew (vpmg.c:3289-3367): multiple Debye-Hueckel spheres on the faces */
static void bc_mdh(orc_pmg *m)
{
    int nx = m->nx, ny = m->ny, nz = m->nz;
    double pre1 = ((O_Vunit_ec) / (4 * O_VPI * O_Vunit_eps0 * m->p.sdie * O_Vunit_kb * m->p.temp)) * (1.0e10);
    double xk = m->xkappa;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                if (!((k == 0) || (k == nz - 1) || (j == 0) || (j == ny - 1) || (i == 0) || (i == nx - 1)))
                    continue;
                double gx = m->xf[i], gy = m->yf[j], gz = m->zf[k];
                double val = 0.0;
                for (int a = 0; a < m->natoms; a++) {
                    double q = O_Vunit_ec * m->achg[a];
                    double dist = sqrt(OSQR(gx - m->apos[3 * a]) + OSQR(gy - m->apos[3 * a + 1]) +
                                       OSQR(gz - m->apos[3 * a + 2]));
                    if (xk > O_VSMALL)
                        val += pre1 * (q / dist) * exp(-xk * (dist - m->arad[a])) / (1 + xk * m->arad[a]);
                    else
                        val += pre1 * (q / dist);
                }
                if (i == 0) m->gxcf[j + ny * k] += val;
                if (i == nx - 1) m->gxcf[j + ny * k + (size_t)ny * nz] += val;
                if (j == 0) m->gycf[i + nx * k] += val;
                if (j == ny - 1) m->gycf[i + nx * k + (size_t)nx * nz] += val;
                if (k == 0) m->gzcf[i + nx * j] += val;
                if (k == nz - 1) m->gzcf[i + nx * j + (size_t)nx * ny] += val;
            }
}

/* multipolebc (vpmg.c:3369-3477): the radial factors of the monopole,
 * dipole and traceless-quadrupole potentials outside a sphere of radius rad
 * and dielectric eps_p in a solvent eps_w with Debye-Hueckel kappa. */
static void multipole_tensor(double r, double kappa, double eps_p, double eps_w, double rad,
                             double tsr[3])
{
    double eps_r = eps_w / eps_p;
    double r2 = r * r, r3 = r2 * r, r5 = r3 * r2;
    tsr[0] = (1.0 / eps_w) / r;
    tsr[1] = (1.0 / eps_w) * (-1.0) / r3;
    tsr[2] = (1.0 / eps_w) * (3.0) / r5;
    if (kappa < O_VSMALL) {
        tsr[1] = (3.0 * eps_r) / (1.0 + 2.0 * eps_r) * tsr[1];
        tsr[2] = (5.0 * eps_r) / (2.0 + 3.0 * eps_r) * tsr[2];
    } else {
        double ka = kappa * rad, ka2 = ka * ka, ka3 = ka2 * ka;
        double kr = kappa * r, kr2 = kr * kr;
        tsr[0] = exp(ka - kr) / (1.0 + ka) * tsr[0];
        tsr[1] = 3.0 * eps_r * exp(ka - kr) * (1.0 + kr) * tsr[1];
        tsr[1] = tsr[1] / (1.0 + ka + eps_r * (2.0 + 2.0 * ka + ka2));
        tsr[2] = 5.0 * eps_r * exp(ka - kr) * (3.0 + 3.0 * kr + kr2) * tsr[2];
        tsr[2] = tsr[2] / (6.0 + 6.0 * ka + 2.0 * ka2 + eps_r * (9.0 + 9.0 * ka + 4.0 * ka2 + ka3));
    }
}

/* bcfl_sdh (vpmg.c:3479-3645): one sphere at the solute centre with the
 * solute radius, carrying the molecule's charge, dipole and traceless
 * quadrupole (atom order sums). */
static void bc_sdh(orc_pmg *m)
{
    int nx = m->nx, ny = m->ny, nz = m->nz;
    double pre = (O_Vunit_ec * O_Vunit_ec) / (4 * O_VPI * O_Vunit_eps0 * O_Vunit_kb * m->p.temp);
    pre = pre * (1.0e10);
    double xk = m->xkappa, size = m->soluteRadius;
    const double *pos = m->center;
    double q = 0.0, dip[3] = {0, 0, 0}, quad[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, tr[9];
    for (int a = 0; a < m->natoms; a++) {
        double xr = m->apos[3 * a] - pos[0], yr = m->apos[3 * a + 1] - pos[1],
               zr = m->apos[3 * a + 2] - pos[2];
        double c = m->achg[a];
        q += c;
        dip[0] += xr * c;
        dip[1] += yr * c;
        dip[2] += zr * c;
        tr[0] = xr * xr * c; tr[1] = xr * yr * c; tr[2] = xr * zr * c;
        tr[3] = yr * xr * c; tr[4] = yr * yr * c; tr[5] = yr * zr * c;
        tr[6] = zr * xr * c; tr[7] = zr * yr * c; tr[8] = zr * zr * c;
        double qave = (tr[0] + tr[4] + tr[8]) / 3.0;
        quad[0] += 1.5 * (tr[0] - qave);
        quad[1] += 1.5 * (tr[1]);
        quad[2] += 1.5 * (tr[2]);
        quad[3] += 1.5 * (tr[3]);
        quad[4] += 1.5 * (tr[4] - qave);
        quad[5] += 1.5 * (tr[5]);
        quad[6] += 1.5 * (tr[6]);
        quad[7] += 1.5 * (tr[7]);
        quad[8] += 1.5 * (tr[8] - qave);
    }
    double ux = dip[0], uy = dip[1], uz = dip[2];
    double qxx = quad[0] / 3.0, qxy = quad[1] / 3.0, qxz = quad[2] / 3.0;
    double qyy = quad[4] / 3.0, qyz = quad[5] / 3.0, qzz = quad[8] / 3.0;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                if (!((k == 0) || (k == nz - 1) || (j == 0) || (j == ny - 1) || (i == 0) || (i == nx - 1)))
                    continue;
                double xr = m->xf[i] - pos[0], yr = m->yf[j] - pos[1], zr = m->zf[k] - pos[2];
                double dist = sqrt(OSQR(xr) + OSQR(yr) + OSQR(zr));
                double t[3];
                multipole_tensor(dist, xk, m->p.pdie, m->p.sdie, size, t);
                double val = pre * q * t[0];
                val -= pre * ux * xr * t[1];
                val -= pre * uy * yr * t[1];
                val -= pre * uz * zr * t[1];
                val += pre * qxx * xr * xr * t[2];
                val += pre * qyy * yr * yr * t[2];
                val += pre * qzz * zr * zr * t[2];
                val += pre * 2.0 * qxy * xr * yr * t[2];
                val += pre * 2.0 * qxz * xr * zr * t[2];
                val += pre * 2.0 * qyz * yr * zr * t[2];
                if (i == 0) m->gxcf[j + ny * k] = val;
                if (i == nx - 1) m->gxcf[j + ny * k + (size_t)ny * nz] = val;
                if (j == 0) m->gycf[i + nx * k] = val;
                if (j == ny - 1) m->gycf[i + nx * k + (size_t)nx * nz] = val;
                if (k == 0) m->gzcf[i + nx * j] = val;
                if (k == nz - 1) m->gzcf[i + nx * j + (size_t)nx * ny] = val;
            }
}

/* ---------------------------------------------------------------- focusing */

static double trilin_old(const ogrid *o, double x, double y, double z, int *ok)
{
    *ok = (x >= (o->xmin - O_VSMALL)) && (y >= (o->ymin - O_VSMALL)) && (z >= (o->zmin - O_VSMALL)) &&
          (x <= (o->xmax + O_VSMALL)) && (y <= (o->ymax + O_VSMALL)) && (z <= (o->zmax + O_VSMALL));
    if (!*ok) return 0.0;
    int nx = o->nx, ny = o->ny;
    double fi = (x - o->xmin) / o->hx, fj = (y - o->ymin) / o->hy, fk = (z - o->zmin) / o->hz;
    int ihi = (int)ceil(fi); if (ihi > (o->nx - 1)) ihi = o->nx - 1;
    int ilo = (int)floor(fi); if (ilo < 0) ilo = 0;
    int jhi = (int)ceil(fj); if (jhi > (o->ny - 1)) jhi = o->ny - 1;
    int jlo = (int)floor(fj); if (jlo < 0) jlo = 0;
    int khi = (int)ceil(fk); if (khi > (o->nz - 1)) khi = o->nz - 1;
    int klo = (int)floor(fk); if (klo < 0) klo = 0;
    double dx = fi - (double)(ilo), dy = fj - (double)(jlo), dz = fk - (double)(klo);
    const double *d = o->data;
    return dx * dy * dz * (d[IJK(ihi, jhi, khi)]) + dx * (1.0 - dy) * dz * (d[IJK(ihi, jlo, khi)]) +
           dx * dy * (1.0 - dz) * (d[IJK(ihi, jhi, klo)]) + dx * (1.0 - dy) * (1.0 - dz) * (d[IJK(ihi, jlo, klo)]) +
           (1.0 - dx) * dy * dz * (d[IJK(ilo, jhi, khi)]) + (1.0 - dx) * (1.0 - dy) * dz * (d[IJK(ilo, jlo, khi)]) +
           (1.0 - dx) * dy * (1.0 - dz) * (d[IJK(ilo, jhi, klo)]) +
           (1.0 - dx) * (1.0 - dy) * (1.0 - dz) * (d[IJK(ilo, jlo, klo)]);
}

/* the old Vpmg's mesh (vpmg.c:1982-1996) */
static ogrid old_mesh(const orc_pmg *o)
{
    ogrid g;
    g.nx = o->nx; g.ny = o->ny; g.nz = o->nz;
    g.hx = o->hx; g.hy = o->hy; g.hz = o->hz;
    g.xmin = o->p.center[0] - ((double)(o->nx - 1) * o->hx) / 2.0;
    g.xmax = o->p.center[0] + ((double)(o->nx - 1) * o->hx) / 2.0;
    g.ymin = o->p.center[1] - ((double)(o->ny - 1) * o->hy) / 2.0;
    g.ymax = o->p.center[1] + ((double)(o->ny - 1) * o->hy) / 2.0;
    g.zmin = o->p.center[2] - ((double)(o->nz - 1) * o->hz) / 2.0;
    g.zmax = o->p.center[2] + ((double)(o->nz - 1) * o->hz) / 2.0;
    g.data = o->u;
    return g;
}

/* ------------------------------------------------------------- maps */

/* Vgrid_value (vgrid.c:179-260).  Vcompare = 10^-(VGRID_DIGITS - 2) = 1e-4
 * (vgrid.c:136, vgrid.h:70).  The reference keeps the indices in size_t, so a
 * point less than one spacing below the map passes its only test (hi < n) with
 * a wrapped lower index and reads outside the array; such a point is off the
 * map here. */
static int map_value(const ogrid *g, const double pt[3], double *value)
{
    const double Vcompare = 1e-4;
    double f[3] = {(pt[0] - g->xmin) / g->hx, (pt[1] - g->ymin) / g->hy, (pt[2] - g->zmin) / g->hz};
    const double lo[3] = {g->xmin, g->ymin, g->zmin}, hi[3] = {g->xmax, g->ymax, g->zmax};
    const int n[3] = {g->nx, g->ny, g->nz};
    long ih[3], il[3];
    for (int d = 0; d < 3; d++) {
        ih[d] = (long)ceil(f[d]);
        il[d] = (long)floor(f[d]);
    }
    for (int d = 0; d < 3; d++)
        if (fabs(pt[d] - lo[d]) < Vcompare) il[d] = 0;
    for (int d = 0; d < 3; d++)
        if (fabs(pt[d] - hi[d]) < Vcompare) ih[d] = n[d] - 1;
    for (int d = 0; d < 3; d++)
        if (ih[d] < 0 || ih[d] >= n[d] || il[d] < 0 || il[d] >= n[d]) return 0;
    int nx = g->nx, ny = g->ny;
    double dx = f[0] - (double)il[0], dy = f[1] - (double)il[1], dz = f[2] - (double)il[2];
    const double *u = g->data;
    long ihi = ih[0], jhi = ih[1], khi = ih[2], ilo = il[0], jlo = il[1], klo = il[2];
    *value = dx * dy * dz * (u[IJK(ihi, jhi, khi)]) + dx * (1.0 - dy) * dz * (u[IJK(ihi, jlo, khi)]) +
             dx * dy * (1.0 - dz) * (u[IJK(ihi, jhi, klo)]) +
             dx * (1.0 - dy) * (1.0 - dz) * (u[IJK(ihi, jlo, klo)]) +
             (1.0 - dx) * dy * dz * (u[IJK(ilo, jhi, khi)]) +
             (1.0 - dx) * (1.0 - dy) * dz * (u[IJK(ilo, jlo, khi)]) +
             (1.0 - dx) * dy * (1.0 - dz) * (u[IJK(ilo, jhi, klo)]) +
             (1.0 - dx) * (1.0 - dy) * (1.0 - dz) * (u[IJK(ilo, jlo, klo)]);
    return 1;
}

static void grid_of(ogrid *g, const int n[3], const double h[3], const double origin[3], double *data)
{
    g->nx = n[0]; g->ny = n[1]; g->nz = n[2];
    g->hx = h[0]; g->hy = h[1]; g->hz = h[2];
    g->xmin = origin[0]; g->ymin = origin[1]; g->zmin = origin[2];
    /* Vgrid_ctor2 (vgrid.c:121-126) */
    g->xmax = origin[0] + (n[0] - 1) * h[0];
    g->ymax = origin[1] + (n[1] - 1) * h[1];
    g->zmax = origin[2] + (n[2] - 1) * h[2];
    g->data = data;
}

int orc_map_value(const int n[3], const double h[3], const double origin[3],
                  const double *data, const double pt[3], double *value)
{
    ogrid g;
    grid_of(&g, n, h, origin, (double *)data);
    return map_value(&g, pt, value);
}

int orc_set_map(orc_pmg *m, int which, const int n[3], const double h[3],
                const double origin[3], const double *data)
{
    if (which < 0 || which > 5 || n[0] < 1 || n[1] < 1 || n[2] < 1 || !(h[0] > 0) ||
        !(h[1] > 0) || !(h[2] > 0) || !data)
        return 0;
    size_t npts = (size_t)n[0] * n[1] * n[2];
    ogrid *g = m->map[which] ? m->map[which] : (ogrid *)calloc(1, sizeof(ogrid));
    free(g->data);
    double *copy = (double *)malloc(npts * sizeof(double));
    memcpy(copy, data, npts * sizeof(double));
    grid_of(g, n, h, origin, copy);
    m->map[which] = g;
    m->usemap[which] = 1;
    return 1;
}

/* fillcoChargeMap (vpmg.c:4938-4985) */
static int fill_charge_map(orc_pmg *m)
{
    int nx = m->nx, ny = m->ny, nz = m->nz;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                double pt[3] = {m->xf[i], m->yf[j], m->zf[k]}, q;
                if (!map_value(m->map[ORC_MAP_CHARGE], pt, &q)) return 0;
                m->charge[IJK(i, j, k)] = q * m->zmagic;
            }
    return 1;
}

/* fillcoCoefMap (vpmg.c:4098-4220) */
static int fill_coef_map(orc_pmg *m)
{
    int nx = m->nx, ny = m->ny, nz = m->nz;
    int salt = m->ionstr > O_VPMGSMALL;
    if (!m->usemap[ORC_MAP_DIELX] || !m->usemap[ORC_MAP_DIELY] || !m->usemap[ORC_MAP_DIELZ] ||
        (!m->usemap[ORC_MAP_KAPPA] && salt))
        return 0;
    double kappamax = -1.0, t;
    if (salt) {
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++) {
                    double pt[3] = {m->xf[i], m->yf[j], m->zf[k]};
                    if (!map_value(m->map[ORC_MAP_KAPPA], pt, &t)) return 0;
                    if (t > kappamax) kappamax = t;
                    if (t < 0.0) return 0;
                }
    }
    if (!(kappamax > 1.0)) kappamax = 1.0;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                size_t c = IJK(i, j, k);
                if (salt) {
                    double pt[3] = {m->xf[i], m->yf[j], m->zf[k]};
                    if (!map_value(m->map[ORC_MAP_KAPPA], pt, &t)) return 0;
                    if (t < O_VPMGSMALL) t = 0.0;
                    m->kappa[c] = t / kappamax;
                } else {
                    m->kappa[c] = 0.0;
                }
                double px[3] = {m->xf[i] + 0.5 * m->hx, m->yf[j], m->zf[k]};
                double py[3] = {m->xf[i], m->yf[j] + 0.5 * m->hy, m->zf[k]};
                double pz[3] = {m->xf[i], m->yf[j], m->zf[k] + 0.5 * m->hz};
                if (!map_value(m->map[ORC_MAP_DIELX], px, &m->epsx[c])) return 0;
                if (!map_value(m->map[ORC_MAP_DIELY], py, &m->epsy[c])) return 0;
                if (!map_value(m->map[ORC_MAP_DIELZ], pz, &m->epsz[c])) return 0;
            }
    return 1;
}

/* bcfl_map (vpmg.c:3953-3990): every grid point on the potential map */
static int bcfl_map_check(orc_pmg *m)
{
    double v;
    for (int k = 0; k < m->nz; k++)
        for (int j = 0; j < m->ny; j++)
            for (int i = 0; i < m->nx; i++) {
                double pt[3] = {m->xf[i], m->yf[j], m->zf[k]};
                if (!map_value(m->map[ORC_MAP_POT], pt, &v)) return 0;
            }
    return 1;
}

/* focusFillBound (vpmg.c:1914-2414) */
static int focus_fill_bound(orc_pmg *m, const ogrid *o)
{
    int nx = m->nx, ny = m->ny, nz = m->nz;
    double xminN = m->p.center[0] - ((double)(nx - 1) * m->hx) / 2.0;
    double xmaxN = m->p.center[0] + ((double)(nx - 1) * m->hx) / 2.0;
    double yminN = m->p.center[1] - ((double)(ny - 1) * m->hy) / 2.0;
    double ymaxN = m->p.center[1] + ((double)(ny - 1) * m->hy) / 2.0;
    double zminN = m->p.center[2] - ((double)(nz - 1) * m->hz) / 2.0;
    double zmaxN = m->p.center[2] + ((double)(nz - 1) * m->hz) / 2.0;
    double xminO = o->xmin, xmaxO = o->xmax, yminO = o->ymin, ymaxO = o->ymax;
    double zminO = o->zmin, zmaxO = o->zmax;
    if (fabs(xminO - xminN) < O_VSMALL) xminN = xminO;
    if (fabs(xmaxO - xmaxN) < O_VSMALL) xmaxN = xmaxO;
    if (fabs(yminO - yminN) < O_VSMALL) yminN = yminO;
    if (fabs(ymaxO - ymaxN) < O_VSMALL) ymaxN = ymaxO;
    if (fabs(zminO - zminN) < O_VSMALL) zminN = zminO;
    if (fabs(zmaxO - zmaxN) < O_VSMALL) zmaxN = zmaxO;
    if ((xmaxN > xmaxO) || (ymaxN > ymaxO) || (zmaxN > zmaxO) || (xminO > xminN) ||
        (yminO > yminN) || (zminO > zminN))
        return 0;
    int ok;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++) {
            double y = yminN + j * m->hy, z = zminN + k * m->hz;
            m->gxcf[j + ny * k] = trilin_old(o, xminN, y, z, &ok); if (!ok) return 0;
            m->gxcf[j + ny * k + (size_t)ny * nz] = trilin_old(o, xmaxN, y, z, &ok); if (!ok) return 0;
        }
    for (int k = 0; k < nz; k++)
        for (int i = 0; i < nx; i++) {
            double x = xminN + i * m->hx, z = zminN + k * m->hz;
            m->gycf[i + nx * k] = trilin_old(o, x, yminN, z, &ok); if (!ok) return 0;
            m->gycf[i + nx * k + (size_t)nx * nz] = trilin_old(o, x, ymaxN, z, &ok); if (!ok) return 0;
        }
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            double x = xminN + i * m->hx, y = yminN + j * m->hy;
            m->gzcf[i + nx * j] = trilin_old(o, x, y, zminN, &ok); if (!ok) return 0;
            m->gzcf[i + nx * j + (size_t)nx * ny] = trilin_old(o, x, y, zmaxN, &ok); if (!ok) return 0;
        }
    return 1;
}

/* Vpmg_setPart (vpmg.c:627-870) with all bflags 0 (sequential focusing) */
static double part_axis(double x, double lo, double hi, double h)
{
    double ok;
    if ((x < (hi - h / 2)) && (x > (lo + h / 2))) ok = 1.0;
    else if (fabs(x - lo) < O_VPMGSMALL) ok = 1.0;
    else if (fabs(x - hi) < O_VPMGSMALL) ok = 1.0;
    else if ((x > (hi + h / 2)) || (x < (lo - h / 2))) ok = 0.0;
    else {
        double x0 = OMAX(x - h / 2, lo), x1 = OMIN(x + h / 2, hi);
        ok = fabs(x1 - x0) / h;
        if (ok < 0.0 && fabs(ok) < O_VPMGSMALL) ok = 0.0;
        if (ok > 1.0 && fabs(ok - 1.0) < O_VPMGSMALL) ok = 1.0;
    }
    return ok;
}

static double atom_axis(double x, double lo, double hi)
{
    if ((x < hi) && (x > lo)) return 1;
    if (fabs(x - lo) < O_VPMGSMALL) return 1;
    if (fabs(x - hi) < O_VPMGSMALL) return 1;
    return 0;
}

static void set_part(orc_pmg *m, const double lo[3], const double hi[3])
{
    int nx = m->nx, ny = m->ny, nz = m->nz;
    double xmin = m->p.center[0] - 0.5 * m->hx * (nx - 1);
    double ymin = m->p.center[1] - 0.5 * m->hy * (ny - 1);
    double zmin = m->p.center[2] - 0.5 * m->hz * (nz - 1);
    for (int a = 0; a < m->natoms; a++) {
        const double *ap = m->apos + 3 * a;
        m->partID[a] = atom_axis(ap[0], lo[0], hi[0]) * atom_axis(ap[1], lo[1], hi[1]) *
                       atom_axis(ap[2], lo[2], hi[2]);
    }
    for (int i = 0; i < nx; i++) {
        double xok = part_axis(i * m->hx + xmin, lo[0], hi[0], m->hx);
        for (int j = 0; j < ny; j++) {
            double yok = part_axis(j * m->hy + ymin, lo[1], hi[1], m->hy);
            for (int k = 0; k < nz; k++) {
                double zok = part_axis(k * m->hz + zmin, lo[2], hi[2], m->hz);
                double v = xok * yok * zok;
                m->pvec[IJK(i, j, k)] = (fabs(v) < O_VPMGSMALL) ? 0.0 : v;
            }
        }
    }
}

static void unset_part(orc_pmg *m) /* vpmg.c:872-890 */
{
    for (size_t i = 0; i < m->n; i++) m->pvec[i] = 1;
    for (int a = 0; a < m->natoms; a++) m->partID[a] = 1;
}

void orc_set_part_domain(orc_pmg *m)
{
    double lo[3], hi[3];
    for (int d = 0; d < 3; d++) {
        double glen = m->p.grid[d] * (double)(m->p.dime[d] - 1);
        lo[d] = m->p.center[d] - 0.5 * glen;
        hi[d] = m->p.center[d] + 0.5 * glen;
    }
    set_part(m, lo, hi);
}

/* extEnergy (vpmg.c:2416-2545) with all bflags 0 */
static void ext_energy(orc_pmg *m, orc_pmg *o)
{
    int nxN = m->nx, nyN = m->ny, nzN = m->nz;
    double lo[3], hi[3];
    lo[0] = m->p.center[0] - ((double)(nxN - 1) * m->hx) / 2.0;
    hi[0] = m->p.center[0] + ((double)(nxN - 1) * m->hx) / 2.0;
    lo[1] = m->p.center[1] - ((double)(nyN - 1) * m->hy) / 2.0;
    hi[1] = m->p.center[1] + ((double)(nyN - 1) * m->hy) / 2.0;
    lo[2] = m->p.center[2] - ((double)(nzN - 1) * m->hz) / 2.0;
    hi[2] = m->p.center[2] + ((double)(nzN - 1) * m->hz) / 2.0;
    set_part(o, lo, hi);
    for (size_t i = 0; i < o->n; i++) {
        if (o->pvec[i] > O_VSMALL) o->pvec[i] = 1.0;
        o->pvec[i] = (1 - (o->pvec[i])) * 1.0;
    }
    for (int a = 0; a < o->natoms; a++) {
        if (o->partID[a] > O_VSMALL) o->partID[a] = 1.0;
        o->partID[a] = (1 - o->partID[a]) * 1.0;
    }
    m->extQm = orc_qm_energy(o, 1);
    m->extQf = orc_qf_energy(o, 1);
    m->extDi = orc_diel_energy(o, 1);
    unset_part(o);
}

/* ---------------------------------------------------------------- objects */

orc_pmg *orc_create(const orc_params *p, int natoms, const double *xyz,
                    const double *rad, const double *chg, orc_pmg *old, int energy_flag)
{
    orc_pmg *m = (orc_pmg *)calloc(1, sizeof(orc_pmg));
    m->p = *p;
    m->natoms = natoms;
    m->apos = (double *)malloc((size_t)3 * natoms * sizeof(double));
    m->arad = (double *)malloc((size_t)natoms * sizeof(double));
    m->achg = (double *)malloc((size_t)natoms * sizeof(double));
    m->partID = (double *)malloc((size_t)natoms * sizeof(double));
    memcpy(m->apos, xyz, (size_t)3 * natoms * sizeof(double));
    memcpy(m->arad, rad, (size_t)natoms * sizeof(double));
    memcpy(m->achg, chg, (size_t)natoms * sizeof(double));
    setup_pbe(m);
    m->nx = p->dime[0]; m->ny = p->dime[1]; m->nz = p->dime[2];
    m->nlev = p->nlev;
    m->hx = p->grid[0]; m->hy = p->grid[1]; m->hz = p->grid[2];
    m->xlen = ((double)(m->nx - 1)) * m->hx;
    m->ylen = ((double)(m->ny - 1)) * m->hy;
    m->zlen = ((double)(m->nz - 1)) * m->hz;
    m->n = (size_t)m->nx * m->ny * m->nz;
    size_t n = m->n;
    m->epsx = (double *)calloc(n, sizeof(double));
    m->epsy = (double *)calloc(n, sizeof(double));
    m->epsz = (double *)calloc(n, sizeof(double));
    m->kappa = (double *)calloc(n, sizeof(double));
    m->charge = (double *)calloc(n, sizeof(double));
    m->u = (double *)calloc(n, sizeof(double));
    m->pvec = (double *)calloc(n, sizeof(double));
    m->gxcf = (double *)calloc((size_t)2 * m->ny * m->nz, sizeof(double));
    m->gycf = (double *)calloc((size_t)2 * m->nx * m->nz, sizeof(double));
    m->gzcf = (double *)calloc((size_t)2 * m->nx * m->ny, sizeof(double));
    m->xf = (double *)calloc(m->nx, sizeof(double));
    m->yf = (double *)calloc(m->ny, sizeof(double));
    m->zf = (double *)calloc(m->nz, sizeof(double));
    if (old) {
        m->p.bcfl = ORC_BCFL_FOCUS;
        ogrid om = old_mesh(old);
        if (!focus_fill_bound(m, &om)) { orc_destroy(old); orc_destroy(m); return NULL; }
        if (energy_flag) ext_energy(m, old);
        orc_destroy(old);
    }
    unset_part(m);
    return m;
}

void orc_destroy(orc_pmg *m)
{
    if (!m) return;
    free(m->apos); free(m->arad); free(m->achg); free(m->partID);
    free(m->epsx); free(m->epsy); free(m->epsz); free(m->kappa); free(m->charge);
    free(m->u); free(m->pvec); free(m->gxcf); free(m->gycf); free(m->gzcf);
    free(m->xf); free(m->yf); free(m->zf);
    for (int w = 0; w < 6; w++)
        if (m->map[w]) { free(m->map[w]->data); free(m->map[w]); }
    omg_free(&m->g);
    free(m);
}

int orc_fillco(orc_pmg *m)
{
    m->xmin = m->p.center[0] - (m->xlen / 2.0);
    m->ymin = m->p.center[1] - (m->ylen / 2.0);
    m->zmin = m->p.center[2] - (m->zlen / 2.0);
    m->xmax = m->p.center[0] + (m->xlen / 2.0);
    m->ymax = m->p.center[1] + (m->ylen / 2.0);
    m->zmax = m->p.center[2] + (m->zlen / 2.0);
    /* vpmg.c:5354-5362: not a Laplacian when coefficient or potential maps are used */
    int coef_maps = m->usemap[ORC_MAP_DIELX] || m->usemap[ORC_MAP_DIELY] ||
                    m->usemap[ORC_MAP_DIELZ] || m->usemap[ORC_MAP_KAPPA];
    int islap = !(coef_maps || m->usemap[ORC_MAP_POT]) && (m->ionstr < O_VPMGSMALL) &&
                (fabs(m->p.pdie - m->p.sdie) < O_VPMGSMALL);
    for (int i = 0; i < m->nx; i++) m->xf[i] = m->xmin + i * m->hx;
    for (int i = 0; i < m->ny; i++) m->yf[i] = m->ymin + i * m->hy;
    for (int i = 0; i < m->nz; i++) m->zf[i] = m->zmin + i * m->hz;
    if (m->usemap[ORC_MAP_CHARGE]) {
        if (!fill_charge_map(m)) return 0;
    } else if (m->p.chgm == ORC_CHGM_SPL2) fill_charge_spl2(m);
    else if (m->p.chgm == ORC_CHGM_SPL4) fill_charge_spl4(m);
    else fill_charge_spl0(m);
    if (!islap && coef_maps) { /* fillcoCoef (vpmg.c:4856-4866) */
        if (!fill_coef_map(m)) return 0;
    } else if (!islap) {
        if (m->p.srfm >= ORC_SRFM_SPL2) fill_spline(m, m->p.srfm);
        else {
            fill_mol_diel(m);
            if (m->p.srfm == ORC_SRFM_SMOL) fill_mol_smooth(m);
            fill_mol_ion(m);
        }
    } else {
        for (size_t i = 0; i < m->n; i++) {
            m->kappa[i] = 0.0;
            m->epsx[i] = m->epsy[i] = m->epsz[i] = m->p.pdie;
        }
    }
    if (m->p.bcfl != ORC_BCFL_FOCUS) {
        memset(m->gxcf, 0, (size_t)2 * m->ny * m->nz * sizeof(double));
        memset(m->gycf, 0, (size_t)2 * m->nx * m->nz * sizeof(double));
        memset(m->gzcf, 0, (size_t)2 * m->nx * m->ny * sizeof(double));
        if (m->p.bcfl == ORC_BCFL_MDH) bc_mdh(m);
        else if (m->p.bcfl == ORC_BCFL_SDH) bc_sdh(m);
        else if (m->p.bcfl == ORC_BCFL_MAP) { /* bcCalc (vpmg.c:4084-4087) */
            if (!m->usemap[ORC_MAP_POT] || !bcfl_map_check(m)) return 0;
            if (!focus_fill_bound(m, m->map[ORC_MAP_POT])) return 0;
        }
        else if (m->p.bcfl != ORC_BCFL_ZERO) return 0;
    }
    m->filled = 1;
    return 1;
}

int orc_solve(orc_pmg *m)
{
    if (!m->filled) return 0;
    size_t n = m->n;
    double *ccf = (double *)malloc(n * sizeof(double));
    for (size_t i = 0; i < n; i++) ccf[i] = (m->zkappa2 > O_VPMGSMALL) ? m->zkappa2 * m->kappa[i] : 0.0;
    omg *g = &m->g;
    omg_free(g);
    g->colour8 = m->p.coarse_smoother;
    g->deep = m->p.deep_coarse;
    g->errtol = (m->p.etol > 0) ? m->p.etol : 1.0e-6;
    g->itmax = 200; g->nu1 = 2; g->nu2 = 2;
    g->newton_itmax = m->p.newton_itmax > 0 ? m->p.newton_itmax : g->itmax;
    /* Vpmg_ctor2 scales the bulk concentrations by 1/(2 I) (vpmg.c:340-377) */
    double zks2 = (m->ionstr > 0.0) ? 0.5 / m->ionstr : 0.0;
    g->nion = m->p.nion;
    for (int i = 0; i < m->p.nion; i++) { g->q[i] = m->p.ionq[i]; g->sconc[i] = zks2 * m->p.ionc[i]; }
    if (!omg_build(g, m->nx, m->ny, m->nz, m->nlev, m->p.pbetype, m->xf, m->yf, m->zf, m->gxcf,
                   m->gycf, m->gzcf, m->epsx, m->epsy, m->epsz, ccf, m->charge)) {
        free(ccf);
        return 0;
    }
    free(ccf);
    memset(m->u, 0, n * sizeof(double));
    if (m->p.pbetype == ORC_NPBE) omg_solve_newton(g, m->u);
    else omg_solve_linear(g, m->u);
    orc_fbound_apply(m->nx, m->ny, m->nz, m->u, m->gxcf, m->gycf, m->gzcf);
    m->solved = 1;
    return 1;
}

/* --------------------------------------------------------------- energies */

static double vcap_exp(double x, int *ichop) /* vcap.c:59-73 */
{
    if (x > O_EXPMAX) { *ichop = 1; return exp(O_EXPMAX); }
    if (x < O_EXPMIN) { *ichop = 1; return exp(O_EXPMIN); }
    *ichop = 0;
    return exp(x);
}

double orc_diel_energy(orc_pmg *m, int extFlag)
{
    int nx = m->nx, ny = m->ny, nz = m->nz;
    double e = 0.0;
    for (int k = 0; k < nz - 1; k++)
        for (int j = 0; j < ny - 1; j++)
            for (int i = 0; i < nx - 1; i++) {
                double pvx = 0.5 * (m->pvec[IJK(i, j, k)] + m->pvec[IJK(i + 1, j, k)]);
                double pvy = 0.5 * (m->pvec[IJK(i, j, k)] + m->pvec[IJK(i, j + 1, k)]);
                double pvz = 0.5 * (m->pvec[IJK(i, j, k)] + m->pvec[IJK(i, j, k + 1)]);
                double ex = m->epsx[IJK(i, j, k)] * pvx * OSQR((m->u[IJK(i, j, k)] - m->u[IJK(i + 1, j, k)]) / m->hx);
                double ey = m->epsy[IJK(i, j, k)] * pvy * OSQR((m->u[IJK(i, j, k)] - m->u[IJK(i, j + 1, k)]) / m->hy);
                double ez = m->epsz[IJK(i, j, k)] * pvz * OSQR((m->u[IJK(i, j, k)] - m->u[IJK(i, j, k + 1)]) / m->hz);
                e += (ex + ey + ez);
            }
    e = 0.5 * e * m->hx * m->hy * m->hz;
    e = e / m->zmagic;
    if (extFlag == 1) e += m->extDi;
    return e;
}

double orc_qm_energy(orc_pmg *m, int extFlag)
{
    if (m->zkappa2 < O_VSMALL) return 0.0;
    double zks2 = 0.5 * m->zkappa2 / m->ionstr;
    double e = 0.0;
    int nchop = 0, ichop;
    if (m->p.pbetype == ORC_NPBE) {
        for (size_t i = 0; i < m->n; i++) {
            if (m->pvec[i] * m->kappa[i] > O_VSMALL)
                for (int j = 0; j < m->p.nion; j++) {
                    e += (m->pvec[i] * m->kappa[i] * zks2 * m->p.ionc[j] *
                          (vcap_exp(-m->p.ionq[j] * m->u[i], &ichop) - 1.0));
                    nchop += ichop;
                }
        }
    } else {
        for (size_t i = 0; i < m->n; i++)
            if (m->pvec[i] * m->kappa[i] > O_VSMALL)
                e += (m->pvec[i] * m->zkappa2 * m->kappa[i] * OSQR(m->u[i]));
        e = 0.5 * e;
    }
    e = e * m->hx * m->hy * m->hz;
    e = e / m->zmagic;
    if (extFlag == 1) e += m->extQm;
    return e;
}

double orc_qf_energy(orc_pmg *m, int extFlag)
{
    double e = 0.0;
    /* Vpmg_qfEnergy (vpmg.c:1687-1702) */
    if (m->usemap[ORC_MAP_CHARGE] || m->p.chgm == ORC_CHGM_SPL2) { /* vpmg.c:1861-1891 */
        for (size_t i = 0; i < m->n; i++) e += (m->pvec[i] * m->u[i] * m->charge[i]);
        e = e * m->hx * m->hy * m->hz / m->zmagic;
        if (extFlag == 1) e += m->extQf;
        return e;
    }
    /* Vpmg_qfEnergyPoint (vpmg.c:1704-1789) */
    int nx = m->nx, ny = m->ny, nz = m->nz;
    for (int a = 0; a < m->natoms; a++) {
        const double *ap = m->apos + 3 * a;
        double fi = (ap[0] - m->xmin) / m->hx, fj = (ap[1] - m->ymin) / m->hy, fk = (ap[2] - m->zmin) / m->hz;
        int ihi = (int)ceil(fi), ilo = (int)floor(fi), jhi = (int)ceil(fj), jlo = (int)floor(fj);
        int khi = (int)ceil(fk), klo = (int)floor(fk);
        if (m->partID[a] > 0) {
            if ((ihi < nx) && (jhi < ny) && (khi < nz) && (ilo >= 0) && (jlo >= 0) && (klo >= 0)) {
                double dx = fi - (double)(ilo), dy = fj - (double)(jlo), dz = fk - (double)(klo);
                const double *u = m->u;
                double uval = dx * dy * dz * u[IJK(ihi, jhi, khi)] + dx * (1.0 - dy) * dz * u[IJK(ihi, jlo, khi)] +
                              dx * dy * (1.0 - dz) * u[IJK(ihi, jhi, klo)] +
                              dx * (1.0 - dy) * (1.0 - dz) * u[IJK(ihi, jlo, klo)] +
                              (1.0 - dx) * dy * dz * u[IJK(ilo, jhi, khi)] +
                              (1.0 - dx) * (1.0 - dy) * dz * u[IJK(ilo, jlo, khi)] +
                              (1.0 - dx) * dy * (1.0 - dz) * u[IJK(ilo, jhi, klo)] +
                              (1.0 - dx) * (1.0 - dy) * (1.0 - dz) * u[IJK(ilo, jlo, klo)];
                e += (uval * m->achg[a] * m->partID[a]);
            }
        }
    }
    if (extFlag) e += m->extQf;
    return e;
}

double orc_energy(orc_pmg *m, int extFlag) /* vpmg.c:1248-1277 */
{
    if ((m->p.pbetype == ORC_NPBE) && (m->ionstr > 0.)) {
        double qm = orc_qm_energy(m, extFlag), qf = orc_qf_energy(m, extFlag),
               di = orc_diel_energy(m, extFlag);
        return qf - di - qm;
    }
    return 0.5 * orc_qf_energy(m, extFlag);
}

/* ----------------------------------------------------------------- forces */

/* Vacc_splineAccGradAtomNorm (vacc.c:314-373) */
static void spline_grad(const double *center, double win, double infrad,
                        const double *apos, double aradius, double *grad)
{
    double w2i = 1.0 / (win * win), w3i = 1.0 / (win * win * win);
    grad[0] = grad[1] = grad[2] = 0.0;
    if (aradius > 0.0) {
        double arad = aradius + infrad;
        double dist = sqrt(OSQR(apos[0] - center[0]) + OSQR(apos[1] - center[1]) + OSQR(apos[2] - center[2]));
        if (dist < (arad - win)) return;
        else if (dist > (arad + win)) return;
        else if ((fabs(dist - (arad - win)) < O_VSMALL) || (fabs(dist - (arad + win)) < O_VSMALL)) return;
        double sm = dist - arad + win, sm2 = OSQR(sm);
        double mychi = 0.75 * sm2 * w2i - 0.25 * sm * sm2 * w3i;
        double mygrad = 1.5 * sm * w2i - 0.75 * sm2 * w3i;
        for (int i = 0; i < 3; i++) grad[i] = -(mygrad / mychi) * ((center[i] - apos[i]) / dist);
    }
}

/* Vacc_splineAccGradAtomNorm3 / Norm4 (vacc.c:1097-1190, 1004-1095): the
 * polynomial windows' normalized gradient, coefficients computed per call
 * from (radius + infrad) +- win as the reference does */
static void poly_grad(int order, const double *center, double win, double infrad,
                      const double *apos, double aradius, double *grad)
{
    grad[0] = grad[1] = grad[2] = 0.0;
    if (aradius > 0.0) {
        double arad = aradius + infrad, c[8];
        win_coef(order, arad - win, arad + win, c);
        double dist = sqrt(OSQR(apos[0] - center[0]) + OSQR(apos[1] - center[1]) + OSQR(apos[2] - center[2]));
        if (dist < (arad - win)) return;
        else if (dist > (arad + win)) return;
        else if ((fabs(dist - (arad - win)) < O_VSMALL) || (fabs(dist - (arad + win)) < O_VSMALL)) return;
        double sm = dist, sm2 = sm * sm, sm3 = sm2 * sm, sm4 = sm3 * sm, sm5 = sm4 * sm;
        double mychi, mygrad;
        if (order == 3) {
            mychi = c[0] + c[1] * sm + c[2] * sm2 + c[3] * sm3 + c[4] * sm4 + c[5] * sm5;
            mygrad = c[1] + 2.0 * c[2] * sm + 3.0 * c[3] * sm2 + 4.0 * c[4] * sm3 + 5.0 * c[5] * sm4;
        } else {
            double sm6 = sm5 * sm, sm7 = sm6 * sm;
            mychi = c[0] + c[1] * sm + c[2] * sm2 + c[3] * sm3 + c[4] * sm4 + c[5] * sm5 +
                    c[6] * sm6 + c[7] * sm7;
            mygrad = c[1] + 2.0 * c[2] * sm + 3.0 * c[3] * sm2 + 4.0 * c[4] * sm3 +
                     5.0 * c[5] * sm4 + 6.0 * c[6] * sm5 + 7.0 * c[7] * sm6;
        }
        if (mychi <= 0.0) return;
        else if (mychi > 1.0) mychi = 1.0;
        for (int i = 0; i < 3; i++) grad[i] = -(mygrad / mychi) * ((center[i] - apos[i]) / dist);
    }
}

/* Vpmg_splineSelect (vpmg.c:1890-1910) */
static void win_grad(int srfm, const double *center, double win, double infrad,
                     const double *apos, double aradius, double *grad)
{
    if (srfm == ORC_SRFM_SPL3 || srfm == ORC_SRFM_SPL4)
        poly_grad(srfm, center, win, infrad, apos, aradius, grad);
    else
        spline_grad(center, win, infrad, apos, aradius, grad);
}

static void qf_force_spl4(orc_pmg *m, int a, double *force) /* vpmg.c:6156-6275 */
{
    int nx = m->nx, ny = m->ny, nz = m->nz;
    double hx = m->hx, hy = m->hy, hzed = m->hz;
    const double *ap = m->apos + 3 * a;
    double e[3] = {0.0, 0.0, 0.0};
    force[0] = force[1] = force[2] = 0.0;
    if (m->partID[a] == 0) return;
    double c = m->achg[a];
    if ((ap[0] <= (m->xmin + 2 * hx)) || (ap[0] >= (m->xmax - 2 * hx)) ||
        (ap[1] <= (m->ymin + 2 * hy)) || (ap[1] >= (m->ymax - 2 * hy)) ||
        (ap[2] <= (m->zmin + 2 * hzed)) || (ap[2] >= (m->zmax - 2 * hzed)))
        return;
    double fi = (ap[0] - m->xmin) / hx, fj = (ap[1] - m->ymin) / hy, fk = (ap[2] - m->zmin) / hzed;
    int ip2 = OMIN((int)ceil(fi) + 2, nx - 1), im2 = OMAX((int)floor(fi) - 2, 0);
    int jp2 = OMIN((int)ceil(fj) + 2, ny - 1), jm2 = OMAX((int)floor(fj) - 2, 0);
    int kp2 = OMIN((int)ceil(fk) + 2, nz - 1), km2 = OMAX((int)floor(fk) - 2, 0);
    for (int ii = im2; ii <= ip2; ii++) {
        double mi = VFCHI4(ii, fi), mx = bspline4(mi), dmx = dbspline4(mi);
        for (int jj = jm2; jj <= jp2; jj++) {
            double mj = VFCHI4(jj, fj), my = bspline4(mj), dmy = dbspline4(mj);
            for (int kk = km2; kk <= kp2; kk++) {
                double mk = VFCHI4(kk, fk), mz = bspline4(mk), dmz = dbspline4(mk);
                double f = m->u[IJK(ii, jj, kk)];
                e[0] += f * dmx * my * mz / hx;
                e[1] += f * mx * dmy * mz / hy;
                e[2] += f * mx * my * dmz / hzed;
            }
        }
    }
    force[0] = e[0] * c;
    force[1] = e[1] * c;
    force[2] = e[2] * c;
}

static void qf_force_spl2(orc_pmg *m, int a, double *f) /* vpmg.c:6043-6154 */
{
    int nx = m->nx, ny = m->ny, nz = m->nz;
    double hx = m->hx, hy = m->hy, hzed = m->hz;
    const double *ap = m->apos + 3 * a;
    f[0] = f[1] = f[2] = 0.0;
    if (m->partID[a] == 0) return;
    if ((ap[0] <= (m->xmin + hx)) || (ap[0] >= (m->xmax - hx)) || (ap[1] <= (m->ymin + hy)) ||
        (ap[1] >= (m->ymax - hy)) || (ap[2] <= (m->zmin + hzed)) || (ap[2] >= (m->zmax - hzed)))
        return;
    double charge = m->achg[a];
    double fi = (ap[0] - m->xmin) / hx, fj = (ap[1] - m->ymin) / hy, fk = (ap[2] - m->zmin) / hzed;
    int ip1 = (int)ceil(fi), ip2 = ip1 + 1, im1 = (int)floor(fi), im2 = im1 - 1;
    int jp1 = (int)ceil(fj), jp2 = jp1 + 1, jm1 = (int)floor(fj), jm2 = jm1 - 1;
    int kp1 = (int)ceil(fk), kp2 = kp1 + 1, km1 = (int)floor(fk), km2 = km1 - 1;
    ip2 = OMIN(ip2, nx - 1); im2 = OMAX(im2, 0);
    jp2 = OMIN(jp2, ny - 1); jm2 = OMAX(jm2, 0);
    kp2 = OMIN(kp2, nz - 1); km2 = OMAX(km2, 0);
    (void)im1; (void)jm1; (void)km1;
    for (int ii = im2; ii <= ip2; ii++) {
        double mx = bspline2(VFCHI(ii, fi)), dmx = dbspline2(VFCHI(ii, fi));
        for (int jj = jm2; jj <= jp2; jj++) {
            double my = bspline2(VFCHI(jj, fj)), dmy = dbspline2(VFCHI(jj, fj));
            for (int kk = km2; kk <= kp2; kk++) {
                double mz = bspline2(VFCHI(kk, fk)), dmz = dbspline2(VFCHI(kk, fk));
                double uu = m->u[IJK(ii, jj, kk)];
                f[0] += (charge * dmx * my * mz * uu) / hx;
                f[1] += (charge * mx * dmy * mz * uu) / hy;
                f[2] += (charge * mx * my * dmz * uu) / hzed;
            }
        }
    }
}

static void qf_force_spl0(orc_pmg *m, int a, double *f) /* vpmg.c:5899-6041 */
{
    int nx = m->nx, ny = m->ny;
    const double *ap = m->apos + 3 * a;
    const double *u = m->u;
    f[0] = f[1] = f[2] = 0.0;
    if (m->partID[a] == 0) return;
    if (off_mesh_strict(m, ap)) return;
    double charge = m->achg[a];
    double fi = (ap[0] - m->xmin) / m->hx, fj = (ap[1] - m->ymin) / m->hy, fk = (ap[2] - m->zmin) / m->hz;
    int ihi = (int)ceil(fi), ilo = (int)floor(fi), jhi = (int)ceil(fj), jlo = (int)floor(fj);
    int khi = (int)ceil(fk), klo = (int)floor(fk);
    double dx = fi - (double)(ilo), dy = fj - (double)(jlo), dz = fk - (double)(klo);
    if ((dx > O_VPMGSMALL) && (fabs(1.0 - dx) > O_VPMGSMALL))
        f[0] = -charge * (dy * dz * u[IJK(ihi, jhi, khi)] + dy * (1 - dz) * u[IJK(ihi, jhi, klo)] +
                          (1 - dy) * dz * u[IJK(ihi, jlo, khi)] + (1 - dy) * (1 - dz) * u[IJK(ihi, jlo, klo)] -
                          dy * dz * u[IJK(ilo, jhi, khi)] - dy * (1 - dz) * u[IJK(ilo, jhi, klo)] -
                          (1 - dy) * dz * u[IJK(ilo, jlo, khi)] - (1 - dy) * (1 - dz) * u[IJK(ilo, jlo, klo)]) /
               m->hx;
    if ((dy > O_VPMGSMALL) && (fabs(1.0 - dy) > O_VPMGSMALL))
        f[1] = -charge * (dx * dz * u[IJK(ihi, jhi, khi)] + dx * (1 - dz) * u[IJK(ihi, jhi, klo)] -
                          dx * dz * u[IJK(ihi, jlo, khi)] - dx * (1 - dz) * u[IJK(ihi, jlo, klo)] +
                          (1 - dx) * dz * u[IJK(ilo, jhi, khi)] + (1 - dx) * (1 - dz) * u[IJK(ilo, jhi, klo)] -
                          (1 - dx) * dz * u[IJK(ilo, jlo, khi)] - (1 - dx) * (1 - dz) * u[IJK(ilo, jlo, klo)]) /
               m->hy;
    if ((dz > O_VPMGSMALL) && (fabs(1.0 - dz) > O_VPMGSMALL))
        f[2] = -charge * (dy * dx * u[IJK(ihi, jhi, khi)] - dy * dx * u[IJK(ihi, jhi, klo)] +
                          (1 - dy) * dx * u[IJK(ihi, jlo, khi)] - (1 - dy) * dx * u[IJK(ihi, jlo, klo)] +
                          dy * (1 - dx) * u[IJK(ilo, jhi, khi)] - dy * (1 - dx) * u[IJK(ilo, jhi, klo)] +
                          (1 - dy) * (1 - dx) * u[IJK(ilo, jlo, khi)] - (1 - dy) * (1 - dx) * u[IJK(ilo, jlo, klo)]) /
               m->hz;
}

static int ib_force(orc_pmg *m, int a, double *force) /* vpmg.c:5440-5603 */
{
    int nx = m->nx, ny = m->ny, nz = m->nz;
    double hx = m->hx, hy = m->hy, hzed = m->hz;
    const double *ap = m->apos + 3 * a;
    double arad = m->arad[a];
    force[0] = force[1] = force[2] = 0.0;
    if (m->p.srfm < ORC_SRFM_SPL2) return 0;   /* spline surfaces only (vpmg.c:5471, 5638) */
    if (m->partID[a] == 0) return 1;
    double irad = m->maxIonRadius, zkappa2 = m->zkappa2, izmagic = 1.0 / m->zmagic;
    if (zkappa2 < O_VPMGSMALL) return 1;
    if (off_mesh_strict(m, ap)) return 1;
    double pos[3] = {ap[0] - m->xmin, ap[1] - m->ymin, ap[2] - m->zmin};
    double rtot = (irad + arad + m->p.swin), rtot2 = OSQR(rtot);
    double dx = rtot + 0.5 * hx;
    int imin = OMAX(0, (int)ceil((pos[0] - dx) / hx));
    int imax = OMIN(nx - 1, (int)floor((pos[0] + dx) / hx));
    for (int i = imin; i <= imax; i++) {
        double dx2 = OSQR(pos[0] - hx * i);
        double dy = (rtot2 > dx2) ? sqrt(rtot2 - dx2) + 0.5 * hy : 0.5 * hy;
        int jmin = OMAX(0, (int)ceil((pos[1] - dy) / hy));
        int jmax = OMIN(ny - 1, (int)floor((pos[1] + dy) / hy));
        for (int j = jmin; j <= jmax; j++) {
            double dy2 = OSQR(pos[1] - hy * j);
            double dz = (rtot2 > (dx2 + dy2)) ? sqrt(rtot2 - dx2 - dy2) + 0.5 * hzed : 0.5 * hzed;
            int kmin = OMAX(0, (int)ceil((pos[2] - dz) / hzed));
            int kmax = OMIN(nz - 1, (int)floor((pos[2] + dz) / hzed));
            for (int k = kmin; k <= kmax; k++) {
                double dz2 = OSQR(k * hzed - pos[2]);
                if ((dz2 + dy2 + dx2) <= rtot2) {
                    double gpos[3] = {i * hx + m->xmin, j * hy + m->ymin, k * hzed + m->zmin};
                    double tg[3];
                    win_grad(m->p.srfm, gpos, m->p.swin, irad, ap, arad, tg);
                    double fmag;
                    if (m->p.pbetype == ORC_NPBE) {
                        fmag = 0.0;
                        int ichop;
                        for (int s = 0; s < m->p.nion; s++)
                            fmag += (m->kappa[IJK(i, j, k)]) * m->p.ionc[s] *
                                    (vcap_exp(-m->p.ionq[s] * m->u[IJK(i, j, k)], &ichop) - 1.0) / m->ionstr;
                    } else {
                        fmag = OSQR(m->u[IJK(i, j, k)]) * (m->kappa[IJK(i, j, k)]);
                    }
                    force[0] += (zkappa2 * fmag * tg[0]);
                    force[1] += (zkappa2 * fmag * tg[1]);
                    force[2] += (zkappa2 * fmag * tg[2]);
                }
            }
        }
    }
    for (int d = 0; d < 3; d++) force[d] = force[d] * 0.5 * hx * hy * hzed * izmagic;
    return 1;
}

static int db_force(orc_pmg *m, int a, double *dbF) /* vpmg.c:5605-5860 */
{
    int nx = m->nx, ny = m->ny;
    double hx = m->hx, hy = m->hy, hzed = m->hz;
    const double *ap = m->apos + 3 * a;
    double arad = m->arad[a], srad = m->solventRadius;
    dbF[0] = dbF[1] = dbF[2] = 0.0;
    if (!m->filled) return 0;
    if (m->p.srfm < ORC_SRFM_SPL2) return 0;   /* spline surfaces only (vpmg.c:5471, 5638) */
    if (m->partID[a] == 0) return 1;
    double epsp = m->p.pdie, epsw = m->p.sdie, izmagic = 1.0 / m->zmagic;
    const double *u = m->u;
    if (fabs(epsp - epsw) < O_VPMGSMALL) return 1;
    double deps = (epsw - epsp), depsi = 1.0 / deps;
    double rtot = (arad + m->p.swin + srad);
    if ((ap[0] <= m->xmin + rtot) || (ap[0] >= m->xmax - rtot) || (ap[1] <= m->ymin + rtot) ||
        (ap[1] >= m->ymax - rtot) || (ap[2] <= m->zmin + rtot) || (ap[2] >= m->zmax - rtot))
        return 1;
    double pos[3] = {ap[0] - m->xmin, ap[1] - m->ymin, ap[2] - m->zmin};
    int imin = (int)floor((pos[0] - rtot) / hx); if (imin < 1) return 0;
    int imax = (int)ceil((pos[0] + rtot) / hx); if (imax > (nx - 2)) return 0;
    int jmin = (int)floor((pos[1] - rtot) / hy); if (jmin < 1) return 0;
    int jmax = (int)ceil((pos[1] + rtot) / hy); if (jmax > (ny - 2)) return 0;
    int kmin = (int)floor((pos[2] - rtot) / hzed); if (kmin < 1) return 0;
    int kmax = (int)ceil((pos[2] + rtot) / hzed); if (kmax > (m->nz - 2)) return 0;
    double win = m->p.swin;
    for (int i = imin; i <= imax; i++)
        for (int j = jmin; j <= jmax; j++)
            for (int k = kmin; k <= kmax; k++) {
                double g[3], dHx[3], dHy[3], dHz[3], dHxm[3], dHym[3], dHzm[3];
                g[0] = (i + 0.5) * hx + m->xmin; g[1] = j * hy + m->ymin; g[2] = k * hzed + m->zmin;
                double Hx = (m->epsx[IJK(i, j, k)] - epsp) * depsi;
                win_grad(m->p.srfm, g, win, 0., ap, arad, dHx);
                for (int l = 0; l < 3; l++) dHx[l] *= Hx;
                g[0] = i * hx + m->xmin; g[1] = (j + 0.5) * hy + m->ymin; g[2] = k * hzed + m->zmin;
                double Hy = (m->epsy[IJK(i, j, k)] - epsp) * depsi;
                win_grad(m->p.srfm, g, win, 0., ap, arad, dHy);
                for (int l = 0; l < 3; l++) dHy[l] *= Hy;
                g[0] = i * hx + m->xmin; g[1] = j * hy + m->ymin; g[2] = (k + 0.5) * hzed + m->zmin;
                double Hz = (m->epsz[IJK(i, j, k)] - epsp) * depsi;
                win_grad(m->p.srfm, g, win, 0., ap, arad, dHz);
                for (int l = 0; l < 3; l++) dHz[l] *= Hz;
                g[0] = (i - 0.5) * hx + m->xmin; g[1] = j * hy + m->ymin; g[2] = k * hzed + m->zmin;
                double Hxm = (m->epsx[IJK(i - 1, j, k)] - epsp) * depsi;
                win_grad(m->p.srfm, g, win, 0., ap, arad, dHxm);
                for (int l = 0; l < 3; l++) dHxm[l] *= Hxm;
                g[0] = i * hx + m->xmin; g[1] = (j - 0.5) * hy + m->ymin; g[2] = k * hzed + m->zmin;
                double Hym = (m->epsy[IJK(i, j - 1, k)] - epsp) * depsi;
                win_grad(m->p.srfm, g, win, 0., ap, arad, dHym);
                for (int l = 0; l < 3; l++) dHym[l] *= Hym;
                g[0] = i * hx + m->xmin; g[1] = j * hy + m->ymin; g[2] = (k - 0.5) * hzed + m->zmin;
                double Hzm = (m->epsz[IJK(i, j, k - 1)] - epsp) * depsi;
                win_grad(m->p.srfm, g, win, 0., ap, arad, dHzm);
                for (int l = 0; l < 3; l++) dHzm[l] *= Hzm;
                double fmag = u[IJK(i, j, k)];
                double tg[3];
                for (int l = 0; l < 3; l++)
                    tg[l] = (dHx[l] * (u[IJK(i + 1, j, k)] - u[IJK(i, j, k)]) +
                             dHxm[l] * (u[IJK(i - 1, j, k)] - u[IJK(i, j, k)])) / OSQR(hx) +
                            (dHy[l] * (u[IJK(i, j + 1, k)] - u[IJK(i, j, k)]) +
                             dHym[l] * (u[IJK(i, j - 1, k)] - u[IJK(i, j, k)])) / OSQR(hy) +
                            (dHz[l] * (u[IJK(i, j, k + 1)] - u[IJK(i, j, k)]) +
                             dHzm[l] * (u[IJK(i, j, k - 1)] - u[IJK(i, j, k)])) / OSQR(hzed);
                dbF[0] += (fmag * tg[0]);
                dbF[1] += (fmag * tg[1]);
                dbF[2] += (fmag * tg[2]);
            }
    for (int d = 0; d < 3; d++) dbF[d] = -dbF[d] * hx * hy * hzed * deps * 0.5 * izmagic;
    return 1;
}

int orc_force(orc_pmg *m, int atom, double qf[3], double ib[3], double db[3])
{
    if (m->p.chgm == ORC_CHGM_SPL2) qf_force_spl2(m, atom, qf);
    else if (m->p.chgm == ORC_CHGM_SPL4) qf_force_spl4(m, atom, qf);
    else qf_force_spl0(m, atom, qf);
    int r1 = ib_force(m, atom, ib);
    int r2 = db_force(m, atom, db);
    return r1 && r2;
}

/* --------------------------------------------------------------- accessors */

int orc_get_array(orc_pmg *m, int which, double *out)
{
    double *src[] = {m->u, m->epsx, m->epsy, m->epsz, m->kappa, m->charge, m->pvec};
    if (which < 0 || which > 6) return 0;
    memcpy(out, src[which], m->n * sizeof(double));
    return 1;
}

int orc_get_boundary(orc_pmg *m, double *gx, double *gy, double *gz)
{
    memcpy(gx, m->gxcf, (size_t)2 * m->ny * m->nz * sizeof(double));
    memcpy(gy, m->gycf, (size_t)2 * m->nx * m->nz * sizeof(double));
    memcpy(gz, m->gzcf, (size_t)2 * m->nx * m->ny * sizeof(double));
    return 1;
}

int orc_stats(orc_pmg *m, double *out)
{
    out[0] = m->g.iters;
    out[1] = m->g.inner_total;
    out[2] = m->g.ntot;
    out[3] = m->g.relres;
    out[4] = m->zkappa2;
    out[5] = m->zmagic;
    out[6] = m->xkappa;
    out[7] = m->ionstr;
    return 1;
}
