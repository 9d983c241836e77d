/*
 * orc_mg.c -- oracle restatement of the PMG multigrid core (src/pmgc/ *.c).
 * Test infrastructure only; see orc.h.  Single-threaded, built with
 * -ffp-contract=off so products and sums round exactly as written.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc_int.h"

const int o_fwd[14][3] = {
    {0, 0, 0},                                   /* oC  */
    {1, 0, 0},   {0, 1, 0},   {0, 0, 1},         /* oE oN uC */
    {1, 1, 0},   {-1, 1, 0},                     /* oNE oNW */
    {1, 0, 1},   {-1, 0, 1},  {0, 1, 1}, {0, -1, 1},  /* uE uW uN uS */
    {1, 1, 1},   {-1, 1, 1},  {1, -1, 1}, {-1, -1, 1} /* uNE uNW uSE uSW */
};

#define IX(i, j, k) OIDX(i, j, k, nx, ny)

/* ---------------------------------------------------------------- vector kit
 * mikpckd.c:57-289: interior-only BLAS-1 on (nx,ny,nz) grids. */
void o_zero(double *x, size_t n) { memset(x, 0, n * sizeof(double)); }

void o_fbound00(int nx, int ny, int nz, double *x) /* mikpckd.c:258 */
{
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++) { x[IX(0, j, k)] = 0.0; x[IX(nx - 1, j, k)] = 0.0; }
    for (int k = 0; k < nz; k++)
        for (int i = 0; i < nx; i++) { x[IX(i, 0, k)] = 0.0; x[IX(i, ny - 1, k)] = 0.0; }
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) { x[IX(i, j, 0)] = 0.0; x[IX(i, j, nz - 1)] = 0.0; }
}

/* VfboundPMG (mikpckd.c:209): g arrays laid out (j,k,face) / (i,k,face) /
 * (i,j,face), faces 0 = low, 1 = high. */
static void o_fbound(int nx, int ny, int nz, double *x, const double *gx,
                     const double *gy, const double *gz)
{
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++) {
            x[IX(0, j, k)] = gx[j + ny * k];
            x[IX(nx - 1, j, k)] = gx[j + ny * k + (size_t)ny * nz];
        }
    for (int k = 0; k < nz; k++)
        for (int i = 0; i < nx; i++) {
            x[IX(i, 0, k)] = gy[i + nx * k];
            x[IX(i, ny - 1, k)] = gy[i + nx * k + (size_t)nx * nz];
        }
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            x[IX(i, j, 0)] = gz[i + nx * j];
            x[IX(i, j, nz - 1)] = gz[i + nx * j + (size_t)nx * ny];
        }
}

double o_nrm1(int nx, int ny, int nz, const double *x) /* mikpckd.c:131 */
{
    double s = 0.0;
    for (int k = 1; k < nz - 1; k++)
        for (int j = 1; j < ny - 1; j++)
            for (int i = 1; i < nx - 1; i++) s += fabs(x[IX(i, j, k)]);
    return s;
}

double o_nrm2(int nx, int ny, int nz, const double *x) /* mikpckd.c:152 */
{
    double s = 0.0;
    for (int k = 1; k < nz - 1; k++)
        for (int j = 1; j < ny - 1; j++)
            for (int i = 1; i < nx - 1; i++) s += x[IX(i, j, k)] * x[IX(i, j, k)];
    return sqrt(s);
}

double o_dot(int nx, int ny, int nz, const double *x, const double *y) /* :173 */
{
    double s = 0.0;
    for (int k = 1; k < nz - 1; k++)
        for (int j = 1; j < ny - 1; j++)
            for (int i = 1; i < nx - 1; i++) s += x[IX(i, j, k)] * y[IX(i, j, k)];
    return s;
}

static void o_axpy(int nx, int ny, int nz, double a, const double *x, double *y)
{ /* mikpckd.c:112 */
    for (int k = 1; k < nz - 1; k++)
        for (int j = 1; j < ny - 1; j++)
            for (int i = 1; i < nx - 1; i++) y[IX(i, j, k)] += a * x[IX(i, j, k)];
}

static void o_copy(int nx, int ny, int nz, const double *x, double *y)
{ /* mikpckd.c:57 */
    for (int k = 1; k < nz - 1; k++)
        for (int j = 1; j < ny - 1; j++)
            for (int i = 1; i < nx - 1; i++) y[IX(i, j, k)] = x[IX(i, j, k)];
}

static void o_scal(int nx, int ny, int nz, double a, double *x)
{
    for (int k = 1; k < nz - 1; k++)
        for (int j = 1; j < ny - 1; j++)
            for (int i = 1; i < nx - 1; i++) x[IX(i, j, k)] *= a;
}

/* ------------------------------------------------------------ fine operator
 * VbuildA_fv (buildAd.c:105-281).  Coefficients from mesh differences, the
 * diagonal summed in the reference order, Dirichlet faces folded into fc in
 * the order e, n, u, w, s, d. */
void orc_buildA_fv(int nx, int ny, int nz, const double *xf, const double *yf,
                   const double *zf, const double *gxcf, const double *gycf,
                   const double *gzcf, const double *a1cf, const double *a2cf,
                   const double *a3cf, const double *ccf, const double *fcf,
                   double *oC, double *cc, double *fc, double *oE, double *oN,
                   double *uC)
{
    const double diag = 1.0;
    for (int k = 1; k < nz - 1; k++) {
        double hzm1 = zf[k] - zf[k - 1], hz = zf[k + 1] - zf[k];
        for (int j = 1; j < ny - 1; j++) {
            double hym1 = yf[j] - yf[j - 1], hy = yf[j + 1] - yf[j];
            for (int i = 1; i < nx - 1; i++) {
                double hxm1 = xf[i] - xf[i - 1], hx = xf[i + 1] - xf[i];
                double c_oE = diag * (hym1 + hy) * (hzm1 + hz) / (4.0 * hx);
                double c_oEm1 = diag * (hym1 + hy) * (hzm1 + hz) / (4.0 * hxm1);
                double c_oN = diag * (hxm1 + hx) * (hzm1 + hz) / (4.0 * hy);
                double c_oNm1 = diag * (hxm1 + hx) * (hzm1 + hz) / (4.0 * hym1);
                double c_uC = diag * (hxm1 + hx) * (hym1 + hy) / (4.0 * hz);
                double c_uCm1 = diag * (hxm1 + hx) * (hym1 + hy) / (4.0 * hzm1);
                double c_fc = diag * (hxm1 + hx) * (hym1 + hy) * (hzm1 + hz) / 8.0;
                size_t p = IX(i, j, k);
                double f = c_fc * fcf[p];
                cc[p] = c_fc * ccf[p];
                oC[p] = c_oE * a1cf[p] + c_oEm1 * a1cf[IX(i - 1, j, k)] +
                        c_oN * a2cf[p] + c_oNm1 * a2cf[IX(i, j - 1, k)] +
                        c_uC * a3cf[p] + c_uCm1 * a3cf[IX(i, j, k - 1)];
                /* a face next to the Dirichlet boundary carries no link and
                 * adds coef*eps*g to the rhs instead */
                if (i != nx - 2) oE[p] = c_oE * a1cf[p];
                else { oE[p] = 0.0; f += c_oE * a1cf[p] * gxcf[j + ny * k + (size_t)ny * nz]; }
                if (j != ny - 2) oN[p] = c_oN * a2cf[p];
                else { oN[p] = 0.0; f += c_oN * a2cf[p] * gycf[i + nx * k + (size_t)nx * nz]; }
                if (k != nz - 2) uC[p] = c_uC * a3cf[p];
                else { uC[p] = 0.0; f += c_uC * a3cf[p] * gzcf[i + nx * j + (size_t)nx * ny]; }
                if (i == 1) f += c_oEm1 * a1cf[IX(i - 1, j, k)] * gxcf[j + ny * k];
                if (j == 1) f += c_oNm1 * a2cf[IX(i, j - 1, k)] * gycf[i + nx * k];
                if (k == 1) f += c_uCm1 * a3cf[IX(i, j, k - 1)] * gzcf[i + nx * j];
                fc[p] = f;
            }
        }
    }
}

/* ------------------------------------------------------------- stencil apply
 * Coefficient s at point q couples q with q + o_fwd[s].  For the 27-point
 * operator the three partial sums follow matvecd.c:518-596 (tmpO over the
 * own plane, tmpU over k+1, tmpD over k-1). */
#define A_(s) (a + (size_t)(s) * n)

/* 7-point: init + six couplings, accumulated left to right in the order of
 * matvecd.c:461-496 (north, south, east, west, down, up). */
static inline double acc7(double init, const double *a, size_t n, int nx, int ny,
                          int i, int j, int k, const double *x)
{
    const double *oE = A_(A_OE), *oN = A_(A_ON), *uC = A_(A_UC);
    double s = init;
    s += oN[IX(i, j, k)] * x[IX(i, j + 1, k)];
    s += oN[IX(i, j - 1, k)] * x[IX(i, j - 1, k)];
    s += oE[IX(i, j, k)] * x[IX(i + 1, j, k)];
    s += oE[IX(i - 1, j, k)] * x[IX(i - 1, j, k)];
    s += uC[IX(i, j, k - 1)] * x[IX(i, j, k - 1)];
    s += uC[IX(i, j, k)] * x[IX(i, j, k + 1)];
    return s;
}

/* 27-point partial sums tmpO / tmpU / tmpD of matvecd.c:518-596 */
static inline void part27(const double *a, size_t n, int nx, int ny, int i, int j,
                          int k, const double *x, double *tO, double *tU, double *tD)
{
    const double *oE = A_(A_OE), *oN = A_(A_ON), *uC = A_(A_UC);
    const double *oNE = A_(A_ONE), *oNW = A_(A_ONW);
    const double *uE = A_(A_UE), *uW = A_(A_UW), *uN = A_(A_UN), *uS = A_(A_US);
    const double *uNE = A_(A_UNE), *uNW = A_(A_UNW), *uSE = A_(A_USE), *uSW = A_(A_USW);
    double o = oN[IX(i, j, k)] * x[IX(i, j + 1, k)];
    o += oN[IX(i, j - 1, k)] * x[IX(i, j - 1, k)];
    o += oE[IX(i, j, k)] * x[IX(i + 1, j, k)];
    o += oE[IX(i - 1, j, k)] * x[IX(i - 1, j, k)];
    o += oNE[IX(i, j, k)] * x[IX(i + 1, j + 1, k)];
    o += oNW[IX(i, j, k)] * x[IX(i - 1, j + 1, k)];
    o += oNW[IX(i + 1, j - 1, k)] * x[IX(i + 1, j - 1, k)];
    o += oNE[IX(i - 1, j - 1, k)] * x[IX(i - 1, j - 1, k)];
    double u = uC[IX(i, j, k)] * x[IX(i, j, k + 1)];
    u += uN[IX(i, j, k)] * x[IX(i, j + 1, k + 1)];
    u += uS[IX(i, j, k)] * x[IX(i, j - 1, k + 1)];
    u += uE[IX(i, j, k)] * x[IX(i + 1, j, k + 1)];
    u += uW[IX(i, j, k)] * x[IX(i - 1, j, k + 1)];
    u += uNE[IX(i, j, k)] * x[IX(i + 1, j + 1, k + 1)];
    u += uNW[IX(i, j, k)] * x[IX(i - 1, j + 1, k + 1)];
    u += uSE[IX(i, j, k)] * x[IX(i + 1, j - 1, k + 1)];
    u += uSW[IX(i, j, k)] * x[IX(i - 1, j - 1, k + 1)];
    double d = uC[IX(i, j, k - 1)] * x[IX(i, j, k - 1)];
    d += uS[IX(i, j + 1, k - 1)] * x[IX(i, j + 1, k - 1)];
    d += uN[IX(i, j - 1, k - 1)] * x[IX(i, j - 1, k - 1)];
    d += uW[IX(i + 1, j, k - 1)] * x[IX(i + 1, j, k - 1)];
    d += uE[IX(i - 1, j, k - 1)] * x[IX(i - 1, j, k - 1)];
    d += uSW[IX(i + 1, j + 1, k - 1)] * x[IX(i + 1, j + 1, k - 1)];
    d += uSE[IX(i - 1, j + 1, k - 1)] * x[IX(i - 1, j + 1, k - 1)];
    d += uNW[IX(i + 1, j - 1, k - 1)] * x[IX(i + 1, j - 1, k - 1)];
    d += uNE[IX(i - 1, j - 1, k - 1)] * x[IX(i - 1, j - 1, k - 1)];
    *tO = o; *tU = u; *tD = d;
}

/* residual r = f + sum a x_nbr - (oC+cc) x  (Vmresid7_1s :461, Vmresid27_1s :518) */
void orc_resid(int nx, int ny, int nz, int st, const double *a, const double *cc,
               const double *fc, const double *x, double *r)
{
    size_t n = (size_t)nx * ny * nz;
    const double *oC = A_(A_OC);
    for (int k = 1; k < nz - 1; k++)
        for (int j = 1; j < ny - 1; j++)
            for (int i = 1; i < nx - 1; i++) {
                size_t p = IX(i, j, k);
                if (st == 7) {
                    r[p] = acc7(fc[p], a, n, nx, ny, i, j, k, x) - (oC[p] + cc[p]) * x[p];
                } else {
                    double tO, tU, tD;
                    part27(a, n, nx, ny, i, j, k, x, &tO, &tU, &tD);
                    r[p] = fc[p] + tO + tU + tD - (oC[p] + cc[p]) * x[p];
                }
            }
}

/* y = (oC+cc) x - sum a x_nbr  (Vmatvec7_1s :98, Vmatvec27_1s :153).  The
 * reference's chain -t1 - t2 - ... equals -(t1 + t2 + ...) exactly. */
void orc_matvec(int nx, int ny, int nz, int st, const double *a, const double *cc,
                const double *x, double *y)
{
    size_t n = (size_t)nx * ny * nz;
    const double *oC = A_(A_OC);
    for (int k = 1; k < nz - 1; k++)
        for (int j = 1; j < ny - 1; j++)
            for (int i = 1; i < nx - 1; i++) {
                size_t p = IX(i, j, k);
                if (st == 7) {
                    const double *oE = A_(A_OE), *oN = A_(A_ON), *uC = A_(A_UC);
                    double s = oN[IX(i, j, k)] * x[IX(i, j + 1, k)];
                    s += oN[IX(i, j - 1, k)] * x[IX(i, j - 1, k)];
                    s += oE[IX(i, j, k)] * x[IX(i + 1, j, k)];
                    s += oE[IX(i - 1, j, k)] * x[IX(i - 1, j, k)];
                    s += uC[IX(i, j, k - 1)] * x[IX(i, j, k - 1)];
                    s += uC[IX(i, j, k)] * x[IX(i, j, k + 1)];
                    y[p] = -s + (oC[p] + cc[p]) * x[p];
                } else {
                    double tO, tU, tD;
                    part27(a, n, nx, ny, i, j, k, x, &tO, &tU, &tD);
                    y[p] = (-tO) + (-tU) + (-tD) + (oC[p] + cc[p]) * x[p];
                }
            }
}

/* Nonlinear term, Vc_vecpmg (mypdec.c:145-223): sum over species of
 * (-c_i q_i) coef exp(clamp(-q_i u)), with the am_zero switch-off. */
void orc_cvec(int n, int nion, const double *q, const double *sconc,
              const double *coef, const double *u, double *out)
{
    for (int p = 0; p < n; p++) out[p] = 0;
    for (int s = 0; s < nion; s++) {
        double zcf2 = -1.0 * sconc[s] * q[s];
        double zu2 = -1.0 * q[s];
        for (int p = 0; p < n; p++) {
            double am_zero = OMIN(O_ZSMALL, fabs(zcf2 * coef[p])) * O_ZLARGE;
            double am_neg = OMAX(OMIN(zu2 * u[p], 0.0), O_SINH_MIN);
            double am_pos = OMIN(OMAX(zu2 * u[p], 0.0), O_SINH_MAX);
            double arg = am_zero * (am_neg + am_pos);
            out[p] = out[p] + zcf2 * coef[p] * exp(arg);
        }
    }
}

/* Derivative, Vdc_vecpmg (mypdec.c:366-433): sum c_i q_i^2 coef exp(...) */
void orc_dcvec(int n, int nion, const double *q, const double *sconc,
               const double *coef, const double *u, double *out)
{
    for (int p = 0; p < n; p++) out[p] = 0.0;
    for (int s = 0; s < nion; s++) {
        double zcf2 = sconc[s] * q[s] * q[s];
        double zu2 = -1.0 * q[s];
        for (int p = 0; p < n; p++) {
            double am_zero = OMIN(O_ZSMALL, fabs(zcf2 * coef[p])) * O_ZLARGE;
            double am_neg = OMAX(OMIN(zu2 * u[p], 0.0), O_SINH_MIN);
            double am_pos = OMIN(OMAX(zu2 * u[p], 0.0), O_SINH_MAX);
            double arg = am_zero * (am_neg + am_pos);
            out[p] += zcf2 * coef[p] * exp(arg);
        }
    }
}

/* Nonlinear residual r = f + sum a x_nbr - oC x - N(x)
 * (Vnmresid7_1s matvecd.c:633-674, Vnmresid27_1s :696-800); w is scratch. */
static void o_nresid(omg *g, int lev, const double *cc, const double *fc,
                     const double *x, double *r, double *w)
{
    olev *L = &g->L[lev];
    int nx = L->nx, ny = L->ny, nz = L->nz;
    size_t n = L->n;
    const double *a = L->a, *oC = A_(A_OC);
    orc_cvec((int)n, g->nion, g->q, g->sconc, cc, x, w);
    for (int k = 1; k < nz - 1; k++)
        for (int j = 1; j < ny - 1; j++)
            for (int i = 1; i < nx - 1; i++) {
                size_t p = IX(i, j, k);
                if (L->st == 7) {
                    r[p] = acc7(fc[p], a, n, nx, ny, i, j, k, x) - oC[p] * x[p] - w[p];
                } else {
                    double tO, tU, tD;
                    part27(a, n, nx, ny, i, j, k, x, &tO, &tU, &tD);
                    r[p] = tO + tU + tD + fc[p] - oC[p] * x[p] - w[p];
                }
            }
}

/* ------------------------------------------------------------- smoothers
 * Vgsrb7x (gsd.c:95-165): red points i+j+k even (1-based) first unless
 * iadjoint.  Vgsrb27x (gsd.c:167-340): same colouring, lexicographic order,
 * in place.  colour8: the 8-colour ordering by 0-based parity (i&1, j&1, k&1)
 * used by the GPU schedule, colours ascending for iadjoint 0 and descending
 * for 1. */
static void gs_point(int nx, int ny, int st, const double *a, size_t n,
                     const double *cc, const double *fc, double *x, int i, int j, int k)
{
    size_t p = IX(i, j, k);
    const double *oC = A_(A_OC);
    if (st == 7) {
        x[p] = acc7(fc[p], a, n, nx, ny, i, j, k, x) / (oC[p] + cc[p]);
    } else {
        double tO, tU, tD;
        part27(a, n, nx, ny, i, j, k, x, &tO, &tU, &tD);
        x[p] = (fc[p] + (tO + tU + tD)) / (oC[p] + cc[p]);
    }
}

void orc_gsrb(int nx, int ny, int nz, int st, const double *a, const double *cc,
              const double *fc, double *x, int itmax, int iadjoint, int colour8)
{
    size_t n = (size_t)nx * ny * nz;
    for (int it = 0; it < itmax; it++) {
        if (colour8 && st == 27) {
            for (int cc8 = 0; cc8 < 8; cc8++) {
                int c = iadjoint ? 7 - cc8 : cc8;
                /* 1-based parities -> 0-based: 1-based i even <=> 0-based odd */
                /* colour c holds the points with 0-based parities
                 * (i&1, j&1, k&1) = (c&1, (c>>1)&1, (c>>2)&1) */
                int ci = c & 1, cj = (c >> 1) & 1, ck = (c >> 2) & 1;
                for (int k = ck ? 1 : 2; k < nz - 1; k += 2)
                    for (int j = cj ? 1 : 2; j < ny - 1; j += 2)
                        for (int i = ci ? 1 : 2; i < nx - 1; i += 2)
                            gs_point(nx, ny, st, a, n, cc, fc, x, i, j, k);
            }
            continue;
        }
        for (int colour = 0; colour < 2; colour++) {
            /* red sweep of the reference: 1-based i starts at 2 + (j+k)%2 */
            int black = colour ^ iadjoint;
            for (int k = 1; k < nz - 1; k++)
                for (int j = 1; j < ny - 1; j++) {
                    int J = j + 1, K = k + 1;   /* 1-based */
                    int ioff = black ? (1 - (J + K + 2) % 2) : ((J + K + 2) % 2);
                    for (int I = 2 + ioff; I <= nx - 1; I += 2)
                        gs_point(nx, ny, st, a, n, cc, fc, x, I - 1, j, k);
                }
        }
    }
}

/* --------------------------------------------------------------- transfers
 * Trilinear prolongation weights (buildPd.c:114-204): 1, 1/2, 1/4, 1/8 for a
 * fine point 0, 1, 2, 3 axis steps from the coarse point.  Columns belonging
 * to boundary coarse points are zero (they are never assigned). */
static const double W1 = 1.0, W2 = 1.0 / 2.0, W4 = 1.0 / 4.0, W8 = 1.0 / 8.0;

/* Vrestrc2 (matvecd.c:801-913): P^T without scaling; input boundary zeroed */
void orc_restrict(int nxf, int nyf, int nzf, int nxc, int nyc, int nzc,
                  double *xin, double *xout)
{
    o_fbound00(nxf, nyf, nzf, xin);
    for (int k = 1; k < nzc - 1; k++) {
        int kk = 2 * k;
        for (int j = 1; j < nyc - 1; j++) {
            int jj = 2 * j;
            for (int i = 1; i < nxc - 1; i++) {
                int ii = 2 * i;
#define XI(a, b, c) xin[OIDX(a, b, c, nxf, nyf)]
                double tO = W1 * XI(ii, jj, kk) + W2 * XI(ii, jj + 1, kk) + W2 * XI(ii, jj - 1, kk) +
                            W2 * XI(ii + 1, jj, kk) + W2 * XI(ii - 1, jj, kk) +
                            W4 * XI(ii + 1, jj + 1, kk) + W4 * XI(ii - 1, jj + 1, kk) +
                            W4 * XI(ii + 1, jj - 1, kk) + W4 * XI(ii - 1, jj - 1, kk);
                double tU = W2 * XI(ii, jj, kk + 1) + W4 * XI(ii, jj + 1, kk + 1) +
                            W4 * XI(ii, jj - 1, kk + 1) + W4 * XI(ii + 1, jj, kk + 1) +
                            W4 * XI(ii - 1, jj, kk + 1) + W8 * XI(ii + 1, jj + 1, kk + 1) +
                            W8 * XI(ii - 1, jj + 1, kk + 1) + W8 * XI(ii + 1, jj - 1, kk + 1) +
                            W8 * XI(ii - 1, jj - 1, kk + 1);
                double tD = W2 * XI(ii, jj, kk - 1) + W4 * XI(ii, jj + 1, kk - 1) +
                            W4 * XI(ii, jj - 1, kk - 1) + W4 * XI(ii + 1, jj, kk - 1) +
                            W4 * XI(ii - 1, jj, kk - 1) + W8 * XI(ii + 1, jj + 1, kk - 1) +
                            W8 * XI(ii - 1, jj + 1, kk - 1) + W8 * XI(ii + 1, jj - 1, kk - 1) +
                            W8 * XI(ii - 1, jj - 1, kk - 1);
#undef XI
                xout[OIDX(i, j, k, nxc, nyc)] = tO + tU + tD;
            }
        }
    }
    o_fbound00(nxc, nyc, nzc, xout);
}

/* VinterpPMG2 (matvecd.c:935-1076).  wc(i) is the weight factor of coarse
 * point i along one axis: 1 inside, 0 on the boundary (pc never assigned). */
void orc_interp(int nxc, int nyc, int nzc, int nxf, int nyf, int nzf,
                double *xin, double *xout)
{
    o_fbound00(nxc, nyc, nzc, xin);
#define XC(a, b, c) xin[OIDX(a, b, c, nxc, nyc)]
#define XF(a, b, c) xout[OIDX(a, b, c, nxf, nyf)]
#define IN(a, b, c) (((a) > 0 && (a) < nxc - 1 && (b) > 0 && (b) < nyc - 1 && (c) > 0 && (c) < nzc - 1) ? 1.0 : 0.0)
    for (int k = 0; k <= nzf - 3; k += 2) {
        int kk = k / 2;
        for (int j = 0; j <= nyf - 3; j += 2) {
            int jj = j / 2;
            for (int i = 0; i <= nxf - 3; i += 2) {
                int ii = i / 2;
                double m000 = IN(ii, jj, kk), m100 = IN(ii + 1, jj, kk), m010 = IN(ii, jj + 1, kk),
                       m001 = IN(ii, jj, kk + 1), m110 = IN(ii + 1, jj + 1, kk),
                       m101 = IN(ii + 1, jj, kk + 1), m011 = IN(ii, jj + 1, kk + 1),
                       m111 = IN(ii + 1, jj + 1, kk + 1);
                XF(i, j, k) = XC(ii, jj, kk);
                XF(i + 1, j, k) = W2 * m000 * XC(ii, jj, kk) + W2 * m100 * XC(ii + 1, jj, kk);
                XF(i, j + 1, k) = W2 * m000 * XC(ii, jj, kk) + W2 * m010 * XC(ii, jj + 1, kk);
                XF(i, j, k + 1) = W2 * m000 * XC(ii, jj, kk) + W2 * m001 * XC(ii, jj, kk + 1);
                XF(i + 1, j + 1, k) = W4 * m000 * XC(ii, jj, kk) + W4 * m100 * XC(ii + 1, jj, kk) +
                                      W4 * m010 * XC(ii, jj + 1, kk) + W4 * m110 * XC(ii + 1, jj + 1, kk);
                XF(i + 1, j, k + 1) = W4 * m000 * XC(ii, jj, kk) + W4 * m100 * XC(ii + 1, jj, kk) +
                                      W4 * m001 * XC(ii, jj, kk + 1) + W4 * m101 * XC(ii + 1, jj, kk + 1);
                XF(i, j + 1, k + 1) = W4 * m000 * XC(ii, jj, kk) + W4 * m010 * XC(ii, jj + 1, kk) +
                                      W4 * m001 * XC(ii, jj, kk + 1) + W4 * m011 * XC(ii, jj + 1, kk + 1);
                XF(i + 1, j + 1, k + 1) =
                    W8 * m000 * XC(ii, jj, kk) + W8 * m100 * XC(ii + 1, jj, kk) +
                    W8 * m010 * XC(ii, jj + 1, kk) + W8 * m110 * XC(ii + 1, jj + 1, kk) +
                    W8 * m001 * XC(ii, jj, kk + 1) + W8 * m101 * XC(ii + 1, jj, kk + 1) +
                    W8 * m011 * XC(ii, jj + 1, kk + 1) + W8 * m111 * XC(ii + 1, jj + 1, kk + 1);
            }
        }
    }
#undef XC
#undef XF
#undef IN
    o_fbound00(nxf, nyf, nzf, xout);
}

/* ---------------------------------------------------------------- Galerkin
 * VbuildG_7 / VbuildG_27 (buildGd.c:450-3115) compute A_c = P^T A P for the
 * operator without the Helmholtz term, with trilinear P and interior-only
 * fine rows.  Restated here as the generic triple product, summed over the 27
 * fine points of the coarse point's P column and the fine stencil. */
static double pweight(int dx, int dy, int dz)
{
    if (dx < -1 || dx > 1 || dy < -1 || dy > 1 || dz < -1 || dz > 1) return 0.0;
    int m = abs(dx) + abs(dy) + abs(dz);
    return m == 0 ? W1 : (m == 1 ? W2 : (m == 2 ? W4 : W8));
}

/* fine operator entry A(f, f+s) for |s|_inf <= 1 (without cc) */
static inline double fine_entry(int st, const double *af, size_t n, int nxf, int nyf,
                                int fi, int fj, int fk, int sx, int sy, int sz)
{
    if (sx == 0 && sy == 0 && sz == 0) return af[OIDX(fi, fj, fk, nxf, nyf)];
    int ns = st == 7 ? 4 : 14;
    for (int s = 1; s < ns; s++) {
        if (o_fwd[s][0] == sx && o_fwd[s][1] == sy && o_fwd[s][2] == sz)
            return -af[(size_t)s * n + OIDX(fi, fj, fk, nxf, nyf)];
        if (o_fwd[s][0] == -sx && o_fwd[s][1] == -sy && o_fwd[s][2] == -sz)
            return -af[(size_t)s * n + OIDX(fi + sx, fj + sy, fk + sz, nxf, nyf)];
    }
    return 0.0;
}

void orc_galerkin(int nxf, int nyf, int nzf, int st, const double *af, int nxc,
                  int nyc, int nzc, double *ac)
{
    size_t nf = (size_t)nxf * nyf * nzf, ncn = (size_t)nxc * nyc * nzc;
    memset(ac, 0, 14 * ncn * sizeof(double));
    (void)nzf;
    for (int K = 1; K < nzc - 1; K++)
        for (int J = 1; J < nyc - 1; J++)
            for (int I = 1; I < nxc - 1; I++) {
                int f0x = 2 * I, f0y = 2 * J, f0z = 2 * K;
                for (int s = 0; s < 14; s++) {
                    int ex = o_fwd[s][0], ey = o_fwd[s][1], ez = o_fwd[s][2];
                    int I2 = I + ex, J2 = J + ey, K2 = K + ez;
                    if (I2 < 1 || I2 > nxc - 2 || J2 < 1 || J2 > nyc - 2 || K2 < 1 || K2 > nzc - 2)
                        continue;  /* link to a boundary coarse point: zero */
                    int g0x = 2 * I2, g0y = 2 * J2, g0z = 2 * K2;
                    double sum = 0.0;
                    for (int dz = -1; dz <= 1; dz++)
                        for (int dy = -1; dy <= 1; dy++)
                            for (int dx = -1; dx <= 1; dx++) {
                                int fx = f0x + dx, fy = f0y + dy, fz = f0z + dz;
                                double p1 = pweight(dx, dy, dz);
                                double t = 0.0;
                                for (int sz = -1; sz <= 1; sz++)
                                    for (int sy = -1; sy <= 1; sy++)
                                        for (int sx = -1; sx <= 1; sx++) {
                                            if (st == 7 && abs(sx) + abs(sy) + abs(sz) > 1) continue;
                                            double p2 = pweight(fx + sx - g0x, fy + sy - g0y, fz + sz - g0z);
                                            if (p2 == 0.0) continue;
                                            t += fine_entry(st, af, nf, nxf, nyf, fx, fy, fz, sx, sy, sz) * p2;
                                        }
                                sum += p1 * t;
                            }
                    ac[(size_t)s * ncn + OIDX(I, J, K, nxc, nyc)] = (s == 0) ? sum : -sum;
                }
            }
}

/* ------------------------------------------------------------ coarse solves */

/* Vbuildband1_7 / Vbuildband1_27 (buildBd.c:119-306): upper band storage
 * abd(lda, n), diagonal in row m+1.  The diagonal is oC alone (no cc), and
 * the 27-point entries are taken from the coefficient at the positions the
 * reference reads (e.g. -oNE(i,j-1,k) for the (i+1,j-1,k) column). */
static void o_buildband(omg *g, olev *L)
{
    int nx = L->nx, ny = L->ny, nz = L->nz;
    size_t n = L->n;
    const double *a = L->a;
    int ni = nx - 2, nj = ny - 2, nk = nz - 2;
    int N = ni * nj * nk;
    int m = (L->st == 7) ? ni * nj : ni * nj + ni + 1;
    int lda = m + 1;
    g->bn = N; g->bm = m; g->blda = lda;
    g->band = (double *)calloc((size_t)lda * N, sizeof(double));
#define AB(r, c) g->band[(size_t)((c) - 1) * lda + ((r) - 1)]
    int jj = 0;
    for (int k = 1; k < nz - 1; k++)
        for (int j = 1; j < ny - 1; j++)
            for (int i = 1; i < nx - 1; i++) {
                jj++;
                struct { int d; int s; int ci, cj, ck; } e[14];
                int ne = 0;
#define PUSH(D, S, CI, CJ, CK) do { e[ne].d = (D); e[ne].s = (S); e[ne].ci = (CI); e[ne].cj = (CJ); e[ne].ck = (CK); ne++; } while (0)
                PUSH(0, A_OC, i, j, k);
                PUSH(1, A_OE, i - 1, j, k);
                PUSH(ni, A_ON, i, j - 1, k);
                if (L->st == 27) {
                    PUSH(ni - 1, A_ONE, i, j - 1, k);
                    PUSH(ni + 1, A_ONW, i, j - 1, k);
                }
                PUSH(ni * nj, A_UC, i, j, k - 1);
                if (L->st == 27) {
                    PUSH(ni * nj - 1, A_UE, i, j, k - 1);
                    PUSH(ni * nj + 1, A_UW, i, j, k - 1);
                    PUSH(ni * nj - ni, A_UN, i, j, k - 1);
                    PUSH(ni * nj + ni, A_US, i, j, k - 1);
                    PUSH(ni * nj - ni - 1, A_UNE, i, j, k - 1);
                    PUSH(ni * nj - ni + 1, A_UNW, i, j, k - 1);
                    PUSH(ni * nj + ni - 1, A_USE, i, j, k - 1);
                    PUSH(ni * nj + ni + 1, A_USW, i, j, k - 1);
                }
#undef PUSH
                for (int q = 0; q < ne; q++) {
                    int ii = jj - e[q].d;
                    int kk = ii - jj + m + 1;
                    double v = a[(size_t)e[q].s * n + IX(e[q].ci, e[q].cj, e[q].ck)];
                    AB(kk, jj) = (e[q].d == 0) ? v : -v;
                }
            }
#undef AB
}

/* LINPACK dpbfa (banded Cholesky), restated from the published algorithm
 * (mlinpckd.c:194-233).  Returns info (0 = SPD). */
static int o_dpbfa(double *abd, int lda, int n, int m)
{
#define AB(r, c) abd[(size_t)((c) - 1) * lda + ((r) - 1)]
    for (int j = 1; j <= n; j++) {
        double s = 0.0;
        int ik = m + 1;
        int jk = OMAX(j - m, 1);
        int mu = OMAX(m + 2 - j, 1);
        for (int k = mu; k <= m; k++) {
            double dot = 0.0;
            for (int l = 0; l < k - mu; l++) dot += AB(ik + l, jk) * AB(mu + l, j);
            double t = (AB(k, j) - dot) / AB(m + 1, jk);
            AB(k, j) = t;
            s += t * t;
            ik--;
            jk++;
        }
        s = AB(m + 1, j) - s;
        if (s <= 0.0) return j;
        AB(m + 1, j) = sqrt(s);
    }
    return 0;
}

/* LINPACK dpbsl (mlinpckd.c:57-83) */
static void o_dpbsl(const double *abd, int lda, int n, int m, double *b)
{
    for (int k = 1; k <= n; k++) {
        int lm = OMIN(k - 1, m), la = m + 1 - lm, lb = k - lm;
        double t = 0.0;
        for (int l = 0; l < lm; l++) t += AB(la + l, k) * b[lb - 1 + l];
        b[k - 1] = (b[k - 1] - t) / AB(m + 1, k);
    }
    for (int kb = 1; kb <= n; kb++) {
        int k = n + 1 - kb;
        int lm = OMIN(k - 1, m), la = m + 1 - lm, lb = k - lm;
        b[k - 1] /= AB(m + 1, k);
        double t = -b[k - 1];
        for (int l = 0; l < lm; l++) b[lb - 1 + l] += t * AB(la + l, k);
    }
#undef AB
}

/* Vcghs (cgd.c:57-123): CG on the coarsest level with the full operator
 * (oC + cc), x starts from zero, stops at ||r||2/||r0||2 <= errtol or itmax. */
static int o_cghs(olev *L, const double *cc, const double *fc, double *x,
                  double *p, double *ap, double *r, int itmax, double errtol)
{
    int nx = L->nx, ny = L->ny, nz = L->nz;
    int iters = 0;
    double rhok1 = 0.0;
    orc_resid(nx, ny, nz, L->st, L->a, cc, fc, x, r);
    double denom = o_nrm2(nx, ny, nz, r);
    if (denom == 0.0) return 0;
    for (;;) {
        double rhok2 = o_dot(nx, ny, nz, r, r);
        double rsnrm = sqrt(rhok2);
        if (rsnrm / denom <= errtol) break;
        if (iters >= itmax) break;
        if (iters == 0) o_copy(nx, ny, nz, r, p);
        else {
            double beta = rhok2 / rhok1;
            double alpha = 1.0 / beta;
            o_axpy(nx, ny, nz, alpha, r, p);
            o_scal(nx, ny, nz, beta, p);
        }
        orc_matvec(nx, ny, nz, L->st, L->a, cc, p, ap);
        double pAp = o_dot(nx, ny, nz, p, ap);
        double alpha = rhok2 / pAp;
        rhok1 = rhok2;
        o_axpy(nx, ny, nz, alpha, p, x);
        o_axpy(nx, ny, nz, -alpha, ap, r);
        iters++;
    }
    return iters;
}

/* deep_coarse variant: dense Cholesky of the full coarsest operator
 * (oC + cc and all couplings), matching the GPU schedule. */
static int o_dense_factor(omg *g, olev *L, const double *cc)
{
    int nx = L->nx, ny = L->ny, nz = L->nz;
    int ni = nx - 2, nj = ny - 2, nk = nz - 2, N = ni * nj * nk;
    size_t n = L->n;
    double *Ad = (double *)calloc((size_t)N * N, sizeof(double));
    for (int k = 1; k < nz - 1; k++)
        for (int j = 1; j < ny - 1; j++)
            for (int i = 1; i < nx - 1; i++) {
                int r = (i - 1) + ni * ((j - 1) + nj * (k - 1));
                Ad[(size_t)r * N + r] = L->a[IX(i, j, k)] + cc[IX(i, j, k)];
                int ns = L->st == 7 ? 4 : 14;
                for (int s = 1; s < ns; s++) {
                    int i2 = i + o_fwd[s][0], j2 = j + o_fwd[s][1], k2 = k + o_fwd[s][2];
                    if (i2 < 1 || i2 > nx - 2 || j2 < 1 || j2 > ny - 2 || k2 < 1 || k2 > nz - 2) continue;
                    int c = (i2 - 1) + ni * ((j2 - 1) + nj * (k2 - 1));
                    double v = -L->a[(size_t)s * n + IX(i, j, k)];
                    Ad[(size_t)r * N + c] = v;
                    Ad[(size_t)c * N + r] = v;
                }
            }
    /* in-place Cholesky, lower triangle */
    for (int j = 0; j < N; j++) {
        double s = Ad[(size_t)j * N + j];
        for (int l = 0; l < j; l++) s -= Ad[(size_t)j * N + l] * Ad[(size_t)j * N + l];
        if (s <= 0.0) { free(Ad); return 1; }
        double d = sqrt(s);
        Ad[(size_t)j * N + j] = d;
        for (int i = j + 1; i < N; i++) {
            double t = Ad[(size_t)i * N + j];
            for (int l = 0; l < j; l++) t -= Ad[(size_t)i * N + l] * Ad[(size_t)j * N + l];
            Ad[(size_t)i * N + j] = t / d;
        }
    }
    free(g->dense);
    g->dense = Ad;
    g->dn = N;
    return 0;
}

static void o_dense_solve(omg *g, olev *L, const double *f, double *x)
{
    int nx = L->nx, ny = L->ny, N = g->dn;
    int ni = nx - 2, nj = ny - 2;
    double *b = (double *)malloc((size_t)N * sizeof(double));
    for (int r = 0; r < N; r++) {
        int i = r % ni + 1, j = (r / ni) % nj + 1, k = r / (ni * nj) + 1;
        b[r] = f[IX(i, j, k)];
    }
    const double *Ad = g->dense;
    for (int r = 0; r < N; r++) {
        double t = b[r];
        for (int l = 0; l < r; l++) t -= Ad[(size_t)r * N + l] * b[l];
        b[r] = t / Ad[(size_t)r * N + r];
    }
    for (int r = N - 1; r >= 0; r--) {
        double t = b[r];
        for (int l = r + 1; l < N; l++) t -= Ad[(size_t)l * N + r] * b[l];
        b[r] = t / Ad[(size_t)r * N + r];
    }
    o_zero(x, L->n);
    for (int r = 0; r < N; r++) {
        int i = r % ni + 1, j = (r / ni) % nj + 1, k = r / (ni * nj) + 1;
        x[IX(i, j, k)] = b[r];
    }
    free(b);
}

/* ---------------------------------------------------------------- V-cycle */

static void o_coarse_solve(omg *g, int lev, double **CC, const double *f, double *x)
{
    olev *L = &g->L[lev];
    int nx = L->nx, ny = L->ny, nz = L->nz;
    if (g->deep) {
        o_dense_solve(g, L, f, x);
        return;
    }
    if (g->mgsolv == 0) {
        /* mgcsd.c:398-417: CG, 100 iterations, errtol = machine epsilon */
        o_zero(x, L->n);
        o_cghs(L, CC[lev], f, x, g->w1, g->w2, g->w3, 100, DBL_EPSILON);
    } else {
        /* mgcsd.c:419-437: banded solve of the interior block */
        int ni = nx - 2, nj = ny - 2, N = g->bn;
        double *b = (double *)malloc((size_t)N * sizeof(double));
        for (int r = 0; r < N; r++) {
            int i = r % ni + 1, j = (r / ni) % nj + 1, k = r / (ni * nj) + 1;
            b[r] = f[IX(i, j, k)];
        }
        o_dpbsl(g->band, g->blda, N, g->bm, b);
        for (int r = 0; r < N; r++) {
            int i = r % ni + 1, j = (r / ni) % nj + 1, k = r / (ni * nj) + 1;
            x[IX(i, j, k)] = b[r];
        }
        free(b);
        o_fbound00(nx, ny, nz, x);
    }
}

static void o_smooth(omg *g, int lev, const double *cc, const double *f, double *x,
                     int nu, int iadjoint, double *r /* residual out or NULL */)
{
    olev *L = &g->L[lev];
    orc_gsrb(L->nx, L->ny, L->nz, L->st, L->a, cc, f, x, nu, iadjoint, g->colour8);
    if (r) orc_resid(L->nx, L->ny, L->nz, L->st, L->a, cc, f, x, r);
}

/* Vmvcs (mgcsd.c:57-565), istop 0 (absolute) or 1 (relative to ||f||1).
 * X[l] per-level solution storage, CC[l] per-level Helmholtz diagonal, f the
 * fine rhs.  Returns ierror (1 = itmax reached). */
static int o_mvcs(omg *g, double **X, double **CC, const double *f, int istop,
                  int itmax, double errtol, int *iters_out)
{
    olev *L = g->L;
    int nlev = g->ntot;
    int nx = L[0].nx, ny = L[0].ny, nz = L[0].nz;
    double rsden = (istop == 0) ? 1.0 : o_nrm1(nx, ny, nz, f);
    if (rsden == 0.0) rsden = 1.0;
    double rsnrm = rsden;
    int iters = 0;
    if (nlev == 1) {
        o_coarse_solve(g, 0, CC, f, X[0]);
        orc_resid(nx, ny, nz, L[0].st, L[0].a, CC[0], f, X[0], g->w1);
        rsnrm = o_nrm1(nx, ny, nz, g->w1);
        *iters_out = 1;
        g->relres = rsnrm / rsden;
        return 0;
    }
    do {
        /* pre-smoothing on the fine level with residual into w1 */
        o_smooth(g, 0, CC[0], f, X[0], g->nu1, 0, g->w1);
        o_copy(nx, ny, nz, g->w1, L[0].w0);
        for (int l = 1; l < nlev; l++) {
            orc_restrict(L[l - 1].nx, L[l - 1].ny, L[l - 1].nz, L[l].nx, L[l].ny, L[l].nz,
                         g->w1, L[l].w0);
            if (l != nlev - 1) {
                o_zero(X[l], L[l].n);
                o_smooth(g, l, CC[l], L[l].w0, X[l], g->nu1, 0, g->w1);
            }
        }
        o_coarse_solve(g, nlev - 1, CC, L[nlev - 1].w0, X[nlev - 1]);
        for (int l = nlev - 2; l >= 0; l--) {
            olev *C = &L[l + 1];
            orc_interp(C->nx, C->ny, C->nz, L[l].nx, L[l].ny, L[l].nz, X[l + 1], g->w1);
            /* Hackbusch-Reusken damping (mgcsd.c:463-475) */
            orc_matvec(C->nx, C->ny, C->nz, C->st, C->a, CC[l + 1], X[l + 1], g->w2);
            double xnum = o_dot(C->nx, C->ny, C->nz, X[l + 1], C->w0);
            double xden = o_dot(C->nx, C->ny, C->nz, X[l + 1], g->w2);
            double xdamp = xnum / xden;
            o_axpy(L[l].nx, L[l].ny, L[l].nz, xdamp, g->w1, X[l]);
            o_smooth(g, l, CC[l], (l == 0) ? f : L[l].w0, X[l], g->nu2, 1, NULL);
        }
        iters++;
        orc_resid(nx, ny, nz, L[0].st, L[0].a, CC[0], f, X[0], g->w1);
        rsnrm = o_nrm1(nx, ny, nz, g->w1);
    } while (iters < itmax && (rsnrm / rsden) > errtol);
    *iters_out = iters;
    g->relres = rsnrm / rsden;
    return iters < itmax ? 0 : 1;
}

/* ----------------------------------------------------------- build / drive */

static void alloc_level(olev *L, int nx, int ny, int nz, int st, int newton)
{
    L->nx = nx; L->ny = ny; L->nz = nz; L->st = st;
    L->n = (size_t)nx * ny * nz;
    L->a = (double *)calloc((st == 7 ? 4 : 14) * L->n, sizeof(double));
    L->cc = (double *)calloc(L->n, sizeof(double));
    L->fc = (double *)calloc(L->n, sizeof(double));
    L->x = (double *)calloc(L->n, sizeof(double));
    L->w0 = (double *)calloc(L->n, sizeof(double));
    if (newton) {
        L->cprime = (double *)calloc(L->n, sizeof(double));
        L->rhs = (double *)calloc(L->n, sizeof(double));
        L->xtmp = (double *)calloc(L->n, sizeof(double));
    }
}

static int coarsen(int n) { return (n - 1) / 2 + 1; }

/* Vbuildops (mgsubd.c:61-258) with mgcoar=2 (Galerkin), mgprol=0, mgdisc=0 */
int omg_build(omg *g, int nx, int ny, int nz, int nlev, int pbetype,
              const double *xf, const double *yf, const double *zf,
              const double *gxcf, const double *gycf, const double *gzcf,
              const double *a1cf, const double *a2cf, const double *a3cf,
              const double *ccf, const double *fcf)
{
    int newton = (pbetype == ORC_NPBE);
    g->nlev = nlev;
    g->mgsolv = newton ? 0 : 1;       /* vpmgp.c:155-166 */
    int ntot = nlev;
    if (g->deep) {
        int cx = nx, cy = ny, cz = nz;
        for (int l = 1; l < nlev; l++) { cx = coarsen(cx); cy = coarsen(cy); cz = coarsen(cz); }
        while (((cx - 1) % 2 == 0) && ((cy - 1) % 2 == 0) && ((cz - 1) % 2 == 0) &&
               coarsen(cx) >= 3 && coarsen(cy) >= 3 && coarsen(cz) >= 3 &&
               (cx - 2) * (cy - 2) * (cz - 2) > 64) {
            cx = coarsen(cx); cy = coarsen(cy); cz = coarsen(cz); ntot++;
        }
    }
    g->ntot = ntot;
    alloc_level(&g->L[0], nx, ny, nz, 7, newton);
    olev *F = &g->L[0];
    orc_buildA_fv(nx, ny, nz, xf, yf, zf, gxcf, gycf, gzcf, a1cf, a2cf, a3cf, ccf, fcf,
                  F->a, F->cc, F->fc, F->a + F->n, F->a + 2 * F->n, F->a + 3 * F->n);
    int cx = nx, cy = ny, cz = nz;
    for (int l = 1; l < ntot; l++) {
        int px = cx, py = cy, pz = cz;
        cx = coarsen(px); cy = coarsen(py); cz = coarsen(pz);
        alloc_level(&g->L[l], cx, cy, cz, 27, newton);
        olev *P = &g->L[l - 1], *C = &g->L[l];
        orc_galerkin(px, py, pz, P->st, P->a, cx, cy, cz, C->a);
        orc_restrict(px, py, pz, cx, cy, cz, P->cc, C->cc);
        orc_restrict(px, py, pz, cx, cy, cz, P->fc, C->fc);
    }
    size_t nf = F->n;
    g->w1 = (double *)calloc(nf, sizeof(double));
    g->w2 = (double *)calloc(nf, sizeof(double));
    g->w3 = (double *)calloc(nf, sizeof(double));
    if (g->deep) {
        if (!newton) return o_dense_factor(g, &g->L[ntot - 1], g->L[ntot - 1].cc) == 0;
        return 1;
    }
    if (g->mgsolv == 1) {
        olev *C = &g->L[ntot - 1];
        o_buildband(g, C);
        if (o_dpbfa(g->band, g->blda, g->bn, g->bm) != 0) g->mgsolv = 0;  /* mgsubd.c:244-255 */
    }
    return 1;
}

void omg_free(omg *g)
{
    for (int l = 0; l < g->ntot; l++) {
        olev *L = &g->L[l];
        free(L->a); free(L->cc); free(L->fc); free(L->x); free(L->w0);
        free(L->cprime); free(L->rhs); free(L->xtmp);
    }
    free(g->w1); free(g->w2); free(g->w3); free(g->band); free(g->dense);
    memset(g, 0, sizeof(*g));
}

/* Vmgdriv2 (mgdrvd.c:190-560): u = 0, V-cycles, caller restores the faces */
int omg_solve_linear(omg *g, double *u)
{
    olev *F = &g->L[0];
    double *X[24], *CC[24];
    X[0] = u;
    CC[0] = F->cc;
    for (int l = 1; l < g->ntot; l++) { X[l] = g->L[l].x; CC[l] = g->L[l].cc; }
    for (int l = 0; l < g->ntot; l++) o_zero(X[l], g->L[l].n);
    int iters = 0;
    int ierr = o_mvcs(g, X, CC, F->fc, 1, g->itmax, g->errtol, &iters);
    g->iters = iters;
    return ierr;
}

/* Vnewton (newtond.c:162-548) with Vgetjac (:550-598) */
int omg_solve_newton(omg *g, double *u)
{
    olev *L = g->L;
    int nx = L[0].nx, ny = L[0].ny, nz = L[0].nz;
    size_t nf = L[0].n;
    double *w0 = (double *)calloc(nf, sizeof(double));
    double *wn1 = (double *)calloc(nf, sizeof(double));  /* Newton's w1 */
    double *wn2 = (double *)calloc(nf, sizeof(double));
    double *wn3 = (double *)calloc(nf, sizeof(double));
    double *X[24], *CP[24];
    o_fbound00(nx, ny, nz, u);
    /* rsden = ||F(0)||1 (newtond.c:218-228) */
    o_zero(wn1, nf);
    o_nresid(g, 0, L[0].cc, L[0].fc, wn1, wn2, wn3);
    double rsden = o_nrm1(nx, ny, nz, wn2);
    if (rsden == 0.0) rsden = 1.0;
    double rsnrm = rsden;
    o_nresid(g, 0, L[0].cc, L[0].fc, u, w0, wn2);
    double xnorm_old = o_nrm1(nx, ny, nz, w0);
    double xnorm_den = rsden, xnorm_new = xnorm_old, xnorm_med = 0.0;
    int idamp = 1, iters = 0, inner_total = 0;
    for (;;) {
        iters++;
        /* Vgetjac: rhs = F(x), cprime = N'(x), both restricted down */
        o_copy(nx, ny, nz, w0, L[0].rhs);
        orc_dcvec((int)nf, g->nion, g->q, g->sconc, L[0].cc, u, L[0].cprime);
        for (int l = 1; l < g->ntot; l++) {
            orc_restrict(L[l - 1].nx, L[l - 1].ny, L[l - 1].nz, L[l].nx, L[l].ny, L[l].nz,
                         L[l - 1].rhs, L[l].rhs);
            orc_restrict(L[l - 1].nx, L[l - 1].ny, L[l - 1].nz, L[l].nx, L[l].ny, L[l].nz,
                         L[l - 1].cprime, L[l].cprime);
        }
        double errtol_s = OMIN(0.9 * xnorm_old, 1.0 * pow(xnorm_old, 2.0));
        for (int l = 0; l < g->ntot; l++) { X[l] = L[l].xtmp; CP[l] = L[l].cprime; o_zero(X[l], L[l].n); }
        if (g->deep) o_dense_factor(g, &L[g->ntot - 1], L[g->ntot - 1].cprime);
        int iters_s = 0;
        o_mvcs(g, X, CP, L[0].rhs, 0, 1000, errtol_s, &iters_s);
        inner_total += iters_s;
        double *xtmp = L[0].xtmp;
        if (idamp == 1) {
            o_copy(nx, ny, nz, u, wn1);
            o_axpy(nx, ny, nz, 1.0, xtmp, wn1);
            o_nresid(g, 0, L[0].cc, L[0].fc, wn1, w0, L[0].rhs);
            xnorm_new = o_nrm1(nx, ny, nz, w0);
            double damp = 1.0;
            int iter_d = 0, itmax_d = 10, mode = 0;
            while (iter_d < itmax_d) {
                if (mode == 0) {
                    if (xnorm_new < xnorm_old) mode = 1;
                } else if (xnorm_new > xnorm_med) break;
                o_copy(nx, ny, nz, wn1, wn2);
                o_copy(nx, ny, nz, w0, wn3);
                xnorm_med = xnorm_new;
                o_copy(nx, ny, nz, u, wn1);
                damp = damp / 2.0;
                o_axpy(nx, ny, nz, damp, xtmp, wn1);
                o_nresid(g, 0, L[0].cc, L[0].fc, wn1, w0, L[0].rhs);
                xnorm_new = o_nrm1(nx, ny, nz, w0);
                iter_d++;
            }
            o_copy(nx, ny, nz, wn2, u);
            o_copy(nx, ny, nz, wn3, w0);
            xnorm_new = xnorm_med;
            xnorm_old = xnorm_new;
            if ((iter_d - 1) == 0) idamp = 0;
        } else {
            o_axpy(nx, ny, nz, 1.0, xtmp, u);
            o_nresid(g, 0, L[0].cc, L[0].fc, u, w0, L[0].rhs);
            xnorm_new = o_nrm1(nx, ny, nz, w0);
            xnorm_old = xnorm_new;
        }
        rsnrm = xnorm_new;
        (void)xnorm_den;
        if ((rsnrm / rsden) <= g->errtol) break;
        if (iters >= g->newton_itmax) break;
    }
    g->iters = iters;
    g->inner_total = inner_total;
    g->relres = rsnrm / rsden;
    free(w0); free(wn1); free(wn2); free(wn3);
    return iters < g->itmax ? 0 : 1;
}

void orc_fbound_apply(int nx, int ny, int nz, double *u, const double *gx,
                      const double *gy, const double *gz)
{
    o_fbound(nx, ny, nz, u, gx, gy, gz);
}
