/*
 * orc_int.h -- internal structures of the oracle (test infrastructure only).
 * The multigrid hierarchy mirrors the per-level storage that the reference
 * packs into its rwork/iwork arrays (src/pmgc/mgsubd.c:261-342), but as
 * separate arrays per level.
 */
#ifndef ORC_INT_H
#define ORC_INT_H

#include <stddef.h>
#include "orc.h"

#define OIDX(i, j, k, nx, ny) ((size_t)(k) * (size_t)(nx) * (size_t)(ny) + (size_t)(j) * (size_t)(nx) + (size_t)(i))
#define OMIN(a, b) ((a) <= (b) ? (a) : (b))
#define OMAX(a, b) ((a) >= (b) ? (a) : (b))
#define OSQR(x) ((x) * (x))

/* MALOC constants used on the path (FETK/MALOC, fetched commit
 * 57195e55351e04ce6ee0ef56a143c996a9aee7e2 per CMakeLists.txt:158-171; the
 * library is absent from the reference tree, values restated from its
 * published maloc_base.h). */
#define O_VSMALL 1.0e-9
#define O_VLARGE 1.0e9
#define O_VPI 3.14159265358979323846
/* src/generic/vhal.h:444 */
#define O_VPMGSMALL 1e-12
/* src/pmgc/mypdec.h:62-67 */
#define O_ZSMALL 1.0e-20
#define O_ZLARGE 1.0e20
#define O_SINH_MIN -85.0
#define O_SINH_MAX 85.0
/* src/generic/vcap.h:72-77 */
#define O_EXPMAX 85.00
#define O_EXPMIN -85.00
/* src/generic/vunit.h:68-116 */
#define O_Vunit_ec 1.6021773e-19
#define O_Vunit_kb 1.3806581e-23
#define O_Vunit_Na 6.0221367e+23
#define O_Vunit_eps0 8.8541878e-12

/* Index of the 14 operator arrays (reference ac layout, gsd.c:74-87). */
enum { A_OC = 0, A_OE, A_ON, A_UC, A_ONE, A_ONW, A_UE, A_UW, A_UN, A_US,
       A_UNE, A_UNW, A_USE, A_USW };

/* Forward offset of stored coefficient s (1..13): coefficient s at point p
 * couples p and p + off[s] (matvecd.c:153-229). */
extern const int o_fwd[14][3];

typedef struct olev {
    int nx, ny, nz;
    size_t n;
    int st;            /* 7 or 27 */
    double *a;         /* 4 or 14 arrays of n */
    double *cc, *fc;   /* Helmholtz diagonal and rhs on this level */
    double *x;         /* solution / correction storage of this level */
    double *w0;        /* restricted residual storage */
    double *cprime, *rhs, *xtmp;  /* Newton storage */
} olev;

typedef struct omg {
    int nlev;          /* reference nlev */
    int ntot;          /* levels allocated (>= nlev with deep_coarse) */
    olev L[24];
    int mgsolv;        /* 1 banded LINPACK, 0 CG */
    double *band; int bn, bm, blda;
    double *dense; int dn;   /* deep_coarse variant: dense Cholesky factor */
    int colour8, deep;
    double *w1, *w2, *w3;    /* fine-size scratch */
    int nion; double q[ORC_MAXION], sconc[ORC_MAXION];
    double errtol;
    int itmax, nu1, nu2;
    int newton_itmax;
    int iters, inner_total; double relres;
} omg;

/* orc_mg.c */
void o_zero(double *x, size_t n);
void o_fbound00(int nx, int ny, int nz, double *x);
double o_nrm1(int nx, int ny, int nz, const double *x);
double o_nrm2(int nx, int ny, int nz, const double *x);
double o_dot(int nx, int ny, int nz, const double *x, const double *y);
int omg_build(omg *g, int nx, int ny, int nz, int nlev, int pbetype,
              const double *xf, const double *yf, const double *zf,
              const double *gxcf, const double *gycf, const double *gzcf,
              const double *a1cf, const double *a2cf, const double *a3cf,
              const double *ccf, const double *fcf);
int omg_solve_linear(omg *g, double *u);
int omg_solve_newton(omg *g, double *u);
void omg_free(omg *g);

#endif
