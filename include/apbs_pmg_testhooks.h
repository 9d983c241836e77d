/*
 * apbs_pmg_testhooks.h -- single-kernel entry points of libapbs_pmg.so used
 * by the kernel-by-kernel parity tests.  Host arrays in, host arrays out; no
 * counterpart in the reference (they expose the stencil kernels that the
 * reference runs inside Vmvcs / Vbuildops).
 */
#ifndef APBS_PMG_TESTHOOKS_H
#define APBS_PMG_TESTHOOKS_H

#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif
/* Vgsrb7x / 27-point 8-colour sweeps (src/pmgc/gsd.c:95-340); a packs 4
 * (stencil 7) or 14 (stencil 27) operator arrays of nx*ny*nz each. */
int pmgtest_gsrb(int nx, int ny, int nz, int stencil, const double *a, const double *cc,
                 const double *fc, double *x, int nu, int iadjoint);
/* Vmresid7_1s / Vmresid27_1s (src/pmgc/matvecd.c:461-596) */
int pmgtest_resid(int nx, int ny, int nz, int stencil, const double *a, const double *cc,
                  const double *fc, const double *x, double *r);
/* Vrestrc2 (src/pmgc/matvecd.c:801-913) */
int pmgtest_restrict(int nxf, int nyf, int nzf, int nxc, int nyc, int nzc, const double *in,
                     double *out);
/* VinterpPMG2 + Vxaxpy (src/pmgc/matvecd.c:935-1076, mgcsd.c:485-487) */
int pmgtest_prolong(int nxc, int nyc, int nzc, int nxf, int nyf, int nzf, const double *xc,
                    double *xf, double damp);
/* VbuildG_7 / VbuildG_27 (src/pmgc/buildGd.c:450-3115) */
int pmgtest_galerkin(int nxf, int nyf, int nzf, int stencil, const double *af, int nxc, int nyc,
                     int nzc, double *ac);
/* The same with the kernel form forced: 0 by level size, 1 one thread per
 * coefficient, 2 one 32-lane group per coefficient, 3 one thread per
 * coefficient with the loops unrolled at compile time (all give the same
 * bits) */
int pmgtest_galerkin_form(int nxf, int nyf, int nzf, int stencil, const double *af, int nxc,
                          int nyc, int nzc, double *ac, int form);
/* VbuildA_fv (src/pmgc/buildAd.c:105-281) */
int pmgtest_buildA(int nx, int ny, int nz, const double *xf, const double *yf, const double *zf,
                   const double *gx, const double *gy, const double *gz, const double *a1,
                   const double *a2, const double *a3, const double *ccf, const double *fcf,
                   double *oC, double *cc, double *fc, double *oE, double *oN, double *uC);
/* Vnmresid7_1s with Vc_vecpmg, and Vdc_vecpmg (src/pmgc/matvecd.c:633-674,
 * src/pmgc/mypdec.c:145-223, 366-433) */
int pmgtest_nonlinear(int nx, int ny, int nz, const double *a, const double *cc,
                      const double *fc, const double *x, int nion, const double *q,
                      const double *sconc, double *out_r, double *out_d);
/* pmgtest_gsrb (7-point) with an explicit tile shape of the fused sweep
 * (rows*100 + planes: 416, 48, 816, 88, 84, 1616, 832). */
int pmgtest_gsrb_tile(int nx, int ny, int nz, const double *a, const double *cc,
                      const double *fc, double *x, int nu, int iadjoint, int tile);
/* Two fused 7-point sweeps from a zero start (the first smoothing of a
 * mixed-precision correction cycle): x's boundary layer must be zero and its
 * interior is not read; tile as pmgtest_gsrb_tile (0: the default shape; a
 * shape without the zero-start form clears x first).  Result in x. */
int pmgtest_gsrb_zero(int nx, int ny, int nz, const double *a, const double *cc,
                      const double *fc, double *x, int iadjoint, int tile);
/* Fused FP32 7-point sweeps on explicit operator arrays (oE, oN, uC, the
 * diagonal D, fc and x: nx*ny*nz floats each): nu sweeps, or with zero set the
 * zero-start pair (x's interior not read); tile as pmgtest_gsrb_tile.  Result
 * in x; returns 0 on bad sizes. */
int pmgtest_fused32(int nx, int ny, int nz, const float *oE, const float *oN, const float *uC,
                    const float *D, const float *fc, float *x, int nu, int iadjoint, int tile,
                    int zero);
/* pmgtest_gsrb (27-point) with the kernel form forced: mode 0 one launch per
 * colour, 1 one launch per plane parity (tile by level size), 2 one launch
 * per colour pair (nx <= 130, else returns 0; the vector kernel where the
 * row allows it), 3 the colour pair with one point per lane, 8 / 16 / 32 one
 * launch per plane parity with that square tile.  All give the same bits. */
int pmgtest_gsrb27_mode(int nx, int ny, int nz, const double *a, const double *cc,
                        const double *fc, double *x, int nu, int iadjoint, int mode);
/* FP32 27-point sweeps in a given form (mode as pmgtest_gsrb27_mode, and 4
 * the colour-pair vector kernel with 8-byte runs) on
 * explicit arrays: a holds the 14 stencil arrays (oC first, unused), D the
 * diagonal; every array nx*ny*nz floats.  Result in x; 0 on bad sizes. */
int pmgtest_gsrb27_32(int nx, int ny, int nz, const float *a, const float *D, const float *fc,
                      float *x, int nu, int iadjoint, int mode);
/* FP32 7-point residual r = fc + sum(a x) - D x on the interior (oE, oN, uC,
 * D, fc, x, r: nx*ny*nz floats; r's boundary is kept).  0 on bad sizes. */
int pmgtest_resid32(int nx, int ny, int nz, const float *oE, const float *oN, const float *uC,
                    const float *D, const float *fc, const float *x, float *r);
/* No reference counterpart: the 27-point sweeps alone on a synthetic n^3
 * coarse-level operator, word 8 or 4, mode as above; average microseconds
 * per sweep in *us_per_sweep. */
int pmgtest_gs27_sweeps(int n, int word, int sweeps, int mode, double *us_per_sweep);
/* No reference counterpart (the reference factors its coarsest operator in
 * band form, mgsubd.c:244-255): the dense inverse of the coarsest level's
 * operator, form 0 (by size: registers for N <= 64) or 1 (LDS / global). */
int pmgtest_dense_invert(int N, const double *M, double *out, int form);
/* No reference counterpart: microseconds per Galerkin call (form as above)
 * on a synthetic nf^3 operator of stencil st (nf odd). */
int pmgtest_galerkin_time(int nf, int st, int form, int reps, double *us);
/* No reference counterpart: reps device copies of n doubles, the known-size
 * calibration for hardware byte counters (FETCH_SIZE / WRITE_SIZE). */
int pmgtest_stream_copy(size_t n, int reps);
/* No reference counterpart: the fused red-black sweep alone on a synthetic
 * n^3 7-point operator, word 8 (FP64) or 4 (FP32), `sweeps` (even) timed
 * launches; average microseconds per launch in *us_per_sweep.  The same
 * launches as the solver's fine level, for counter and timing runs. */
int pmgtest_fused_sweeps(int n, int word, int sweeps, int tile, double *us_per_sweep);
/* No reference counterpart: on != 0 makes the multigrid hierarchies a Vpmg
 * (passed as void *) allocates from then on start with NaN in every array's
 * interior (the boundary layers are cleared as usual), to show that every
 * interior entry is written before it is read. */
int pmgtest_poison_hierarchy(void *pmg, int on);
#ifdef __cplusplus
}
#endif
#endif
