// testhooks.cpp -- single-kernel entry points for kernel-by-kernel parity
// tests (host arrays in, host arrays out).  Not part of the Vpmg API.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "pmg_device.h"

#define TCHECK(x)                                                                          \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) {                                                                \
      fprintf(stderr, "pmgtest: HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, \
              __LINE__);                                                                   \
      abort();                                                                             \
    }                                                                                      \
  } while (0)

namespace {
double* up(const double* h, size_t n) {
  double* d = nullptr;
  TCHECK(hipMalloc(&d, (n ? n : 1) * sizeof(double)));
  if (h) TCHECK(hipMemcpy(d, h, n * sizeof(double), hipMemcpyHostToDevice));
  return d;
}
void down(double* h, const double* d, size_t n) {
  TCHECK(hipMemcpy(h, d, n * sizeof(double), hipMemcpyDeviceToHost));
}
struct OpsHolder {
  std::vector<double*> ptr;
  pmg::Ops<double> ops{};
  OpsHolder(const double* a, size_t n, int st) {
    const int na = st == 7 ? 4 : 14;
    for (int s = 0; s < 14; s++) ops.a[s] = nullptr;
    for (int s = 0; s < na; s++) {
      ptr.push_back(up(a + (size_t)s * n, n));
      ops.a[s] = ptr.back();
    }
  }
  ~OpsHolder() {
    for (double* p : ptr) TCHECK(hipFree(p));
  }
};
}  // namespace

__global__ void k_stream_copy(const double* __restrict__ a, double* __restrict__ b, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

// The fused red-black sweep alone on an n^3 7-point operator (diagonally
// dominant, values irrelevant to its memory traffic): one untimed sweep, then
// `sweeps` (even) timed launches on the null stream.
template <class T>
double fused_sweeps_us(int n, int sweeps, int tile) {
  const size_t N = (size_t)n * n * n;
  std::vector<T> h(N);
  T* d[8];
  const T val[8] = {T(6.5), T(1), T(1), T(1), T(6.5), T(1), T(0), T(0)};  // oC oE oN uC D fc x t
  for (int a = 0; a < 8; a++) {
    TCHECK(hipMalloc(&d[a], N * sizeof(T)));
    for (size_t i = 0; i < N; i++) h[i] = val[a];
    TCHECK(hipMemcpy(d[a], h.data(), N * sizeof(T), hipMemcpyHostToDevice));
  }
  pmg::Ops<T> ops{};
  for (int a = 0; a < 4; a++) ops.a[a] = d[a];
  pmg::gsrb7_fused<T>(0, n, n, n, ops, d[4], d[5], d[6], d[7], 2, 0, tile);
  hipEvent_t e0, e1;
  TCHECK(hipEventCreate(&e0));
  TCHECK(hipEventCreate(&e1));
  TCHECK(hipEventRecord(e0, 0));
  pmg::gsrb7_fused<T>(0, n, n, n, ops, d[4], d[5], d[6], d[7], sweeps, 0, tile);
  TCHECK(hipEventRecord(e1, 0));
  TCHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  TCHECK(hipEventElapsedTime(&ms, e0, e1));
  TCHECK(hipEventDestroy(e0));
  TCHECK(hipEventDestroy(e1));
  for (int a = 0; a < 8; a++) TCHECK(hipFree(d[a]));
  return 1e3 * ms / sweeps;
}

// The 27-point 8-colour sweeps alone on an n^3 coarse-level operator: mode 0
// one launch per colour, mode 1 one per plane parity (tile by level size),
// 8 / 16 / 32 one per plane parity with that tile.  One untimed call, then
// one timed call of `sweeps` sweeps; microseconds per sweep.
template <class T>
double gs27_sweeps_us(int n, int sweeps, int mode) {
  const size_t N = (size_t)n * n * n;
  std::vector<T> h(N);
  T* a[14];
  T *D, *fc, *x, *xs;
  for (int s = 0; s < 14; s++) {
    TCHECK(hipMalloc(&a[s], N * sizeof(T)));
    for (size_t i = 0; i < N; i++) h[i] = s == 0 ? T(26) : T(1);
    TCHECK(hipMemcpy(a[s], h.data(), N * sizeof(T), hipMemcpyHostToDevice));
  }
  T** bufs[4] = {&D, &fc, &x, &xs};
  const T val[4] = {T(26), T(1), T(0), T(0)};
  for (int b = 0; b < 4; b++) {
    TCHECK(hipMalloc(bufs[b], N * sizeof(T)));
    for (size_t i = 0; i < N; i++) h[i] = val[b];
    TCHECK(hipMemcpy(*bufs[b], h.data(), N * sizeof(T), hipMemcpyHostToDevice));
  }
  pmg::Ops<T> ops{};
  for (int s = 0; s < 14; s++) ops.a[s] = a[s];
  auto run = [&](int nu) {
    if (mode == 0)
      pmg::gsrb27<T>(0, n, n, n, ops, D, fc, x, nu, 0);
    else if (mode >= 2 && mode <= 4) {
      const pmg::LevelArgs<T> c{ops, D, fc, x, nullptr, nullptr};
      pmg::gsrb27_pairs_form<T>(0, n, n, n, &c, 1, nu, 0, mode - 2);
    }
    else
      pmg::gsrb27_planes<T>(0, n, n, n, ops, D, fc, x, xs, nu, 0, mode);
  };
  run(2);
  hipEvent_t e0, e1;
  TCHECK(hipEventCreate(&e0));
  TCHECK(hipEventCreate(&e1));
  TCHECK(hipEventRecord(e0, 0));
  run(sweeps);
  TCHECK(hipEventRecord(e1, 0));
  TCHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  TCHECK(hipEventElapsedTime(&ms, e0, e1));
  TCHECK(hipEventDestroy(e0));
  TCHECK(hipEventDestroy(e1));
  for (int s = 0; s < 14; s++) TCHECK(hipFree(a[s]));
  for (int b = 0; b < 4; b++) TCHECK(hipFree(*bufs[b]));
  return 1e3 * ms / sweeps;
}

// Galerkin coarse operator of a synthetic nf^3 operator (stencil st) in the
// given kernel form: one untimed call, then reps timed calls; microseconds
// per call.
double galerkin_us(int nf, int st, int form, int reps) {
  const int nc = (nf - 1) / 2 + 1;
  const size_t Nf = (size_t)nf * nf * nf, Nc = (size_t)nc * nc * nc;
  const int na = st == 7 ? 4 : 14;
  std::vector<double> h(Nf);
  pmg::Ops<double> ops{};
  for (int s = 0; s < 14; s++) ops.a[s] = nullptr;
  for (int s = 0; s < na; s++) {
    for (size_t i = 0; i < Nf; i++) h[i] = s == 0 ? 26.0 : 1.0 + 1e-3 * (double)(i % 7);
    ops.a[s] = up(h.data(), Nf);
  }
  double* ac[14];
  for (int s = 0; s < 14; s++) ac[s] = up(nullptr, Nc);
  pmg::galerkin(0, st, nf, nf, nf, ops, nc, nc, nc, ac, form);
  hipEvent_t e0, e1;
  TCHECK(hipEventCreate(&e0));
  TCHECK(hipEventCreate(&e1));
  TCHECK(hipEventRecord(e0, 0));
  for (int r = 0; r < reps; r++) pmg::galerkin(0, st, nf, nf, nf, ops, nc, nc, nc, ac, form);
  TCHECK(hipEventRecord(e1, 0));
  TCHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  TCHECK(hipEventElapsedTime(&ms, e0, e1));
  TCHECK(hipEventDestroy(e0));
  TCHECK(hipEventDestroy(e1));
  for (int s = 0; s < na; s++) TCHECK(hipFree((void*)ops.a[s]));
  for (int s = 0; s < 14; s++) TCHECK(hipFree(ac[s]));
  return 1e3 * ms / reps;
}

extern "C" {

// Calibration for hardware byte counters: reps copies of n doubles (n*8
// bytes read and n*8 bytes written per launch) between two device buffers.
int pmgtest_stream_copy(size_t n, int reps) {
  double *a = up(nullptr, n), *b = up(nullptr, n);
  TCHECK(hipMemset(a, 0, n * sizeof(double)));
  TCHECK(hipMemset(b, 0, n * sizeof(double)));
  for (int r = 0; r < reps; r++) k_stream_copy<<<4096, 256>>>(a, b, n);
  TCHECK(hipDeviceSynchronize());
  TCHECK(hipFree(a));
  TCHECK(hipFree(b));
  return 1;
}

// Average microseconds per fused sweep launch on an n^3 grid; word 8 (FP64)
// or 4 (FP32); sweeps even and positive; tile 0 = the library's default.
int pmgtest_fused_sweeps(int n, int word, int sweeps, int tile, double* us_per_sweep) {
  if (n < 3 || sweeps <= 0 || (sweeps & 1) || (word != 4 && word != 8)) return 0;
  *us_per_sweep = word == 8 ? fused_sweeps_us<double>(n, sweeps, tile)
                            : fused_sweeps_us<float>(n, sweeps, tile);
  return 1;
}

int gsrb_impl(int nx, int ny, int nz, int st, const double* a, const double* cc, const double* fc,
              double* x, int nu, int iadjoint, int tile);

// Inverse of a dense N x N matrix (row-major) in the given form of
// dense_invert (0 by size, 1 without the register form)
int pmgtest_dense_invert(int N, const double* M, double* out, int form) {
  if (N <= 0 || N > 4096 || form < 0 || form > 1) return 0;
  const size_t nn = (size_t)N * N;
  double *dM = up(M, nn), *dw = up(nullptr, (size_t)N);
  pmg::dense_invert(0, dM, dw, N, form);
  TCHECK(hipDeviceSynchronize());
  down(out, dM, nn);
  TCHECK(hipFree(dM));
  TCHECK(hipFree(dw));
  return 1;
}

// Average microseconds per Galerkin call on a synthetic nf^3 level (see
// galerkin_us); form as pmgtest_galerkin_form
int pmgtest_galerkin_time(int nf, int st, int form, int reps, double* us) {
  if (nf < 5 || (nf & 1) == 0 || (st != 7 && st != 27) || reps <= 0 || form < 0 || form > 4) return 0;
  *us = galerkin_us(nf, st, form, reps);
  return 1;
}

// Average microseconds per 27-point sweep on an n^3 level (see gs27_sweeps_us)
int pmgtest_gs27_sweeps(int n, int word, int sweeps, int mode, double* us_per_sweep) {
  if (n < 3 || sweeps <= 0 || (word != 4 && word != 8)) return 0;
  *us_per_sweep = word == 8 ? gs27_sweeps_us<double>(n, sweeps, mode)
                            : gs27_sweeps_us<float>(n, sweeps, mode);
  return 1;
}

// 27-point sweeps with the kernel form forced: mode 0 one launch per colour,
// 1 per plane parity (tile by level size), 2 per colour pair (rows of up to
// 128 interior points; the vector kernel where the row allows it), 3 per
// colour pair with one point per lane, 8 / 16 / 32 per plane parity with
// that tile
int pmgtest_gsrb27_mode(int nx, int ny, int nz, const double* a, const double* cc,
                        const double* fc, double* x, int nu, int iadjoint, int mode) {
  return gsrb_impl(nx, ny, nz, 27, a, cc, fc, x, nu, iadjoint, mode);
}

// x <- nu sweeps (7-point red-black or 27-point 8-colour), D = oC + cc
int pmgtest_gsrb(int nx, int ny, int nz, int st, const double* a, const double* cc,
                 const double* fc, double* x, int nu, int iadjoint) {
  return gsrb_impl(nx, ny, nz, st, a, cc, fc, x, nu, iadjoint, 0);
}

// the same with an explicit tile shape of the fused 7-point sweep
int pmgtest_gsrb_tile(int nx, int ny, int nz, const double* a, const double* cc,
                      const double* fc, double* x, int nu, int iadjoint, int tile) {
  return gsrb_impl(nx, ny, nz, 7, a, cc, fc, x, nu, iadjoint, tile);
}

// Two fused 7-point sweeps from a zero start (the zero_x form of the V-cycle's
// first smoothing): x's boundary layer must be zero, its interior is ignored
// (the caller may fill it with anything); tile as pmgtest_gsrb_tile (0: the
// default shape).  The result is written back to x.
int pmgtest_gsrb_zero(int nx, int ny, int nz, const double* a, const double* cc,
                      const double* fc, double* x, int iadjoint, int tile) {
  const size_t n = (size_t)nx * ny * nz;
  OpsHolder A(a, n, 7);
  double *dcc = up(cc, n), *dfc = up(fc, n), *dx = up(x, n), *dD = up(nullptr, n);
  pmg::add_diag<double>(0, n, A.ops.a[0], dcc, dD);
  double* dt = up(nullptr, n);  // the partner: zero boundary layer
  pmg::fill<double>(0, dt, n, 0.0);
  const pmg::LevelArgs<double> c{A.ops, dD, dfc, dx, dt, nullptr};
  pmg::gsrb7_fused_batch<double>(0, nx, ny, nz, &c, 1, 2, iadjoint, tile, true);
  TCHECK(hipDeviceSynchronize());
  down(x, dx, n);
  for (double* p : {dcc, dfc, dx, dD, dt}) TCHECK(hipFree(p));
  return 1;
}

}  // extern "C"

int gsrb_impl(int nx, int ny, int nz, int st, const double* a, const double* cc, const double* fc,
              double* x, int nu, int iadjoint, int tile) {
  const size_t n = (size_t)nx * ny * nz;
  OpsHolder A(a, n, st);
  double *dcc = up(cc, n), *dfc = up(fc, n), *dx = up(x, n), *dD = up(nullptr, n);
  pmg::add_diag<double>(0, n, A.ops.a[0], dcc, dD);
  if (st == 7) {
    // the fused sweep kernel of the V-cycle; its partner buffer carries the
    // boundary of x
    double* dt = up(x, n);
    pmg::gsrb7_fused<double>(0, nx, ny, nz, A.ops, dD, dfc, dx, dt, nu, iadjoint, tile);
    TCHECK(hipDeviceSynchronize());
    TCHECK(hipFree(dt));
  } else if (tile == 0) {
    pmg::gsrb27<double>(0, nx, ny, nz, A.ops, dD, dfc, dx, nu, iadjoint);
  } else if (tile == 2 || tile == 3) {
    if (nx > 130) return 0;
    const pmg::LevelArgs<double> c{A.ops, dD, dfc, dx, nullptr, nullptr};
    pmg::gsrb27_pairs_form<double>(0, nx, ny, nz, &c, 1, nu, iadjoint, tile - 2);
  } else {
    // the plane-parity sweeps; the partner buffer carries the boundary of x
    double* dt = up(x, n);
    pmg::gsrb27_planes<double>(0, nx, ny, nz, A.ops, dD, dfc, dx, dt, nu, iadjoint,
                               tile == 1 ? 0 : tile);
    TCHECK(hipDeviceSynchronize());
    TCHECK(hipFree(dt));
  }
  TCHECK(hipDeviceSynchronize());
  down(x, dx, n);
  TCHECK(hipFree(dcc));
  TCHECK(hipFree(dfc));
  TCHECK(hipFree(dx));
  TCHECK(hipFree(dD));
  return 1;
}

extern "C" {

int pmgtest_resid(int nx, int ny, int nz, int st, const double* a, const double* cc,
                  const double* fc, const double* x, double* r) {
  const size_t n = (size_t)nx * ny * nz;
  OpsHolder A(a, n, st);
  double *dcc = up(cc, n), *dfc = up(fc, n), *dx = up(x, n), *dD = up(nullptr, n), *dr = up(r, n);
  pmg::add_diag<double>(0, n, A.ops.a[0], dcc, dD);
  pmg::resid<double>(0, st, nx, ny, nz, A.ops, dD, dfc, dx, dr);
  TCHECK(hipDeviceSynchronize());
  down(r, dr, n);
  for (double* p : {dcc, dfc, dx, dD, dr}) TCHECK(hipFree(p));
  return 1;
}

int pmgtest_restrict(int nxf, int nyf, int nzf, int nxc, int nyc, int nzc, const double* in,
                     double* out) {
  const size_t nf = (size_t)nxf * nyf * nzf, nc = (size_t)nxc * nyc * nzc;
  double *di = up(in, nf), *dout = up(out, nc);
  pmg::restrict3<double>(0, nxf, nyf, nzf, nxc, nyc, nzc, di, dout);
  TCHECK(hipDeviceSynchronize());
  down(out, dout, nc);
  TCHECK(hipFree(di));
  TCHECK(hipFree(dout));
  return 1;
}

// xf_out = xf + damp * P xc on the fine interior
int pmgtest_prolong(int nxc, int nyc, int nzc, int nxf, int nyf, int nzf, const double* xc,
                    double* xf, double damp) {
  const size_t nf = (size_t)nxf * nyf * nzf, nc = (size_t)nxc * nyc * nzc;
  double *dc = up(xc, nc), *df = up(xf, nf), *dd = up(&damp, 1);
  pmg::prolong_add<double>(0, nxc, nyc, nzc, nxf, nyf, nzf, dc, df, dd);
  TCHECK(hipDeviceSynchronize());
  down(xf, df, nf);
  for (double* p : {dc, df, dd}) TCHECK(hipFree(p));
  return 1;
}

int pmgtest_galerkin_form(int nxf, int nyf, int nzf, int st, const double* af, int nxc, int nyc,
                          int nzc, double* ac, int form) {
  if (form < 0 || form > 4) return 0;
  const size_t nf = (size_t)nxf * nyf * nzf, nc = (size_t)nxc * nyc * nzc;
  OpsHolder A(af, nf, st);
  double* out[14];
  for (int s = 0; s < 14; s++) {
    out[s] = up(nullptr, nc);
    TCHECK(hipMemset(out[s], 0, nc * sizeof(double)));
  }
  pmg::galerkin(0, st, nxf, nyf, nzf, A.ops, nxc, nyc, nzc, out, form);
  TCHECK(hipDeviceSynchronize());
  for (int s = 0; s < 14; s++) {
    down(ac + (size_t)s * nc, out[s], nc);
    TCHECK(hipFree(out[s]));
  }
  return 1;
}

int pmgtest_galerkin(int nxf, int nyf, int nzf, int st, const double* af, int nxc, int nyc,
                     int nzc, double* ac) {
  return pmgtest_galerkin_form(nxf, nyf, nzf, st, af, nxc, nyc, nzc, ac, 0);
}

int pmgtest_buildA(int nx, int ny, int nz, const double* xf, const double* yf, const double* zf,
                   const double* gx, const double* gy, const double* gz, const double* a1,
                   const double* a2, const double* a3, const double* ccf, const double* fcf,
                   double* oC, double* cc, double* fc, double* oE, double* oN, double* uC) {
  const size_t n = (size_t)nx * ny * nz;
  double *dxf = up(xf, nx), *dyf = up(yf, ny), *dzf = up(zf, nz);
  double *dgx = up(gx, 2 * (size_t)ny * nz), *dgy = up(gy, 2 * (size_t)nx * nz),
         *dgz = up(gz, 2 * (size_t)nx * ny);
  double *d1 = up(a1, n), *d2 = up(a2, n), *d3 = up(a3, n), *dccf = up(ccf, n), *dfcf = up(fcf, n);
  double *doC = up(oC, n), *dcc = up(cc, n), *dfc = up(fc, n), *doE = up(oE, n), *doN = up(oN, n),
         *duC = up(uC, n);
  pmg::buildA_fv(0, nx, ny, nz, dxf, dyf, dzf, dgx, dgy, dgz, d1, d2, d3, dccf, dfcf, doC, dcc, dfc,
                 doE, doN, duC);
  TCHECK(hipDeviceSynchronize());
  down(oC, doC, n);
  down(cc, dcc, n);
  down(fc, dfc, n);
  down(oE, doE, n);
  down(oN, doN, n);
  down(uC, duC, n);
  for (double* p : {dxf, dyf, dzf, dgx, dgy, dgz, d1, d2, d3, dccf, dfcf, doC, dcc, dfc, doE, doN, duC})
    TCHECK(hipFree(p));
  return 1;
}

// nonlinear term pieces: out_r = nonlinear residual (7-point), out_d = N'(x)
int pmgtest_nonlinear(int nx, int ny, int nz, const double* a, const double* cc, const double* fc,
                      const double* x, int nion, const double* q, const double* sconc,
                      double* out_r, double* out_d) {
  const size_t n = (size_t)nx * ny * nz;
  OpsHolder A(a, n, 7);
  pmg::Ions I{};
  I.nion = nion;
  for (int i = 0; i < nion; i++) {
    I.q[i] = q[i];
    I.sconc[i] = sconc[i];
  }
  double *dcc = up(cc, n), *dfc = up(fc, n), *dx = up(x, n), *dr = up(out_r, n), *dd = up(out_d, n);
  pmg::nresid7(0, nx, ny, nz, A.ops, dcc, dfc, dx, dr, I);
  pmg::dcvec(0, n, dcc, dx, dd, I);
  TCHECK(hipDeviceSynchronize());
  down(out_r, dr, n);
  down(out_d, dd, n);
  for (double* p : {dcc, dfc, dx, dr, dd}) TCHECK(hipFree(p));
  return 1;
}

// Fused FP32 7-point sweeps (nu of them, or the zero-start pair when zero is
// set, x's interior then ignored) on the given operator arrays: oE, oN, uC,
// D, fc and x are nx*ny*nz floats; tile as pmgtest_gsrb_tile.  The partner
// buffer carries x's boundary.  Result in x.
int pmgtest_fused32(int nx, int ny, int nz, const float* oE, const float* oN, const float* uC,
                    const float* D, const float* fc, float* x, int nu, int iadjoint, int tile,
                    int zero) {
  if (nx < 3 || ny < 3 || nz < 3 || (!zero && nu < 1)) return 0;
  const size_t n = (size_t)nx * ny * nz;
  auto upf = [&](const float* h) {
    float* d = nullptr;
    TCHECK(hipMalloc(&d, n * sizeof(float)));
    TCHECK(hipMemcpy(d, h, n * sizeof(float), hipMemcpyHostToDevice));
    return d;
  };
  float *dE = upf(oE), *dN = upf(oN), *dU = upf(uC), *dD = upf(D), *df = upf(fc), *dx = upf(x),
        *dt = upf(x);
  pmg::Ops<float> ops{};
  ops.a[1] = dE;
  ops.a[2] = dN;
  ops.a[3] = dU;
  const pmg::LevelArgs<float> c{ops, dD, df, dx, dt, nullptr};
  pmg::gsrb7_fused_batch<float>(0, nx, ny, nz, &c, 1, zero ? 2 : nu, iadjoint, tile, zero != 0);
  TCHECK(hipDeviceSynchronize());
  TCHECK(hipMemcpy(x, dx, n * sizeof(float), hipMemcpyDeviceToHost));
  for (float* q : {dE, dN, dU, dD, df, dx, dt}) TCHECK(hipFree(q));
  return 1;
}

// FP32 27-point sweeps in a given form on explicit arrays (see the header)
int pmgtest_gsrb27_32(int nx, int ny, int nz, const float* a, const float* D, const float* fc,
                      float* x, int nu, int iadjoint, int mode) {
  if (nx < 3 || ny < 3 || nz < 3 || nu < 1) return 0;
  if (mode >= 2 && mode <= 4 && nx > 130) return 0;
  const size_t n = (size_t)nx * ny * nz;
  auto upf = [&](const float* h) {
    float* d = nullptr;
    TCHECK(hipMalloc(&d, n * sizeof(float)));
    TCHECK(hipMemcpy(d, h, n * sizeof(float), hipMemcpyHostToDevice));
    return d;
  };
  pmg::Ops<float> ops{};
  for (int s = 0; s < 14; s++) ops.a[s] = upf(a + (size_t)s * n);
  float *dD = upf(D), *df = upf(fc), *dx = upf(x), *dt = upf(x);
  const pmg::LevelArgs<float> c{ops, dD, df, dx, dt, nullptr};
  if (mode == 0)
    pmg::gsrb27_batch<float>(0, nx, ny, nz, &c, 1, nu, iadjoint);
  else if (mode >= 2 && mode <= 4)
    pmg::gsrb27_pairs_form<float>(0, nx, ny, nz, &c, 1, nu, iadjoint, mode - 2);
  else
    pmg::gsrb27_planes_batch<float>(0, nx, ny, nz, &c, 1, nu, iadjoint, mode == 1 ? 0 : mode);
  TCHECK(hipDeviceSynchronize());
  TCHECK(hipMemcpy(x, dx, n * sizeof(float), hipMemcpyDeviceToHost));
  for (int s = 0; s < 14; s++) TCHECK(hipFree((void*)ops.a[s]));
  for (float* q : {dD, df, dx, dt}) TCHECK(hipFree(q));
  return 1;
}

// FP32 7-point residual r = f + sum(a x) - D x on explicit arrays (see the
// header); r's boundary is left as the caller passed it
int pmgtest_resid32(int nx, int ny, int nz, const float* oE, const float* oN, const float* uC,
                    const float* D, const float* fc, const float* x, float* r) {
  if (nx < 3 || ny < 3 || nz < 3) return 0;
  const size_t n = (size_t)nx * ny * nz;
  auto upf = [&](const float* h) {
    float* d = nullptr;
    TCHECK(hipMalloc(&d, n * sizeof(float)));
    TCHECK(hipMemcpy(d, h, n * sizeof(float), hipMemcpyHostToDevice));
    return d;
  };
  float *dE = upf(oE), *dN = upf(oN), *dU = upf(uC), *dD = upf(D), *df = upf(fc), *dx = upf(x),
        *dr = upf(r);
  pmg::Ops<float> ops{};
  ops.a[1] = dE;
  ops.a[2] = dN;
  ops.a[3] = dU;
  pmg::resid<float>(0, 7, nx, ny, nz, ops, dD, df, dx, dr);
  TCHECK(hipDeviceSynchronize());
  TCHECK(hipMemcpy(r, dr, n * sizeof(float), hipMemcpyDeviceToHost));
  for (float* q : {dE, dN, dU, dD, df, dx, dr}) TCHECK(hipFree(q));
  return 1;
}

}  // extern "C"
