// fill_device.hip -- coefficient maps, boundary values, energies and forces.
//
// Order-sensitive maps (spline charges, the multiplicative spline dielectric
// and kappa windows) are gathers: each grid point visits the atoms that reach
// it in ascending atom index, which is the order in which the reference's
// per-atom scatter loops (src/mg/vpmg.c:4631-4854, 4986-5248) update that
// point.  The result is deterministic and rounds like the reference.
#include "fill_device.h"

#include <cmath>
#include <cstdio>
#include <mutex>
#include <vector>

namespace pmg {
namespace {

constexpr int kB = 256;
constexpr int kParts = 1024;
static_assert(kScratchParts == kParts, "the static buffer holds one partial per reduction block");
constexpr double VSMALL = 1.0e-9;
constexpr double VPMGSMALL = 1e-12;

__host__ __device__ inline size_t IX(int i, int j, int k, int nx, int ny) {
  return ((size_t)k * ny + j) * nx + i;
}

inline unsigned nb(size_t n) {
  size_t b = (n + kB - 1) / kB;
  return (unsigned)(b == 0 ? 1 : b);
}

// ----------------------------------------------------------------- splines
__device__ inline double bspline2(double x) {  // vpmg.c:5091-5104
  double m2m, m2, m3;
  if ((x >= 0.0) && (x <= 2.0)) m2m = 1.0 - fabs(x - 1.0); else m2m = 0.0;
  if ((x >= 1.0) && (x <= 3.0)) m2 = 1.0 - fabs(x - 2.0); else m2 = 0.0;
  if ((x >= 0.0) && (x <= 3.0)) m3 = 0.5 * x * m2m + 0.5 * (3.0 - x) * m2; else m3 = 0.0;
  return m3;
}

__device__ inline double dbspline2(double x) {  // vpmg.c:5106-5121
  double m2m, m2;
  if ((x >= 0.0) && (x <= 2.0)) m2m = 1.0 - fabs(x - 1.0); else m2m = 0.0;
  if ((x >= 1.0) && (x <= 3.0)) m2 = 1.0 - fabs(x - 2.0); else m2 = 0.0;
  return m2m - m2;
}

__device__ inline double vfchi(int iint, double iflt) { return 1.5 + ((double)(iint) - (iflt)); }

// The 5th-order B-spline of chgm spl4 and its first two derivatives
// (bspline4 / dbspline4 / d2bspline4, vpmg.c:6731-6822), argument
// VFCHI4(i, f) = 2.5 + (i - f) (vpmg.h:939).
__device__ inline double vfchi4(int iint, double iflt) { return 2.5 + ((double)(iint) - (iflt)); }

__device__ inline double bspline4(double x) {
  const double one6 = 1.0 / 6.0, one8 = 1.0 / 8.0, one24 = 1.0 / 24.0;
  const double thirteen24 = 13.0 / 24.0, fourtyseven24 = 47.0 / 24.0, seventeen24 = 17.0 / 24.0;
  double m, m2;
  if ((x > 0.0) && (x <= 1.0)) {
    m = x * x;
    return one24 * m * m;
  } else if ((x > 1.0) && (x <= 2.0)) {
    m = x - 1.0;
    m2 = m * m;
    return -one8 + one6 * x + m2 * (0.25 + one6 * m - one6 * m2);
  } else if ((x > 2.0) && (x <= 3.0)) {
    m = x - 2.0;
    m2 = m * m;
    return -thirteen24 + 0.5 * x + m2 * (-0.25 - 0.5 * m + 0.25 * m2);
  } else if ((x > 3.0) && (x <= 4.0)) {
    m = x - 3.0;
    m2 = m * m;
    return fourtyseven24 - 0.5 * x + m2 * (-0.25 + 0.5 * m - one6 * m2);
  } else if ((x > 4.0) && (x <= 5.0)) {
    m = x - 4.0;
    m2 = m * m;
    return seventeen24 - one6 * x + m2 * (0.25 - one6 * m + one24 * m2);
  }
  return 0.0;
}

__device__ inline double dbspline4(double x) {
  const double one6 = 1.0 / 6.0, one3 = 1.0 / 3.0, two3 = 2.0 / 3.0, thirteen6 = 13.0 / 6.0;
  double m, m2;
  if ((x > 0.0) && (x <= 1.0)) {
    m2 = x * x;
    return one6 * x * m2;
  } else if ((x > 1.0) && (x <= 2.0)) {
    m = x - 1.0;
    m2 = m * m;
    return -one3 + 0.5 * x + m2 * (0.5 - two3 * m);
  } else if ((x > 2.0) && (x <= 3.0)) {
    m = x - 2.0;
    m2 = m * m;
    return 1.5 - 0.5 * x + m2 * (-1.5 + m);
  } else if ((x > 3.0) && (x <= 4.0)) {
    m = x - 3.0;
    m2 = m * m;
    return 1.0 - 0.5 * x + m2 * (1.5 - two3 * m);
  } else if ((x > 4.0) && (x <= 5.0)) {
    m = x - 4.0;
    m2 = m * m;
    return -thirteen6 + 0.5 * x + m2 * (-0.5 + one6 * m);
  }
  return 0.0;
}

__device__ inline double d2bspline4(double x) {
  double m, m2;
  if ((x > 0.0) && (x <= 1.0)) {
    return 0.5 * x * x;
  } else if ((x > 1.0) && (x <= 2.0)) {
    m = x - 1.0;
    m2 = m * m;
    return -0.5 + x - 2.0 * m2;
  } else if ((x > 2.0) && (x <= 3.0)) {
    m = x - 2.0;
    m2 = m * m;
    return 5.5 - 3.0 * x + 3.0 * m2;
  } else if ((x > 3.0) && (x <= 4.0)) {
    m = x - 3.0;
    m2 = m * m;
    return -9.5 + 3.0 * x - 2.0 * m2;
  } else if ((x > 4.0) && (x <= 5.0)) {
    m = x - 4.0;
    m2 = m * m;
    return 4.5 - x + 0.5 * m2;
  }
  return 0.0;
}

__device__ inline int imin2(int a, int b) { return a <= b ? a : b; }
__device__ inline int imax2(int a, int b) { return a >= b ? a : b; }

// ----------------------------------------------------------------- charges
// Point r (0 or 1) of thread tid in tile `tile`: a workgroup of kB = 256
// threads covers the tile's 8 x 8 x 8 points, two planes of 4 apart per thread.
static_assert(kTile == 8 && kB == 256, "tile gathers: 2 points per thread");
__device__ inline void tile_point(const Tiles& T, int tile, int tid, int r, int& i, int& j,
                                  int& k) {
  const int tx = tile % T.tdim[0], ty = (tile / T.tdim[0]) % T.tdim[1],
            tz = tile / (T.tdim[0] * T.tdim[1]);
  i = tx * kTile + tid % kTile;
  j = ty * kTile + (tid / kTile) % kTile;
  k = tz * kTile + tid / (kTile * kTile) + 4 * r;
}

// One workgroup per tile: every point sums the contributions of the tile's
// atoms, in ascending atom index (the order of the reference's per-atom loop).
__global__ void __launch_bounds__(kB)
    k_charge(Geo g, AtomsDev A, Tiles T, int mode, double zmagic, double* charge) {
  const int tile = blockIdx.x;
  const int q0 = T.start[tile], q1 = T.start[tile + 1];
  for (int r = 0; r < 2; r++) {
  int i, j, k;
  tile_point(T, tile, threadIdx.x, r, i, j, k);
  if (i >= g.nx || j >= g.ny || k >= g.nz) continue;
  double sum = 0.0;
  for (int q = q0; q < q1; q++) {
    const int a = T.list[q];
    const double* ap = A.pos + 3 * a;
    const double fi = (ap[0] - g.xmin) / g.hx, fj = (ap[1] - g.ymin) / g.hy,
                 fk = (ap[2] - g.zmin) / g.hz;
    const double c = A.chg[a] * zmagic / (g.hx * g.hy * g.hz);
    if (mode == 2) {
      // chgm spl4: fillcoPermanentMultipole (vpmg.c:6835-7050) with the atom's
      // charge only; its dipole and quadrupole terms are exact zeros here and
      // are kept so each contribution rounds as the reference's expression
      if ((ap[0] <= (g.xmin - 2 * g.hx)) || (ap[0] >= (g.xmax + 2 * g.hx)) ||
          (ap[1] <= (g.ymin - 2 * g.hy)) || (ap[1] >= (g.ymax + 2 * g.hy)) ||
          (ap[2] <= (g.zmin - 2 * g.hz)) || (ap[2] >= (g.zmax + 2 * g.hz)))
        continue;
      const int ip2 = imin2((int)ceil(fi) + 2, g.nx - 1), im2 = imax2((int)floor(fi) - 2, 0);
      const int jp2 = imin2((int)ceil(fj) + 2, g.ny - 1), jm2 = imax2((int)floor(fj) - 2, 0);
      const int kp2 = imin2((int)ceil(fk) + 2, g.nz - 1), km2 = imax2((int)floor(fk) - 2, 0);
      if (i < im2 || i > ip2 || j < jm2 || j > jp2 || k < km2 || k > kp2) continue;
      const double f = zmagic / (g.hx * g.hy * g.hz);
      const double c4 = A.chg[a] * f;
      const double zero = 0.0;
      const double mi = vfchi4(i, fi), mj = vfchi4(j, fj), mk = vfchi4(k, fk);
      const double mx = bspline4(mi), dmx = dbspline4(mi), d2mx = d2bspline4(mi);
      const double my = bspline4(mj), dmy = dbspline4(mj), d2my = d2bspline4(mj);
      const double mz = bspline4(mk), dmz = dbspline4(mk), d2mz = d2bspline4(mk);
      const double v = mx * my * mz * c4 - dmx * my * mz * zero - mx * dmy * mz * zero -
                       mx * my * dmz * zero + d2mx * my * mz * zero + dmx * dmy * mz * zero +
                       mx * d2my * mz * zero + dmx * my * dmz * zero + mx * dmy * dmz * zero +
                       mx * my * d2mz * zero;
      sum += v;
    } else if (mode == 1) {
      if ((ap[0] <= (g.xmin - g.hx)) || (ap[0] >= (g.xmax + g.hx)) || (ap[1] <= (g.ymin - g.hy)) ||
          (ap[1] >= (g.ymax + g.hy)) || (ap[2] <= (g.zmin - g.hz)) || (ap[2] >= (g.zmax + g.hz)))
        continue;
      const int ip2 = imin2((int)ceil(fi) + 1, g.nx - 1), im2 = imax2((int)floor(fi) - 1, 0);
      const int jp2 = imin2((int)ceil(fj) + 1, g.ny - 1), jm2 = imax2((int)floor(fj) - 1, 0);
      const int kp2 = imin2((int)ceil(fk) + 1, g.nz - 1), km2 = imax2((int)floor(fk) - 1, 0);
      if (i < im2 || i > ip2 || j < jm2 || j > jp2 || k < km2 || k > kp2) continue;
      const double mx = bspline2(vfchi(i, fi)), my = bspline2(vfchi(j, fj)),
                   mz = bspline2(vfchi(k, fk));
      sum += (c * mx * my * mz);
    } else {
      if ((ap[0] <= g.xmin) || (ap[0] >= g.xmax) || (ap[1] <= g.ymin) || (ap[1] >= g.ymax) ||
          (ap[2] <= g.zmin) || (ap[2] >= g.zmax))
        continue;
      const int ihi = (int)ceil(fi), ilo = (int)floor(fi), jhi = (int)ceil(fj),
                jlo = (int)floor(fj), khi = (int)ceil(fk), klo = (int)floor(fk);
      const double dx = fi - (double)(ilo), dy = fj - (double)(jlo), dz = fk - (double)(klo);
      // the eight statements of vpmg.c:5078-5085, in order
      if (i == ihi && j == jhi && k == khi) sum += (dx * dy * dz * c);
      if (i == ihi && j == jlo && k == khi) sum += (dx * (1.0 - dy) * dz * c);
      if (i == ihi && j == jhi && k == klo) sum += (dx * dy * (1.0 - dz) * c);
      if (i == ihi && j == jlo && k == klo) sum += (dx * (1.0 - dy) * (1.0 - dz) * c);
      if (i == ilo && j == jhi && k == khi) sum += ((1.0 - dx) * dy * dz * c);
      if (i == ilo && j == jlo && k == khi) sum += ((1.0 - dx) * (1.0 - dy) * dz * c);
      if (i == ilo && j == jhi && k == klo) sum += ((1.0 - dx) * dy * (1.0 - dz) * c);
      if (i == ilo && j == jlo && k == klo) sum += ((1.0 - dx) * (1.0 - dy) * (1.0 - dz) * c);
    }
  }
  charge[IX(i, j, k, g.nx, g.ny)] = sum;
  }
}

// ------------------------------------------------------------ spline coefs
// Polynomial windows of srfm spl3 / spl4 (fillcoCoefSpline3 vpmg.c:7525,
// fillcoCoefSpline4 vpmg.c:7052): chi(s) = sum_n c_n s^n of the distance s,
// clamped to [0, 1], with s^2 taken as given (the reference squares s in the
// x and y branches and reuses the squared distance in the kappa and z ones).
// The per-atom coefficients come from the host (kSplCoef doubles per atom:
// dielectric, ion for the maps, ion for the ion-boundary force).
constexpr int kSplCoef = 24;

template <int kOrd>
__device__ inline double poly_chi(const double* c, double sm, double sm2) {
  const double sm3 = sm2 * sm, sm4 = sm3 * sm, sm5 = sm4 * sm;
  double v;
  if (kOrd == 3) {
    v = c[0] + c[1] * sm + c[2] * sm2 + c[3] * sm3 + c[4] * sm4 + c[5] * sm5;
  } else {
    const double sm6 = sm5 * sm, sm7 = sm6 * sm;
    v = c[0] + c[1] * sm + c[2] * sm2 + c[3] * sm3 + c[4] * sm4 + c[5] * sm5 + c[6] * sm6 +
        c[7] * sm7;
  }
  if (v > 1.0) {
    v = 1.0;
  } else if (v < 0.0) {
    v = 0.0;
  }
  return v;
}

template <int kOrd>
__global__ void __launch_bounds__(kB)
    k_coef_spline(Geo g, AtomsDev A, Tiles T, double win, double irad, double ionmask, double epsw,
                  double epsp, const double* pc, double* kappa, double* epsx, double* epsy,
                  double* epsz) {
  const int tile = blockIdx.x;
  const int q0 = T.start[tile], q1 = T.start[tile + 1];
  const double w2i = 1.0 / (win * win), w3i = 1.0 / (win * win * win);
  const double hx = g.hx, hy = g.hy, hzed = g.hz;
  for (int r = 0; r < 2; r++) {
  int i, j, k;
  tile_point(T, tile, threadIdx.x, r, i, j, k);
  if (i >= g.nx || j >= g.ny || k >= g.nz) continue;
  double ka = 1.0, ex = 1.0, ey = 1.0, ez = 1.0;
  for (int q = q0; q < q1; q++) {
    const int a = T.list[q];
    const double* ap = A.pos + 3 * a;
    const double arad = A.rad[a];
    if ((ap[0] <= g.xmin) || (ap[0] >= g.xmax) || (ap[1] <= g.ymin) || (ap[1] >= g.ymax) ||
        (ap[2] <= g.zmin) || (ap[2] >= g.zmax))
      continue;
    if (!(arad > VPMGSMALL)) continue;
    const double p0 = ap[0] - g.xmin, p1 = ap[1] - g.ymin, p2 = ap[2] - g.zmin;
    const double itot = irad + arad + win, itot2 = itot * itot;
    const double ictot = fmax(0.0, (irad + arad - win)), ictot2 = ictot * ictot;
    const double stot = arad + win, stot2 = stot * stot;
    const double sctot = fmax(0.0, (arad - win)), sctot2 = sctot * sctot;
    const double rtot = itot >= stot ? itot : stot;
    const double dx = rtot + 0.5 * hx, dy = rtot + 0.5 * hy, dz = rtot + 0.5 * hzed;
    const int imin = imax2(0, (int)floor((p0 - dx) / hx)), imax = imin2(g.nx - 1, (int)ceil((p0 + dx) / hx));
    const int jmin = imax2(0, (int)floor((p1 - dy) / hy)), jmax = imin2(g.ny - 1, (int)ceil((p1 + dy) / hy));
    const int kmin = imax2(0, (int)floor((p2 - dz) / hzed)), kmax = imin2(g.nz - 1, (int)ceil((p2 + dz) / hzed));
    if (i < imin || i > imax || j < jmin || j > jmax || k < kmin || k > kmax) continue;
    const double dx2 = (p0 - hx * i) * (p0 - hx * i);
    const double dy2 = (p1 - hy * j) * (p1 - hy * j);
    const double dz2 = (p2 - k * hzed) * (p2 - k * hzed);
    if (ka > VPMGSMALL) {
      const double d2 = dz2 + dy2 + dx2;
      if (d2 <= ictot2) ka = 0.0;
      if ((d2 < itot2) && (d2 > ictot2)) {
        if (kOrd == 2) {
          const double dist = sqrt(d2), sm = dist - (arad + irad) + win, sm2 = sm * sm;
          ka *= 0.75 * sm2 * w2i - 0.25 * sm * sm2 * w3i;
        } else {
          const double dist = sqrt(d2);
          ka *= poly_chi<kOrd>(pc + kSplCoef * a + 8, dist, d2);
        }
      }
    }
    if (ex > VPMGSMALL) {
      const double q = p0 - (i + 0.5) * hx;
      const double d2 = dz2 + dy2 + q * q;
      if (d2 >= stot2) ex *= 1.0;
      if (d2 <= sctot2) ex = 0.0;
      if ((d2 > sctot2) && (d2 < stot2)) {
        if (kOrd == 2) {
          const double dist = sqrt(d2), sm = dist - arad + win, sm2 = sm * sm;
          ex *= 0.75 * sm2 * w2i - 0.25 * sm * sm2 * w3i;
        } else {
          const double sm = sqrt(d2);
          ex *= poly_chi<kOrd>(pc + kSplCoef * a, sm, sm * sm);
        }
      }
    }
    if (ey > VPMGSMALL) {
      const double q = p1 - (j + 0.5) * hy;
      const double d2 = dz2 + dx2 + q * q;
      if (d2 >= stot2) ey *= 1.0;
      if (d2 <= sctot2) ey = 0.0;
      if ((d2 > sctot2) && (d2 < stot2)) {
        if (kOrd == 2) {
          const double dist = sqrt(d2), sm = dist - arad + win, sm2 = sm * sm;
          ey *= 0.75 * sm2 * w2i - 0.25 * sm * sm2 * w3i;
        } else {
          const double sm = sqrt(d2);
          ey *= poly_chi<kOrd>(pc + kSplCoef * a, sm, sm * sm);
        }
      }
    }
    if (ez > VPMGSMALL) {
      const double q = p2 - (k + 0.5) * hzed;
      const double d2 = dy2 + dx2 + q * q;
      if (d2 >= stot2) ez *= 1.0;
      if (d2 <= sctot2) ez = 0.0;
      if ((d2 > sctot2) && (d2 < stot2)) {
        if (kOrd == 2) {
          const double dist = sqrt(d2), sm = dist - arad + win, sm2 = sm * sm;
          ez *= 0.75 * sm2 * w2i - 0.25 * sm * sm2 * w3i;
        } else {
          const double sm = sqrt(d2);
          ez *= poly_chi<kOrd>(pc + kSplCoef * a, sm, d2);
        }
      }
    }
  }
  const size_t p = IX(i, j, k, g.nx, g.ny);
  kappa[p] = ionmask * ka;
  epsx[p] = (epsw - epsp) * ex + epsp;
  epsy[p] = (epsw - epsp) * ey + epsp;
  epsz[p] = (epsw - epsp) * ez + epsp;
  }
}

// ------------------------------------------------------------ SAS / markSphere
// Accessibility of the SAS points of one atom (Vacc_atomSurf with
// ivdwAccExclus, vacc.c:78-128, 866-922): a point is kept unless another atom
// b has |pt - b|^2 < (r_b + probe)^2.  One workgroup per atom: the atoms that
// can exclude any point of this atom's sphere (|a - b| < r_a + r_b + 2 probe,
// plus a margin) are gathered once into LDS and ordered nearest first, then
// every point is tested against that list until one excludes it (most points
// are buried, and the nearest atoms exclude them first).  The test is a
// boolean "any", so the list order does not change a flag; the arithmetic of
// each test is the oracle's.  A list longer than kSasCap falls back to the
// cell list in global memory.
constexpr int kSasCap = 256;

__device__ inline bool sas_free_global(const AtomsDev& A, const Cells& C, int self, double px,
                                       double py, double pz, double probe) {
  const int cx = (int)floor((px - C.lo[0]) / C.cs);
  const int cy = (int)floor((py - C.lo[1]) / C.cs);
  const int cz = (int)floor((pz - C.lo[2]) / C.cs);
  for (int dz = -1; dz <= 1; dz++)
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++) {
        const int a = cx + dx, b = cy + dy, c = cz + dz;
        if (a < 0 || b < 0 || c < 0 || a >= C.dim[0] || b >= C.dim[1] || c >= C.dim[2]) continue;
        const int id = a + C.dim[0] * (b + C.dim[1] * c);
        for (int q = C.start[id]; q < C.start[id + 1]; q++) {
          const int n = C.list[q];
          if (n == self) continue;
          const double* bp = A.pos + 3 * n;
          const double d2 = (px - bp[0]) * (px - bp[0]) + (py - bp[1]) * (py - bp[1]) +
                            (pz - bp[2]) * (pz - bp[2]);
          const double rr = A.rad[n] + probe;
          if (d2 < rr * rr) return false;
        }
      }
  return true;
}

__global__ void __launch_bounds__(kB)
    k_sas_flags(AtomsDev A, Cells C, const double* ref, int nref, double probe,
                unsigned char* flags) {
  __shared__ double sx[kSasCap], sy[kSasCap], sz[kSasCap], sr2[kSasCap], sd[kSasCap];
  __shared__ double tx[kSasCap], ty[kSasCap], tz[kSasCap], tr2[kSasCap];
  __shared__ int cnt;
  const int a = blockIdx.x;
  const int tid = threadIdx.x;
  unsigned char* fl = flags + (size_t)a * nref;
  const double arad = A.rad[a];
  if (arad < VSMALL) {
    for (int q = tid; q < nref; q += kB) fl[q] = 0;
    return;
  }
  if (tid == 0) cnt = 0;
  __syncthreads();
  const double* ap = A.pos + 3 * a;
  const double ax = ap[0], ay = ap[1], az = ap[2];
  const int cx = (int)floor((ax - C.lo[0]) / C.cs);
  const int cy = (int)floor((ay - C.lo[1]) / C.cs);
  const int cz = (int)floor((az - C.lo[2]) / C.cs);
  // 125 cells (+-2 around the atom's cell), one per thread
  if (tid < 125) {
    const int x = cx + tid % 5 - 2, y = cy + (tid / 5) % 5 - 2, z = cz + tid / 25 - 2;
    if (x >= 0 && y >= 0 && z >= 0 && x < C.dim[0] && y < C.dim[1] && z < C.dim[2]) {
      const int id = x + C.dim[0] * (y + C.dim[1] * z);
      for (int q = C.start[id]; q < C.start[id + 1]; q++) {
        const int b = C.list[q];
        if (b == a) continue;
        const double* bp = A.pos + 3 * b;
        const double rr = A.rad[b] + probe;
        const double lim = arad + probe + rr + 1e-6;
        const double d2 = (ax - bp[0]) * (ax - bp[0]) + (ay - bp[1]) * (ay - bp[1]) +
                          (az - bp[2]) * (az - bp[2]);
        if (d2 < lim * lim) {
          const int slot = atomicAdd(&cnt, 1);
          if (slot < kSasCap) {
            sx[slot] = bp[0];
            sy[slot] = bp[1];
            sz[slot] = bp[2];
            sr2[slot] = rr * rr;
            sd[slot] = d2;
          }
        }
      }
    }
  }
  __syncthreads();
  const int n = cnt;
  if (n <= kSasCap) {
    // nearest first: entry e goes to its rank by (distance, slot)
    for (int e = tid; e < n; e += kB) {
      const double de = sd[e];
      int r = 0;
      for (int f = 0; f < n; f++) r += (sd[f] < de || (sd[f] == de && f < e)) ? 1 : 0;
      tx[r] = sx[e];
      ty[r] = sy[e];
      tz[r] = sz[e];
      tr2[r] = sr2[e];
    }
    __syncthreads();
  }
  const double rad = arad + probe;
  for (int q = tid; q < nref; q += kB) {
    const double px = rad * (ref[3 * q]) + ax;
    const double py = rad * (ref[3 * q + 1]) + ay;
    const double pz = rad * (ref[3 * q + 2]) + az;
    unsigned char ok = 1;
    if (n <= kSasCap) {
      for (int b = 0; b < n; b++) {
        const double d2 = (px - tx[b]) * (px - tx[b]) + (py - ty[b]) * (py - ty[b]) +
                          (pz - tz[b]) * (pz - tz[b]);
        if (d2 < tr2[b]) {
          ok = 0;
          break;
        }
      }
    } else {
      ok = sas_free_global(A, C, a, px, py, pz, probe) ? 1 : 0;
    }
    fl[q] = ok;
  }
}

// Accessible SAS points, written in flag order without atomics: kSasParts
// blocks own contiguous ranges of the flags; pass 1 counts each range, the
// host turns the counts into offsets, pass 2 writes each range's points at
// its offset (wavefront ballot prefix within the block).
constexpr int kSasParts = 1024;

__device__ inline void sas_range(long n, long& b0, long& b1) {
  const long chunk = (n + kSasParts - 1) / kSasParts;
  b0 = (long)blockIdx.x * chunk;
  b1 = b0 + chunk < n ? b0 + chunk : n;
}

__global__ void __launch_bounds__(kB)
    k_sas_count(const unsigned char* flags, long n, int* counts) {
  __shared__ int sh[kB / 64];
  long b0, b1;
  sas_range(n, b0, b1);
  int c = 0;
  for (long t = b0 + threadIdx.x; t < b1; t += kB) c += flags[t] ? 1 : 0;
  for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);
  if ((threadIdx.x % 64) == 0) sh[threadIdx.x / 64] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
    for (int w = 0; w < kB / 64; w++) tot += sh[w];
    counts[blockIdx.x] = tot;
  }
}

__global__ void __launch_bounds__(kB)
    k_sas_compact(AtomsDev A, const double* ref, int nref, double probe,
                  const unsigned char* flags, const int* offsets, double* xyz) {
  __shared__ int wsum[kB / 64];
  const long n = (long)A.n * nref;
  long b0, b1;
  sas_range(n, b0, b1);
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
  int base = offsets[blockIdx.x];
  for (long s0 = b0; s0 < b1; s0 += kB) {
    const long t = s0 + threadIdx.x;
    const bool ok = t < b1 && flags[t];
    const unsigned long long m = __ballot(ok);
    if (lane == 0) wsum[w] = (int)__popcll(m);
    __syncthreads();
    int before = 0, all = 0;
    for (int v = 0; v < kB / 64; v++) {
      if (v < w) before += wsum[v];
      all += wsum[v];
    }
    if (ok) {
      const int slot = base + before + (int)__popcll(m & ((1ull << lane) - 1ull));
      const int a = (int)(t / nref), q = (int)(t % nref);
      const double rad = A.rad[a] + probe;
      const double* ap = A.pos + 3 * a;
      xyz[3 * (size_t)slot] = rad * (ref[3 * q]) + ap[0];
      xyz[3 * (size_t)slot + 1] = rad * (ref[3 * q + 1]) + ap[1];
      xyz[3 * (size_t)slot + 2] = rad * (ref[3 * q + 2]) + ap[2];
    }
    base += all;
    __syncthreads();
  }
}

// markSphere (vpmg.c:6444-6491) as a scatter of flags: every grid point of a
// sphere gets byte 1 in its target's flag map.  Every writer of a byte writes
// the same value, so the result does not depend on the order of the writes.
// One 64-lane wavefront per sphere; for each target the lanes cover the
// sphere's bounding box column by column and test every point of it.  The point set is the
// reference loop's: its row skip ((dx2 + dy2) > rtot2) only drops points
// that also fail the final test, because rounding is monotone.
constexpr int kMarkWaves = kB / 64;

__global__ void __launch_bounds__(kB)
    k_mark_flags(Geo g, int npts, const double* pos, const double* rad, MarkTargets T,
                 unsigned char* flags) {
  const int a = blockIdx.x * kMarkWaves + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  if (a >= npts) return;
  const double* tp = pos + 3 * a;
  const size_t n = (size_t)g.nx * g.ny * g.nz;
  for (int ti = 0; ti < T.nt; ti++) {
    const MarkTarget tg = T.t[ti];
    const double add = tg.add;
    double rtot = add;
    if (tg.mode != 2) {
      const double ar = rad[a];
      if (tg.mode == 0) {
        if ((tp[0] <= g.xmin) || (tp[0] >= g.xmax) || (tp[1] <= g.ymin) || (tp[1] >= g.ymax) ||
            (tp[2] <= g.zmin) || (tp[2] >= g.zmax))
          continue;
        if (!(ar > VSMALL)) continue;
      } else {
        if (!(ar > VSMALL)) continue;
        if ((tp[0] < (g.xmin - add - ar)) || (tp[0] > (g.xmax + add + ar)) ||
            (tp[1] < (g.ymin - add - ar)) || (tp[1] > (g.ymax + add + ar)) ||
            (tp[2] < (g.zmin - add - ar)) || (tp[2] > (g.zmax + add + ar)))
          continue;
      }
      rtot = ar + add;
    }
    const double xs = g.xmin + (tg.which == 0 ? 0.5 * g.hx : 0.0);
    const double ys = g.ymin + (tg.which == 1 ? 0.5 * g.hy : 0.0);
    const double zs = g.zmin + (tg.which == 2 ? 0.5 * g.hz : 0.0);
    unsigned char* fl = flags + (size_t)ti * n;
    const double posx = tp[0] - xs, posy = tp[1] - ys, posz = tp[2] - zs;
    const double rtot2 = rtot * rtot;
    const double xr = rtot + 0.5 * g.hx, yr = rtot + 0.5 * g.hy, zr = rtot + 0.5 * g.hz;
    const int imin = imax2(0, (int)ceil((posx - xr) / g.hx));
    const int jmin = imax2(0, (int)ceil((posy - yr) / g.hy));
    const int kmin = imax2(0, (int)ceil((posz - zr) / g.hz));
    const int imax = imin2(g.nx - 1, (int)floor((posx + xr) / g.hx));
    const int jmax = imin2(g.ny - 1, (int)floor((posy + yr) / g.hy));
    const int kmax = imin2(g.nz - 1, (int)floor((posz + zr) / g.hz));
    const int ni = imax - imin + 1, nj = jmax - jmin + 1, nk = kmax - kmin + 1;
    if (ni <= 0 || nj <= 0 || nk <= 0) continue;
    // Lane e of a panel of 64 (i, j) columns of the box keeps its column and
    // walks the planes k: its column's squared offsets are formed once, the
    // plane's per step, and the sum keeps the reference's order
    // (dz2 + dy2) + dx2.
    const int ncol = ni * nj;
    const size_t nxy = (size_t)g.nx * g.ny;
    for (int e = lane; e < ncol; e += 64) {
      const int ii = e % ni, jj = e / ni;
      const double fi = imin + ii, fj = jmin + jj;
      const double dx2 = (posx - g.hx * fi) * (posx - g.hx * fi);
      const double dy2 = (posy - g.hy * fj) * (posy - g.hy * fj);
      size_t q = IX(imin + ii, jmin + jj, kmin, g.nx, g.ny);
      for (int kk = 0; kk < nk; kk++, q += nxy) {
        const double fk = kmin + kk;
        const double dz2 = (posz - g.hz * fk) * (posz - g.hz * fk);
        if ((dz2 + dy2 + dx2) <= rtot2) fl[q] = 1;
      }
    }
  }
}

// The solute bits of every grid point: bit w (map w = x, y, z shifted grid)
// is set inside an atom sphere (atom flags) and not within a probe radius of
// an SAS point (sas flags, absent when the probe radius is 0), where the
// unsmoothed dielectric value is epsp.
__global__ void __launch_bounds__(kB)
    k_solute_bits(Geo g, const unsigned char* fa, const unsigned char* fsas,
                  unsigned char* bits) {
  const int t = blockIdx.x * kB + threadIdx.x;
  const int k = blockIdx.y;
  const int i = t % g.nx, j = t / g.nx;
  if (j >= g.ny) return;
  const size_t n = (size_t)g.nx * g.ny * g.nz;
  const size_t p = IX(i, j, k, g.nx, g.ny);
  unsigned char b = 0;
  for (int w = 0; w < 3; w++)
    if (fa[w * n + p] != 0 && !(fsas && fsas[w * n + p] != 0)) b |= (unsigned char)(1 << w);
  bits[p] = b;
}

// The maps of the molecular surface from the solute bits, one grid point per
// thread.  The harmonic mean takes 1 / epsp or 1 / epsw for each neighbour
// value, the same quotients as the reference's 1.0 / a (vpmg.c:4500-4628),
// summed in the reference's order; the bits of the 13 points the three means
// read are loaded once.
__global__ void __launch_bounds__(kB)
    k_mol_maps(Geo g, const unsigned char* __restrict__ sb, const unsigned char* fk, double epsp,
               double epsw, double ionmask, int smooth, double* epsx, double* epsy, double* epsz,
               double* kappa) {
  const int t = blockIdx.x * kB + threadIdx.x;
  const int k = blockIdx.y;
  const int i = t % g.nx, j = t / g.nx;
  if (j >= g.ny) return;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const size_t nxy = (size_t)nx * ny;
  const size_t p = IX(i, j, k, nx, ny);
  kappa[p] = (fk && fk[p]) ? 0.0 : ionmask;
  const int c = sb[p];
  if (!smooth) {
    epsx[p] = (c & 1) ? epsp : epsw;
    epsy[p] = (c & 2) ? epsp : epsw;
    epsz[p] = (c & 4) ? epsp : epsw;
    return;
  }
  const bool xlo = i > 0, xhi = i < nx - 1, ylo = j > 0, yhi = j < ny - 1, zlo = k > 0,
             zhi = k < nz - 1;
  const int xm = xlo ? sb[p - 1] : 0, xp = xhi ? sb[p + 1] : 0;
  const int ym = ylo ? sb[p - nx] : 0, yp = yhi ? sb[p + nx] : 0;
  const int zm = zlo ? sb[p - nxy] : 0, zp = zhi ? sb[p + nxy] : 0;
  const int xp_ym = xhi && ylo ? sb[p + 1 - nx] : 0, xp_zm = xhi && zlo ? sb[p + 1 - nxy] : 0;
  const int xm_yp = xlo && yhi ? sb[p - 1 + nx] : 0, yp_zm = yhi && zlo ? sb[p + nx - nxy] : 0;
  const int xm_zp = xlo && zhi ? sb[p - 1 + nxy] : 0, ym_zp = ylo && zhi ? sb[p - nx + nxy] : 0;
  const double rp = 1.0 / epsp, rw = 1.0 / epsw;
  auto inv = [&](int w, int bits) { return ((bits >> w) & 1) ? rp : rw; };
  double frac;
  int np;
  frac = inv(0, c);
  frac += inv(1, c);
  frac += inv(2, c);
  np = 3;
  if (ylo) { frac += inv(1, ym); np += 1; }
  if (zlo) { frac += inv(2, zm); np += 1; }
  if (xhi) {
    frac += inv(1, xp);
    frac += inv(2, xp);
    np += 2;
    if (ylo) { frac += inv(1, xp_ym); np += 1; }
    if (zlo) { frac += inv(2, xp_zm); np += 1; }
  }
  epsx[p] = np / frac;

  frac = inv(1, c);
  frac += inv(0, c);
  frac += inv(2, c);
  np = 3;
  if (xlo) { frac += inv(0, xm); np += 1; }
  if (zlo) { frac += inv(2, zm); np += 1; }
  if (yhi) {
    frac += inv(0, yp);
    frac += inv(2, yp);
    np += 2;
    if (xlo) { frac += inv(0, xm_yp); np += 1; }
    if (zlo) { frac += inv(2, yp_zm); np += 1; }
  }
  epsy[p] = np / frac;

  frac = inv(2, c);
  frac += inv(0, c);
  frac += inv(1, c);
  np = 3;
  if (xlo) { frac += inv(0, xm); np += 1; }
  if (ylo) { frac += inv(1, ym); np += 1; }
  if (zhi) {
    frac += inv(0, zp);
    frac += inv(1, zp);
    np += 2;
    if (xlo) { frac += inv(0, xm_zp); np += 1; }
    if (ylo) { frac += inv(1, ym_zp); np += 1; }
  }
  epsz[p] = np / frac;
}

// -------------------------------------------------------------- boundaries
// face f: 0 x-low, 1 x-high, 2 y-low, 3 y-high, 4 z-low, 5 z-high
__device__ inline void face_point(const Geo& g, int f, int t, int& i, int& j, int& k) {
  if (f < 2) {
    j = t % g.ny;
    k = t / g.ny;
    i = f == 0 ? 0 : g.nx - 1;
  } else if (f < 4) {
    i = t % g.nx;
    k = t / g.nx;
    j = f == 2 ? 0 : g.ny - 1;
  } else {
    i = t % g.nx;
    j = t / g.nx;
    k = f == 4 ? 0 : g.nz - 1;
  }
}

__device__ inline double* face_slot(const Geo& g, int f, int i, int j, int k, double* gx, double* gy,
                                    double* gz) {
  if (f < 2) return gx + j + (size_t)g.ny * k + (f == 1 ? (size_t)g.ny * g.nz : 0);
  if (f < 4) return gy + i + (size_t)g.nx * k + (f == 3 ? (size_t)g.nx * g.nz : 0);
  return gz + i + (size_t)g.nx * j + (f == 5 ? (size_t)g.nx * g.ny : 0);
}

__device__ inline int face_count(const Geo& g, int f) {
  return f < 2 ? g.ny * g.nz : (f < 4 ? g.nx * g.nz : g.nx * g.ny);
}

// bcflnew / MDH (vpmg.c:3289-3367): the potential of every atom at every
// boundary point.  Atoms are split into nch contiguous chunks; blockIdx.z is
// the chunk, whose sequential partial sum goes to part[chunk][face][t], and
// k_bc_mdh_sum adds the partials in chunk order.  With one chunk the sum is
// the reference's atom-order sum.
//
// The reference's term pre1 * (q / d) * exp(-kappa (d - r)) / (1 + kappa r)
// is evaluated as c_a * exp(-kappa d) / d with the per-atom factor
// c_a = pre1 q exp(kappa r) / (1 + kappa r) (q = ec * charge), and 1 / d from
// the hardware reciprocal square root refined by two Newton steps, so a term
// costs one exponential and no division or square root.  Each term is within
// a few ulp of the reference's expression.
__global__ void __launch_bounds__(kB)
    k_bc_mdh_atoms(AtomsDev A, double pre1, double xkappa, double* fac) {
  const int a = blockIdx.x * kB + threadIdx.x;
  if (a >= A.n) return;
  const double ec = 1.6021773e-19;  // Vunit_ec, src/generic/vunit.h:80
  const double q = ec * A.chg[a];
  if (xkappa > VSMALL) {
    const double r = A.rad[a];
    fac[a] = pre1 * q * exp(xkappa * r) / (1 + xkappa * r);
  } else {
    fac[a] = pre1 * q;
  }
}

// 1 / sqrt(d2) for d2 > 0: v_rsq_f64 and two Newton steps y += y (1 - d2 y^2) / 2
__device__ inline double rsqrt_refined(double d2) {
  double y = __builtin_amdgcn_rsq(d2);
  double e = fma(-d2 * y, y, 1.0);
  y = fma(0.5 * y, e, y);
  e = fma(-d2 * y, y, 1.0);
  return fma(0.5 * y, e, y);
}

__global__ void __launch_bounds__(kB)
    k_bc_mdh(Geo g, const double* xf, const double* yf, const double* zf, AtomsDev A,
             const double* fac, double xkappa, int nch, int mx, double* part) {
  const int f = blockIdx.y;
  const int t = blockIdx.x * kB + threadIdx.x;
  if (t >= face_count(g, f)) return;
  int i, j, k;
  face_point(g, f, t, i, j, k);
  const double px = xf[i], py = yf[j], pz = zf[k];
  const int c = blockIdx.z;
  const int per = (A.n + nch - 1) / nch;
  const int a0 = c * per, a1 = min(A.n, a0 + per);
  const bool screened = xkappa > VSMALL;
  auto term = [&](int a) {
    const double* ap = A.pos + 3 * a;
    const double dx = px - ap[0], dy = py - ap[1], dz = pz - ap[2];
    const double d2 = fma(dx, dx, fma(dy, dy, dz * dz));
    const double y = rsqrt_refined(d2);
    const double v = fac[a] * y;
    return screened ? v * exp(-xkappa * (d2 * y)) : v;
  };
  // four independent terms in flight (their dependency chains interleave),
  // added in atom order
  double val = 0.0;
  int a = a0;
  for (; a + 3 < a1; a += 4) {
    const double t0 = term(a), t1 = term(a + 1), t2 = term(a + 2), t3 = term(a + 3);
    val += t0;
    val += t1;
    val += t2;
    val += t3;
  }
  for (; a < a1; a++) val += term(a);
  part[((size_t)c * 6 + f) * mx + t] = val;
}

// The same sum from FP32 terms, added in FP64 (the fast-boundary mode of the
// mixed-precision solve).  Atom and boundary coordinates are taken relative
// to the box centre before rounding to FP32, so a distance carries an
// absolute error of a few 1e-6 A; 1 / d is the hardware reciprocal square
// root and exp(-kappa d) the hardware 2^x of -kappa log2(e) d, each about
// one ulp.  A term is within about 1e-6 relative of the FP64 term (the
// exponent's argument, at most a few tens, dominates); the partial sums and
// their order are those of k_bc_mdh.
__global__ void __launch_bounds__(kB)
    k_bc_mdh_atoms32(AtomsDev A, double pre1, double xkappa, double cx, double cy, double cz,
                     float4* at) {
  const int a = blockIdx.x * kB + threadIdx.x;
  if (a >= A.n) return;
  const double ec = 1.6021773e-19;  // Vunit_ec, src/generic/vunit.h:80
  const double q = ec * A.chg[a];
  double fac = pre1 * q;
  if (xkappa > VSMALL) {
    const double r = A.rad[a];
    fac = pre1 * q * exp(xkappa * r) / (1 + xkappa * r);
  }
  const double* ap = A.pos + 3 * a;
  at[a] = make_float4((float)(ap[0] - cx), (float)(ap[1] - cy), (float)(ap[2] - cz), (float)fac);
}

__global__ void __launch_bounds__(kB)
    k_bc_mdh32(Geo g, const double* xf, const double* yf, const double* zf, double cx, double cy,
               double cz, int natoms, const float4* at, float nk2, int nch, int mx,
               double* part) {
  const int f = blockIdx.y;
  const int t = blockIdx.x * kB + threadIdx.x;
  if (t >= face_count(g, f)) return;
  int i, j, k;
  face_point(g, f, t, i, j, k);
  const float px = (float)(xf[i] - cx), py = (float)(yf[j] - cy), pz = (float)(zf[k] - cz);
  const int c = blockIdx.z;
  const int per = (natoms + nch - 1) / nch;
  const int a0 = c * per, a1 = min(natoms, a0 + per);
  const bool screened = nk2 != 0.0f;
  auto term = [&](int a) {
    const float4 q = at[a];
    const float dx = px - q.x, dy = py - q.y, dz = pz - q.z;
    const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    const float y = __builtin_amdgcn_rsqf(d2);
    const float v = q.w * y;
    return screened ? v * __builtin_amdgcn_exp2f(nk2 * (d2 * y)) : v;
  };
  double val = 0.0;
  int a = a0;
  for (; a + 3 < a1; a += 4) {
    const float t0 = term(a), t1 = term(a + 1), t2 = term(a + 2), t3 = term(a + 3);
    val += (double)t0;
    val += (double)t1;
    val += (double)t2;
    val += (double)t3;
  }
  for (; a < a1; a++) val += (double)term(a);
  part[((size_t)c * 6 + f) * mx + t] = val;
}

__global__ void __launch_bounds__(kB)
    k_bc_mdh_sum(Geo g, int nch, int mx, const double* part, double* gx, double* gy, double* gz) {
  const int f = blockIdx.y;
  const int t = blockIdx.x * kB + threadIdx.x;
  if (t >= face_count(g, f)) return;
  int i, j, k;
  face_point(g, f, t, i, j, k);
  double val = part[(size_t)f * mx + t];
  for (int c = 1; c < nch; c++) val += part[((size_t)c * 6 + f) * mx + t];
  *face_slot(g, f, i, j, k, gx, gy, gz) = val;
}

// multipolebc (vpmg.c:3369-3477): radial factors of the monopole, dipole and
// traceless-quadrupole potentials outside a dielectric sphere.
__device__ inline void multipole_tensor(double r, double kappa, double eps_p, double eps_w,
                                        double rad, double tsr[3]) {
  const double eps_r = eps_w / eps_p;
  const double r2 = r * r, r3 = r2 * r, r5 = r3 * r2;
  tsr[0] = (1.0 / eps_w) / r;
  tsr[1] = (1.0 / eps_w) * (-1.0) / r3;
  tsr[2] = (1.0 / eps_w) * (3.0) / r5;
  if (kappa < VSMALL) {
    tsr[1] = (3.0 * eps_r) / (1.0 + 2.0 * eps_r) * tsr[1];
    tsr[2] = (5.0 * eps_r) / (2.0 + 3.0 * eps_r) * tsr[2];
  } else {
    const double ka = kappa * rad, ka2 = ka * ka, ka3 = ka2 * ka;
    const double kr = kappa * r, kr2 = kr * kr;
    tsr[0] = exp(ka - kr) / (1.0 + ka) * tsr[0];
    tsr[1] = 3.0 * eps_r * exp(ka - kr) * (1.0 + kr) * tsr[1];
    tsr[1] = tsr[1] / (1.0 + ka + eps_r * (2.0 + 2.0 * ka + ka2));
    tsr[2] = 5.0 * eps_r * exp(ka - kr) * (3.0 + 3.0 * kr + kr2) * tsr[2];
    tsr[2] = tsr[2] / (6.0 + 6.0 * ka + 2.0 * ka2 + eps_r * (9.0 + 9.0 * ka + 4.0 * ka2 + ka3));
  }
}

// bcfl_sdh (vpmg.c:3479-3645) on the boundary faces.
__global__ void __launch_bounds__(kB)
    k_bc_sdh(Geo g, const double* xf, const double* yf, const double* zf, SdhMoments M,
             double* gx, double* gy, double* gz) {
  const int f = blockIdx.y;
  const int t = blockIdx.x * kB + threadIdx.x;
  if (t >= face_count(g, f)) return;
  int i, j, k;
  face_point(g, f, t, i, j, k);
  const double xr = xf[i] - M.pos[0], yr = yf[j] - M.pos[1], zr = zf[k] - M.pos[2];
  const double dist = sqrt(xr * xr + yr * yr + zr * zr);
  double tsr[3];
  multipole_tensor(dist, M.xkappa, M.eps_p, M.eps_w, M.size, tsr);
  const double pre = M.pre;
  double val = pre * M.q * tsr[0];
  val -= pre * M.u[0] * xr * tsr[1];
  val -= pre * M.u[1] * yr * tsr[1];
  val -= pre * M.u[2] * zr * tsr[1];
  val += pre * M.qxx * xr * xr * tsr[2];
  val += pre * M.qyy * yr * yr * tsr[2];
  val += pre * M.qzz * zr * zr * tsr[2];
  val += pre * 2.0 * M.qxy * xr * yr * tsr[2];
  val += pre * 2.0 * M.qxz * xr * zr * tsr[2];
  val += pre * 2.0 * M.qyz * yr * zr * tsr[2];
  *face_slot(g, f, i, j, k, gx, gy, gz) = val;
}

__device__ inline double trilin(const OldGrid& o, double x, double y, double z, int* bad) {
  const double VS = VSMALL;
  const bool ok = (x >= (o.xmin - VS)) && (y >= (o.ymin - VS)) && (z >= (o.zmin - VS)) &&
                  (x <= (o.xmax + VS)) && (y <= (o.ymax + VS)) && (z <= (o.zmax + VS));
  if (!ok) {
    atomicExch(bad, 1);
    return 0.0;
  }
  const double fi = (x - o.xmin) / o.hx, fj = (y - o.ymin) / o.hy, fk = (z - o.zmin) / o.hz;
  int ihi = (int)ceil(fi); if (ihi > (o.nx - 1)) ihi = o.nx - 1;
  int ilo = (int)floor(fi); if (ilo < 0) ilo = 0;
  int jhi = (int)ceil(fj); if (jhi > (o.ny - 1)) jhi = o.ny - 1;
  int jlo = (int)floor(fj); if (jlo < 0) jlo = 0;
  int khi = (int)ceil(fk); if (khi > (o.nz - 1)) khi = o.nz - 1;
  int klo = (int)floor(fk); if (klo < 0) klo = 0;
  const double dx = fi - (double)(ilo), dy = fj - (double)(jlo), dz = fk - (double)(klo);
  const double* d = o.u;
  const int nx = o.nx, ny = o.ny;
  return dx * dy * dz * (d[IX(ihi, jhi, khi, nx, ny)]) + dx * (1.0 - dy) * dz * (d[IX(ihi, jlo, khi, nx, ny)]) +
         dx * dy * (1.0 - dz) * (d[IX(ihi, jhi, klo, nx, ny)]) +
         dx * (1.0 - dy) * (1.0 - dz) * (d[IX(ihi, jlo, klo, nx, ny)]) +
         (1.0 - dx) * dy * dz * (d[IX(ilo, jhi, khi, nx, ny)]) +
         (1.0 - dx) * (1.0 - dy) * dz * (d[IX(ilo, jlo, khi, nx, ny)]) +
         (1.0 - dx) * dy * (1.0 - dz) * (d[IX(ilo, jhi, klo, nx, ny)]) +
         (1.0 - dx) * (1.0 - dy) * (1.0 - dz) * (d[IX(ilo, jlo, klo, nx, ny)]);
}

__global__ void __launch_bounds__(kB)
    k_focus(Geo g, double xminN, double yminN, double zminN, double xmaxN, double ymaxN,
            double zmaxN, OldGrid o, double* gx, double* gy, double* gz, int* bad) {
  const int f = blockIdx.y;
  const int t = blockIdx.x * kB + threadIdx.x;
  if (t >= face_count(g, f)) return;
  int i, j, k;
  face_point(g, f, t, i, j, k);
  double x, y, z;
  if (f < 2) {
    x = f == 0 ? xminN : xmaxN;
    y = yminN + j * g.hy;
    z = zminN + k * g.hz;
  } else if (f < 4) {
    x = xminN + i * g.hx;
    y = f == 2 ? yminN : ymaxN;
    z = zminN + k * g.hz;
  } else {
    x = xminN + i * g.hx;
    y = yminN + j * g.hy;
    z = f == 4 ? zminN : zmaxN;
  }
  *face_slot(g, f, i, j, k, gx, gy, gz) = trilin(o, x, y, z, bad);
}

// -------------------------------------------------------------- partition
__device__ inline double part_axis(double x, double lo, double hi, double h) {
  double ok;
  if ((x < (hi - h / 2)) && (x > (lo + h / 2))) ok = 1.0;
  else if (fabs(x - lo) < VPMGSMALL) ok = 1.0;
  else if (fabs(x - hi) < VPMGSMALL) ok = 1.0;
  else if ((x > (hi + h / 2)) || (x < (lo - h / 2))) ok = 0.0;
  else {
    const double x0 = fmax(x - h / 2, lo), x1 = fmin(x + h / 2, hi);
    ok = fabs(x1 - x0) / h;
    if (ok < 0.0 && fabs(ok) < VPMGSMALL) ok = 0.0;
    if (ok > 1.0 && fabs(ok - 1.0) < VPMGSMALL) ok = 1.0;
  }
  return ok;
}

__global__ void __launch_bounds__(kB)
    k_set_part(Geo g, double xm, double ym, double zm, double lo0, double lo1, double lo2,
               double hi0, double hi1, double hi2, int invert, double* pvec) {
  const int t = blockIdx.x * kB + threadIdx.x;
  const int k = blockIdx.y;
  const int i = t % g.nx, j = t / g.nx;
  if (j >= g.ny) return;
  const double xok = part_axis(i * g.hx + xm, lo0, hi0, g.hx);
  const double yok = part_axis(j * g.hy + ym, lo1, hi1, g.hy);
  const double zok = part_axis(k * g.hz + zm, lo2, hi2, g.hz);
  const double v0 = xok * yok * zok;
  double v = (fabs(v0) < VPMGSMALL) ? 0.0 : v0;
  if (invert) {
    if (v > VSMALL) v = 1.0;
    v = (1 - v) * 1.0;
  }
  pvec[IX(i, j, k, g.nx, g.ny)] = v;
}

// ---------------------------------------------------------------- energies
__device__ inline double bsum(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int w = kB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

__global__ void __launch_bounds__(kB) k_fin(const double* part, double* out) {
  __shared__ double sh[kB];
  double a = 0.0;
  for (int t = threadIdx.x; t < kParts; t += kB) a += part[t];
  double s = bsum(a, sh);
  if (threadIdx.x == 0) out[0] = s;
}

__global__ void __launch_bounds__(kB)
    k_e_qf(size_t n, const double* pvec, const double* u, const double* charge, double* part) {
  __shared__ double sh[kB];
  double acc = 0.0;
  for (size_t t = (size_t)blockIdx.x * kB + threadIdx.x; t < n; t += (size_t)kParts * kB)
    acc += (pvec[t] * u[t] * charge[t]);
  double s = bsum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kB)
    k_e_diel(Geo g, const double* pv, const double* u, const double* ex, const double* ey,
             const double* ez, double* part) {
  __shared__ double sh[kB];
  const int nx = g.nx, ny = g.ny;
  const long ni = nx - 1, nij = ni * (ny - 1), n = nij * (g.nz - 1);
  double acc = 0.0;
  for (long t = (long)blockIdx.x * kB + threadIdx.x; t < n; t += (long)kParts * kB) {
    const int i = (int)(t % ni), j = (int)((t / ni) % (ny - 1)), k = (int)(t / nij);
    const size_t p = IX(i, j, k, nx, ny);
    const double pvx = 0.5 * (pv[p] + pv[p + 1]);
    const double pvy = 0.5 * (pv[p] + pv[p + nx]);
    const double pvz = 0.5 * (pv[p] + pv[p + (size_t)nx * ny]);
    const double dxu = (u[p] - u[p + 1]) / g.hx, dyu = (u[p] - u[p + nx]) / g.hy,
                 dzu = (u[p] - u[p + (size_t)nx * ny]) / g.hz;
    const double nrgx = ex[p] * pvx * (dxu * dxu);
    const double nrgy = ey[p] * pvy * (dyu * dyu);
    const double nrgz = ez[p] * pvz * (dzu * dzu);
    acc += (nrgx + nrgy + nrgz);
  }
  double s = bsum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__device__ inline double capexp(double x, int* ichop) {  // vcap.c:59-73
  if (x > 85.0) { *ichop = 1; return exp(85.0); }
  if (x < -85.0) { *ichop = 1; return exp(-85.0); }
  *ichop = 0;
  return exp(x);
}

__global__ void __launch_bounds__(kB)
    k_e_qm(size_t n, int nonlin, const double* pvec, const double* kappa, const double* u,
           double zks2, double zkappa2, IonBulk I, double* part, int* nchop) {
  __shared__ double sh[kB];
  double acc = 0.0;
  int chops = 0;
  for (size_t t = (size_t)blockIdx.x * kB + threadIdx.x; t < n; t += (size_t)kParts * kB) {
    if (pvec[t] * kappa[t] > VSMALL) {
      if (nonlin) {
        for (int s = 0; s < I.nion; s++) {
          int ic;
          acc += (pvec[t] * kappa[t] * zks2 * I.c[s] * (capexp(-I.q[s] * u[t], &ic) - 1.0));
          chops += ic;
        }
      } else {
        acc += (pvec[t] * zkappa2 * kappa[t] * (u[t] * u[t]));
      }
    }
  }
  if (chops) atomicAdd(nchop, chops);
  double s = bsum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_zero_int(int* p) { *p = 0; }

// one thread per atom, partial per atom, then fixed-order sum
__global__ void __launch_bounds__(kB)
    k_e_qf_point(Geo g, AtomsDev A, const double* u, double* per_atom) {
  const int a = blockIdx.x * kB + threadIdx.x;
  if (a >= A.n) return;
  const double* ap = A.pos + 3 * a;
  const int nx = g.nx, ny = g.ny;
  const double fi = (ap[0] - g.xmin) / g.hx, fj = (ap[1] - g.ymin) / g.hy, fk = (ap[2] - g.zmin) / g.hz;
  const int ihi = (int)ceil(fi), ilo = (int)floor(fi), jhi = (int)ceil(fj), jlo = (int)floor(fj);
  const int khi = (int)ceil(fk), klo = (int)floor(fk);
  double e = 0.0;
  if (A.partID[a] > 0 && (ihi < g.nx) && (jhi < g.ny) && (khi < g.nz) && (ilo >= 0) &&
      (jlo >= 0) && (klo >= 0)) {
    const double dx = fi - (double)(ilo), dy = fj - (double)(jlo), dz = fk - (double)(klo);
    const double uval =
        dx * dy * dz * u[IX(ihi, jhi, khi, nx, ny)] + dx * (1.0 - dy) * dz * u[IX(ihi, jlo, khi, nx, ny)] +
        dx * dy * (1.0 - dz) * u[IX(ihi, jhi, klo, nx, ny)] +
        dx * (1.0 - dy) * (1.0 - dz) * u[IX(ihi, jlo, klo, nx, ny)] +
        (1.0 - dx) * dy * dz * u[IX(ilo, jhi, khi, nx, ny)] +
        (1.0 - dx) * (1.0 - dy) * dz * u[IX(ilo, jlo, khi, nx, ny)] +
        (1.0 - dx) * dy * (1.0 - dz) * u[IX(ilo, jhi, klo, nx, ny)] +
        (1.0 - dx) * (1.0 - dy) * (1.0 - dz) * u[IX(ilo, jlo, klo, nx, ny)];
    e = (uval * A.chg[a] * A.partID[a]);
  }
  per_atom[a] = e;
}

// sequential sum over atoms in index order (one thread): same order as the
// reference's loop (vpmg.c:1740-1785)
__global__ void k_seq_sum(const double* v, int n, double* out) {
  if (threadIdx.x != 0) return;
  double s = 0.0;
  for (int a = 0; a < n; a++) s += v[a];
  out[0] = s;
}

// ------------------------------------------------------------------ forces
__global__ void k_force_qf(Geo g, AtomsDev A, int mode, const double* u, double* out) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= A.n) return;
  const double* ap = A.pos + 3 * a;
  const int nx = g.nx, ny = g.ny;
  double f0 = 0.0, f1 = 0.0, f2 = 0.0;
  const double charge = A.chg[a];
  if (A.partID[a] != 0) {
    if (mode == 2) {  // qfForceSpline4 (vpmg.c:6156-6275)
      if (!((ap[0] <= (g.xmin + 2 * g.hx)) || (ap[0] >= (g.xmax - 2 * g.hx)) ||
            (ap[1] <= (g.ymin + 2 * g.hy)) || (ap[1] >= (g.ymax - 2 * g.hy)) ||
            (ap[2] <= (g.zmin + 2 * g.hz)) || (ap[2] >= (g.zmax - 2 * g.hz)))) {
        const double fi = (ap[0] - g.xmin) / g.hx, fj = (ap[1] - g.ymin) / g.hy,
                     fk = (ap[2] - g.zmin) / g.hz;
        const int ip2 = imin2((int)ceil(fi) + 2, g.nx - 1), im2 = imax2((int)floor(fi) - 2, 0);
        const int jp2 = imin2((int)ceil(fj) + 2, g.ny - 1), jm2 = imax2((int)floor(fj) - 2, 0);
        const int kp2 = imin2((int)ceil(fk) + 2, g.nz - 1), km2 = imax2((int)floor(fk) - 2, 0);
        double e0 = 0.0, e1 = 0.0, e2 = 0.0;
        for (int ii = im2; ii <= ip2; ii++) {
          const double mi = vfchi4(ii, fi), mx = bspline4(mi), dmx = dbspline4(mi);
          for (int jj = jm2; jj <= jp2; jj++) {
            const double mj = vfchi4(jj, fj), my = bspline4(mj), dmy = dbspline4(mj);
            for (int kk = km2; kk <= kp2; kk++) {
              const double mk = vfchi4(kk, fk), mz = bspline4(mk), dmz = dbspline4(mk);
              const double f = u[IX(ii, jj, kk, nx, ny)];
              e0 += f * dmx * my * mz / g.hx;
              e1 += f * mx * dmy * mz / g.hy;
              e2 += f * mx * my * dmz / g.hz;
            }
          }
        }
        f0 = e0 * charge;
        f1 = e1 * charge;
        f2 = e2 * charge;
      }
    } else if (mode == 1) {  // qfForceSpline2 (vpmg.c:6043-6154)
      if (!((ap[0] <= (g.xmin + g.hx)) || (ap[0] >= (g.xmax - g.hx)) || (ap[1] <= (g.ymin + g.hy)) ||
            (ap[1] >= (g.ymax - g.hy)) || (ap[2] <= (g.zmin + g.hz)) || (ap[2] >= (g.zmax - g.hz)))) {
        const double fi = (ap[0] - g.xmin) / g.hx, fj = (ap[1] - g.ymin) / g.hy,
                     fk = (ap[2] - g.zmin) / g.hz;
        const int ip2 = imin2((int)ceil(fi) + 1, g.nx - 1), im2 = imax2((int)floor(fi) - 1, 0);
        const int jp2 = imin2((int)ceil(fj) + 1, g.ny - 1), jm2 = imax2((int)floor(fj) - 1, 0);
        const int kp2 = imin2((int)ceil(fk) + 1, g.nz - 1), km2 = imax2((int)floor(fk) - 1, 0);
        for (int ii = im2; ii <= ip2; ii++) {
          const double mx = bspline2(vfchi(ii, fi)), dmx = dbspline2(vfchi(ii, fi));
          for (int jj = jm2; jj <= jp2; jj++) {
            const double my = bspline2(vfchi(jj, fj)), dmy = dbspline2(vfchi(jj, fj));
            for (int kk = km2; kk <= kp2; kk++) {
              const double mz = bspline2(vfchi(kk, fk)), dmz = dbspline2(vfchi(kk, fk));
              const double uu = u[IX(ii, jj, kk, nx, ny)];
              f0 += (charge * dmx * my * mz * uu) / g.hx;
              f1 += (charge * mx * dmy * mz * uu) / g.hy;
              f2 += (charge * mx * my * dmz * uu) / g.hz;
            }
          }
        }
      }
    } else if (!((ap[0] <= g.xmin) || (ap[0] >= g.xmax) || (ap[1] <= g.ymin) ||
                 (ap[1] >= g.ymax) || (ap[2] <= g.zmin) || (ap[2] >= g.zmax))) {
      // qfForceSpline1 (vpmg.c:5899-6041)
      const double fi = (ap[0] - g.xmin) / g.hx, fj = (ap[1] - g.ymin) / g.hy, fk = (ap[2] - g.zmin) / g.hz;
      const int ihi = (int)ceil(fi), ilo = (int)floor(fi), jhi = (int)ceil(fj), jlo = (int)floor(fj);
      const int khi = (int)ceil(fk), klo = (int)floor(fk);
      const double dx = fi - (double)(ilo), dy = fj - (double)(jlo), dz = fk - (double)(klo);
#define U(a_, b_, c_) u[IX(a_, b_, c_, nx, ny)]
      if ((dx > VPMGSMALL) && (fabs(1.0 - dx) > VPMGSMALL))
        f0 = -charge * (dy * dz * U(ihi, jhi, khi) + dy * (1 - dz) * U(ihi, jhi, klo) +
                        (1 - dy) * dz * U(ihi, jlo, khi) + (1 - dy) * (1 - dz) * U(ihi, jlo, klo) -
                        dy * dz * U(ilo, jhi, khi) - dy * (1 - dz) * U(ilo, jhi, klo) -
                        (1 - dy) * dz * U(ilo, jlo, khi) - (1 - dy) * (1 - dz) * U(ilo, jlo, klo)) / g.hx;
      if ((dy > VPMGSMALL) && (fabs(1.0 - dy) > VPMGSMALL))
        f1 = -charge * (dx * dz * U(ihi, jhi, khi) + dx * (1 - dz) * U(ihi, jhi, klo) -
                        dx * dz * U(ihi, jlo, khi) - dx * (1 - dz) * U(ihi, jlo, klo) +
                        (1 - dx) * dz * U(ilo, jhi, khi) + (1 - dx) * (1 - dz) * U(ilo, jhi, klo) -
                        (1 - dx) * dz * U(ilo, jlo, khi) - (1 - dx) * (1 - dz) * U(ilo, jlo, klo)) / g.hy;
      if ((dz > VPMGSMALL) && (fabs(1.0 - dz) > VPMGSMALL))
        f2 = -charge * (dy * dx * U(ihi, jhi, khi) - dy * dx * U(ihi, jhi, klo) +
                        (1 - dy) * dx * U(ihi, jlo, khi) - (1 - dy) * dx * U(ihi, jlo, klo) +
                        dy * (1 - dx) * U(ilo, jhi, khi) - dy * (1 - dx) * U(ilo, jhi, klo) +
                        (1 - dy) * (1 - dx) * U(ilo, jlo, khi) - (1 - dy) * (1 - dx) * U(ilo, jlo, klo)) / g.hz;
#undef U
    }
  }
  out[3 * a] = f0;
  out[3 * a + 1] = f1;
  out[3 * a + 2] = f2;
}

// Vacc_splineAccGradAtomNorm (vacc.c:314-373)
__device__ inline void spline_grad(const double* c, double win, double infrad, const double* ap,
                                   double aradius, double* grad) {
  const double w2i = 1.0 / (win * win), w3i = 1.0 / (win * win * win);
  grad[0] = grad[1] = grad[2] = 0.0;
  if (aradius > 0.0) {
    const double arad = aradius + infrad;
    const double dist = sqrt((ap[0] - c[0]) * (ap[0] - c[0]) + (ap[1] - c[1]) * (ap[1] - c[1]) +
                             (ap[2] - c[2]) * (ap[2] - c[2]));
    if (dist < (arad - win)) return;
    if (dist > (arad + win)) return;
    if ((fabs(dist - (arad - win)) < VSMALL) || (fabs(dist - (arad + win)) < VSMALL)) return;
    const double sm = dist - arad + win, sm2 = sm * sm;
    const double mychi = 0.75 * sm2 * w2i - 0.25 * sm * sm2 * w3i;
    const double mygrad = 1.5 * sm * w2i - 0.75 * sm2 * w3i;
    for (int i = 0; i < 3; i++) grad[i] = -(mygrad / mychi) * ((c[i] - ap[i]) / dist);
  }
}

// Vacc_splineAccGradAtomNorm3 / Norm4 (vacc.c:1097, 1004) with the window
// coefficients of (radius + infrad) +- win precomputed on the host.
template <int kOrd>
__device__ inline void poly_grad(const double* cf, double win, double infrad, const double* c,
                                 const double* ap, double aradius, double* grad) {
  grad[0] = grad[1] = grad[2] = 0.0;
  if (aradius > 0.0) {
    const double arad = aradius + infrad;
    const double dist = sqrt((ap[0] - c[0]) * (ap[0] - c[0]) + (ap[1] - c[1]) * (ap[1] - c[1]) +
                             (ap[2] - c[2]) * (ap[2] - c[2]));
    if (dist < (arad - win)) return;
    if (dist > (arad + win)) return;
    if ((fabs(dist - (arad - win)) < VSMALL) || (fabs(dist - (arad + win)) < VSMALL)) return;
    const double sm = dist, sm2 = sm * sm, sm3 = sm2 * sm, sm4 = sm3 * sm;
    double mychi, mygrad;
    if (kOrd == 3) {
      const double sm5 = sm4 * sm;
      mychi = cf[0] + cf[1] * sm + cf[2] * sm2 + cf[3] * sm3 + cf[4] * sm4 + cf[5] * sm5;
      mygrad = cf[1] + 2.0 * cf[2] * sm + 3.0 * cf[3] * sm2 + 4.0 * cf[4] * sm3 +
               5.0 * cf[5] * sm4;
    } else {
      const double sm5 = sm4 * sm, sm6 = sm5 * sm, sm7 = sm6 * sm;
      mychi = cf[0] + cf[1] * sm + cf[2] * sm2 + cf[3] * sm3 + cf[4] * sm4 + cf[5] * sm5 +
              cf[6] * sm6 + cf[7] * sm7;
      mygrad = cf[1] + 2.0 * cf[2] * sm + 3.0 * cf[3] * sm2 + 4.0 * cf[4] * sm3 +
               5.0 * cf[5] * sm4 + 6.0 * cf[6] * sm5 + 7.0 * cf[7] * sm6;
    }
    if (mychi <= 0.0) return;
    if (mychi > 1.0) mychi = 1.0;
    for (int i = 0; i < 3; i++) grad[i] = -(mygrad / mychi) * ((c[i] - ap[i]) / dist);
  }
}

// The gradient of the surface method in use (Vpmg_splineSelect, vpmg.c:1890-1910).
__device__ inline void win_grad(int ord, const double* cf, const double* c, double win,
                                double infrad, const double* ap, double aradius, double* grad) {
  if (ord == 3)
    poly_grad<3>(cf, win, infrad, c, ap, aradius, grad);
  else if (ord == 4)
    poly_grad<4>(cf, win, infrad, c, ap, aradius, grad);
  else
    spline_grad(c, win, infrad, ap, aradius, grad);
}

// Ion-boundary force (vpmg.c:5440-5603): one block per atom.  The box is
// walked by a single thread per atom in the reference's i, j, k order so
// the sum rounds the same way.
__global__ void k_force_ib(Geo g, AtomsDev A, double win, double irad, double zkappa2,
                           double izmagic, int nonlin, double ionstr, IonBulk I,
                           const double* kappa, const double* u, double* out, int ord,
                           const double* pc) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= A.n) return;
  const double* ap = A.pos + 3 * a;
  const double arad = A.rad[a];
  double f[3] = {0.0, 0.0, 0.0};
  const double hx = g.hx, hy = g.hy, hzed = g.hz;
  if (A.partID[a] != 0 && !(zkappa2 < VPMGSMALL) &&
      !((ap[0] <= g.xmin) || (ap[0] >= g.xmax) || (ap[1] <= g.ymin) || (ap[1] >= g.ymax) ||
        (ap[2] <= g.zmin) || (ap[2] >= g.zmax))) {
    const double p0 = ap[0] - g.xmin, p1 = ap[1] - g.ymin, p2 = ap[2] - g.zmin;
    const double rtot = (irad + arad + win), rtot2 = rtot * rtot;
    const double dx = rtot + 0.5 * hx;
    const int imin = imax2(0, (int)ceil((p0 - dx) / hx));
    const int imax = imin2(g.nx - 1, (int)floor((p0 + dx) / hx));
    for (int i = imin; i <= imax; i++) {
      const double dx2 = (p0 - hx * i) * (p0 - hx * i);
      const double dy = (rtot2 > dx2) ? sqrt(rtot2 - dx2) + 0.5 * hy : 0.5 * hy;
      const int jmin = imax2(0, (int)ceil((p1 - dy) / hy));
      const int jmax = imin2(g.ny - 1, (int)floor((p1 + dy) / hy));
      for (int j = jmin; j <= jmax; j++) {
        const double dy2 = (p1 - hy * j) * (p1 - hy * j);
        const double dz = (rtot2 > (dx2 + dy2)) ? sqrt(rtot2 - dx2 - dy2) + 0.5 * hzed : 0.5 * hzed;
        const int kmin = imax2(0, (int)ceil((p2 - dz) / hzed));
        const int kmax = imin2(g.nz - 1, (int)floor((p2 + dz) / hzed));
        for (int k = kmin; k <= kmax; k++) {
          const double dz2 = (k * hzed - p2) * (k * hzed - p2);
          if ((dz2 + dy2 + dx2) <= rtot2) {
            const double gp[3] = {i * hx + g.xmin, j * hy + g.ymin, k * hzed + g.zmin};
            double tg[3];
            win_grad(ord, pc ? pc + kSplCoef * a + 16 : nullptr, gp, win, irad, ap, arad, tg);
            const size_t p = IX(i, j, k, g.nx, g.ny);
            double fmag;
            if (nonlin) {
              fmag = 0.0;
              for (int s = 0; s < I.nion; s++) {
                int ic;
                fmag += (kappa[p]) * I.c[s] * (capexp(-I.q[s] * u[p], &ic) - 1.0) / ionstr;
              }
            } else {
              fmag = (u[p] * u[p]) * (kappa[p]);
            }
            f[0] += (zkappa2 * fmag * tg[0]);
            f[1] += (zkappa2 * fmag * tg[1]);
            f[2] += (zkappa2 * fmag * tg[2]);
          }
        }
      }
    }
    for (int d = 0; d < 3; d++) f[d] = f[d] * 0.5 * hx * hy * hzed * izmagic;
  }
  out[3 * a] = f[0];
  out[3 * a + 1] = f[1];
  out[3 * a + 2] = f[2];
}

// Dielectric-boundary force (vpmg.c:5605-5860), one thread per atom, box in
// the reference's loop order.
__global__ void k_force_db(Geo g, AtomsDev A, double win, double srad, double epsp, double epsw,
                           double izmagic, const double* epsx, const double* epsy,
                           const double* epsz, const double* u, double* out, int* rcs, int ord,
                           const double* pc) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= A.n) return;
  const double* ap = A.pos + 3 * a;
  const double arad = A.rad[a];
  const int nx = g.nx, ny = g.ny;
  const double hx = g.hx, hy = g.hy, hzed = g.hz;
  double F[3] = {0.0, 0.0, 0.0};
  int rc = 1;
  const double rtot = (arad + win + srad);
  bool go = A.partID[a] != 0 && !(fabs(epsp - epsw) < VPMGSMALL) &&
            !((ap[0] <= g.xmin + rtot) || (ap[0] >= g.xmax - rtot) || (ap[1] <= g.ymin + rtot) ||
              (ap[1] >= g.ymax - rtot) || (ap[2] <= g.zmin + rtot) || (ap[2] >= g.zmax - rtot));
  int imin = 0, imax = -1, jmin = 0, jmax = -1, kmin = 0, kmax = -1;
  if (go) {
    const double p0 = ap[0] - g.xmin, p1 = ap[1] - g.ymin, p2 = ap[2] - g.zmin;
    imin = (int)floor((p0 - rtot) / hx);
    imax = (int)ceil((p0 + rtot) / hx);
    jmin = (int)floor((p1 - rtot) / hy);
    jmax = (int)ceil((p1 + rtot) / hy);
    kmin = (int)floor((p2 - rtot) / hzed);
    kmax = (int)ceil((p2 + rtot) / hzed);
    if (imin < 1 || imax > (nx - 2) || jmin < 1 || jmax > (ny - 2) || kmin < 1 || kmax > (g.nz - 2)) {
      rc = 0;
      go = false;
    }
  }
  if (go) {
    const double* cf = pc ? pc + kSplCoef * a : nullptr;
    const double deps = (epsw - epsp), depsi = 1.0 / deps;
    for (int i = imin; i <= imax; i++)
      for (int j = jmin; j <= jmax; j++)
        for (int k = kmin; k <= kmax; k++) {
          double gp[3], dHx[3], dHy[3], dHz[3], dHxm[3], dHym[3], dHzm[3];
          gp[0] = (i + 0.5) * hx + g.xmin; gp[1] = j * hy + g.ymin; gp[2] = k * hzed + g.zmin;
          const double Hx = (epsx[IX(i, j, k, nx, ny)] - epsp) * depsi;
          win_grad(ord, cf, gp, win, 0., ap, arad, dHx);
          for (int l = 0; l < 3; l++) dHx[l] *= Hx;
          gp[0] = i * hx + g.xmin; gp[1] = (j + 0.5) * hy + g.ymin; gp[2] = k * hzed + g.zmin;
          const double Hy = (epsy[IX(i, j, k, nx, ny)] - epsp) * depsi;
          win_grad(ord, cf, gp, win, 0., ap, arad, dHy);
          for (int l = 0; l < 3; l++) dHy[l] *= Hy;
          gp[0] = i * hx + g.xmin; gp[1] = j * hy + g.ymin; gp[2] = (k + 0.5) * hzed + g.zmin;
          const double Hz = (epsz[IX(i, j, k, nx, ny)] - epsp) * depsi;
          win_grad(ord, cf, gp, win, 0., ap, arad, dHz);
          for (int l = 0; l < 3; l++) dHz[l] *= Hz;
          gp[0] = (i - 0.5) * hx + g.xmin; gp[1] = j * hy + g.ymin; gp[2] = k * hzed + g.zmin;
          const double Hxm = (epsx[IX(i - 1, j, k, nx, ny)] - epsp) * depsi;
          win_grad(ord, cf, gp, win, 0., ap, arad, dHxm);
          for (int l = 0; l < 3; l++) dHxm[l] *= Hxm;
          gp[0] = i * hx + g.xmin; gp[1] = (j - 0.5) * hy + g.ymin; gp[2] = k * hzed + g.zmin;
          const double Hym = (epsy[IX(i, j - 1, k, nx, ny)] - epsp) * depsi;
          win_grad(ord, cf, gp, win, 0., ap, arad, dHym);
          for (int l = 0; l < 3; l++) dHym[l] *= Hym;
          gp[0] = i * hx + g.xmin; gp[1] = j * hy + g.ymin; gp[2] = (k - 0.5) * hzed + g.zmin;
          const double Hzm = (epsz[IX(i, j, k - 1, nx, ny)] - epsp) * depsi;
          win_grad(ord, cf, gp, win, 0., ap, arad, dHzm);
          for (int l = 0; l < 3; l++) dHzm[l] *= Hzm;
          const size_t p = IX(i, j, k, nx, ny);
          const double fmag = u[p];
          for (int l = 0; l < 3; l++) {
            const double tg = (dHx[l] * (u[p + 1] - u[p]) + dHxm[l] * (u[p - 1] - u[p])) / (hx * hx) +
                              (dHy[l] * (u[p + nx] - u[p]) + dHym[l] * (u[p - nx] - u[p])) / (hy * hy) +
                              (dHz[l] * (u[p + (size_t)nx * ny] - u[p]) +
                               dHzm[l] * (u[p - (size_t)nx * ny] - u[p])) / (hzed * hzed);
            F[l] += (fmag * tg);
          }
        }
    for (int d = 0; d < 3; d++) F[d] = -F[d] * hx * hy * hzed * deps * 0.5 * izmagic;
  }
  out[3 * a] = F[0];
  out[3 * a + 1] = F[1];
  out[3 * a + 2] = F[2];
  rcs[a] = rc;
}

// ------------------------------------------------------------------ maps
__global__ void __launch_bounds__(kB)
    k_map_sample(Geo g, MapDev m, const double* xf, const double* yf, const double* zf, double ox,
                 double oy, double oz, int mode, double scale, double* out, double* part,
                 int* bad) {
  __shared__ double smax[kB];
  const int t = blockIdx.x * kB + threadIdx.x;
  const int k = blockIdx.y;
  const int i = t % g.nx, j = t / g.nx;
  double mx = -1.0;
  if (j < g.ny) {
    double v = 0.0;
    if (!map_value(m, xf[i] + ox, yf[j] + oy, zf[k] + oz, &v)) {
      atomicAdd(bad, 1);
    } else {
      if (mode == 1) {
        if (v < 0.0) atomicAdd(bad + 1, 1);
        mx = v;
      }
      if (out) out[IX(i, j, k, g.nx, g.ny)] = mode == 2 ? v * scale : v;
    }
  }
  if (mode == 1) {   // block maximum, one partial per block
    smax[threadIdx.x] = mx;
    __syncthreads();
    for (int w = kB / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
      __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.y * gridDim.x + blockIdx.x] = smax[0];
  }
}

__global__ void k_kappa_scale(size_t n, double kappamax, double* kappa) {
  const size_t i = (size_t)blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  double tk = kappa[i];
  if (tk < VPMGSMALL) tk = 0.0;
  kappa[i] = (tk / kappamax);
}

}  // namespace

// ================================================================ wrappers

// The no-allocation entry points' buffer (layout in fill_device.h).  In host
// code the name `scratch` is a host-side stand-in, not a device address: only
// the address scratch_address() looks up is passed to kernels and copies.
static __device__ __align__(256) double scratch[kScratchDoubles];

static dim3 grid_all(const Geo& g) {
  return dim3((unsigned)(((long)g.nx * g.ny + kB - 1) / kB), (unsigned)g.nz);
}

static unsigned n_tiles(const Tiles& T) {
  return (unsigned)((long)T.tdim[0] * T.tdim[1] * T.tdim[2]);
}

void charge_map(hipStream_t s, const Geo& g, const AtomsDev& A, const Tiles& T, int mode,
                double zmagic, int, double* charge) {
  k_charge<<<n_tiles(T), kB, 0, s>>>(g, A, T, mode, zmagic, charge);
}

void coef_spline(hipStream_t s, const Geo& g, const AtomsDev& A, const Tiles& T, double win,
                 double irad, double ionmask, double epsw, double epsp, double* kappa,
                 double* epsx, double* epsy, double* epsz, int order, const double* coef) {
  if (order == 3)
    k_coef_spline<3><<<n_tiles(T), kB, 0, s>>>(g, A, T, win, irad, ionmask, epsw, epsp, coef,
                                               kappa, epsx, epsy, epsz);
  else if (order == 4)
    k_coef_spline<4><<<n_tiles(T), kB, 0, s>>>(g, A, T, win, irad, ionmask, epsw, epsp, coef,
                                               kappa, epsx, epsy, epsz);
  else
    k_coef_spline<2><<<n_tiles(T), kB, 0, s>>>(g, A, T, win, irad, ionmask, epsw, epsp, nullptr,
                                               kappa, epsx, epsy, epsz);
}

void sas_flags(hipStream_t s, const AtomsDev& A, const Cells& C, const double* ref, int nref,
               double probe, unsigned char* flags) {
  if (A.n <= 0) return;
  k_sas_flags<<<(unsigned)A.n, kB, 0, s>>>(A, C, ref, nref, probe, flags);
}

static void sas_check(hipError_t e) {
  if (e != hipSuccess) {
    fprintf(stderr, "sas_compact: %s\n", hipGetErrorString(e));
    abort();
  }
}

int sas_compact(hipStream_t s, const AtomsDev& A, const double* ref, int nref, double probe,
                const unsigned char* flags, int* scratch, double** xyz_out,
                double* (*alloc)(size_t)) {
  const long n = (long)A.n * nref;
  k_sas_count<<<kSasParts, kB, 0, s>>>(flags, n, scratch);
  std::vector<int> cnt(kSasParts), off(kSasParts);
  sas_check(hipMemcpyAsync(cnt.data(), scratch, kSasParts * sizeof(int), hipMemcpyDeviceToHost,
                           s));
  sas_check(hipStreamSynchronize(s));
  int tot = 0;
  for (int b = 0; b < kSasParts; b++) {
    off[b] = tot;
    tot += cnt[b];
  }
  double* xyz = alloc(3 * (size_t)(tot > 0 ? tot : 1));
  sas_check(hipMemcpyAsync(scratch, off.data(), kSasParts * sizeof(int), hipMemcpyHostToDevice,
                           s));
  k_sas_compact<<<kSasParts, kB, 0, s>>>(A, ref, nref, probe, flags, scratch, xyz);
  sas_check(hipStreamSynchronize(s));
  *xyz_out = xyz;
  return tot;
}

void mark_flags(hipStream_t s, const Geo& g, int npts, const double* pos, const double* rad,
                const MarkTargets& T, unsigned char* flags) {
  if (npts <= 0 || T.nt <= 0) return;
  k_mark_flags<<<(npts + kMarkWaves - 1) / kMarkWaves, kB, 0, s>>>(g, npts, pos, rad, T, flags);
}

void mol_maps(hipStream_t s, const Geo& g, const unsigned char* atom_flags,
              const unsigned char* sas_flags, const unsigned char* kappa_flags, double epsp,
              double epsw, double ionmask, int smooth, double* epsx, double* epsy, double* epsz,
              double* kappa, unsigned char* bits) {
  k_solute_bits<<<grid_all(g), kB, 0, s>>>(g, atom_flags, sas_flags, bits);
  k_mol_maps<<<grid_all(g), kB, 0, s>>>(g, bits, kappa_flags, epsp, epsw, ionmask, smooth, epsx,
                                        epsy, epsz, kappa);
}

int bc_mdh_chunks(int natoms) { return natoms >= 4096 ? 16 : (natoms >= 512 ? natoms / 256 : 1); }

size_t bc_mdh_scratch(const Geo& g, int natoms) {
  size_t mx = (size_t)g.ny * g.nz;
  if ((size_t)g.nx * g.nz > mx) mx = (size_t)g.nx * g.nz;
  if ((size_t)g.nx * g.ny > mx) mx = (size_t)g.nx * g.ny;
  // partial sums, then the per-atom factors (FP64) or float4 atoms (two doubles each)
  return (size_t)bc_mdh_chunks(natoms) * 6 * mx + 2 * (size_t)(natoms > 0 ? natoms : 1);
}

void bc_mdh(hipStream_t s, const Geo& g, const double* xf, const double* yf, const double* zf,
            const AtomsDev& A, double pre1, double xkappa, double* part, double* gx, double* gy,
            double* gz, bool fp32_terms) {
  int mx = g.ny * g.nz;
  if (g.nx * g.nz > mx) mx = g.nx * g.nz;
  if (g.nx * g.ny > mx) mx = g.nx * g.ny;
  const int nch = bc_mdh_chunks(A.n);
  // the per-atom data follow the partial sums in the scratch buffer (6 * mx
  // * nch doubles: an even count, so the float4 array is 16-byte aligned)
  double* fac = part + (size_t)nch * 6 * mx;
  const dim3 grid((mx + kB - 1) / kB, 6, nch);
  if (fp32_terms) {
    const double cx = 0.5 * (g.xmin + g.xmax), cy = 0.5 * (g.ymin + g.ymax),
                 cz = 0.5 * (g.zmin + g.zmax);
    float4* at = reinterpret_cast<float4*>(fac);
    if (A.n > 0)
      k_bc_mdh_atoms32<<<(A.n + kB - 1) / kB, kB, 0, s>>>(A, pre1, xkappa, cx, cy, cz, at);
    const float nk2 = xkappa > VSMALL ? (float)(-xkappa * 1.4426950408889634) : 0.0f;
    k_bc_mdh32<<<grid, kB, 0, s>>>(g, xf, yf, zf, cx, cy, cz, A.n, at, nk2, nch, mx, part);
  } else {
    if (A.n > 0) k_bc_mdh_atoms<<<(A.n + kB - 1) / kB, kB, 0, s>>>(A, pre1, xkappa, fac);
    k_bc_mdh<<<grid, kB, 0, s>>>(g, xf, yf, zf, A, fac, xkappa, nch, mx, part);
  }
  k_bc_mdh_sum<<<dim3((mx + kB - 1) / kB, 6), kB, 0, s>>>(g, nch, mx, part, gx, gy, gz);
}

void bc_sdh(hipStream_t s, const Geo& g, const double* xf, const double* yf, const double* zf,
            const SdhMoments& M, double* gx, double* gy, double* gz) {
  int mx = g.ny * g.nz;
  if (g.nx * g.nz > mx) mx = g.nx * g.nz;
  if (g.nx * g.ny > mx) mx = g.nx * g.ny;
  k_bc_sdh<<<dim3((mx + kB - 1) / kB, 6), kB, 0, s>>>(g, xf, yf, zf, M, gx, gy, gz);
}

void focus_bound(hipStream_t s, const Geo& g, double xminN, double yminN, double zminN,
                 double xmaxN, double ymaxN, double zmaxN, const OldGrid& o, double* gx,
                 double* gy, double* gz, int* bad) {
  int mx = g.ny * g.nz;
  if (g.nx * g.nz > mx) mx = g.nx * g.nz;
  if (g.nx * g.ny > mx) mx = g.nx * g.ny;
  k_focus<<<dim3((mx + kB - 1) / kB, 6), kB, 0, s>>>(g, xminN, yminN, zminN, xmaxN, ymaxN, zmaxN, o,
                                                     gx, gy, gz, bad);
}

void set_part(hipStream_t s, const Geo& g, double xcen0, double ycen0, double zcen0,
              const double lo[3], const double hi[3], int invert, double* pvec) {
  k_set_part<<<grid_all(g), kB, 0, s>>>(g, xcen0, ycen0, zcen0, lo[0], lo[1], lo[2], hi[0], hi[1],
                                        hi[2], invert, pvec);
}

void energy_qf_volume(hipStream_t s, size_t n, const double* pvec, const double* u,
                      const double* charge, double* part, double* out) {
  k_e_qf<<<kParts, kB, 0, s>>>(n, pvec, u, charge, part);
  k_fin<<<1, kB, 0, s>>>(part, out);
}

void energy_diel(hipStream_t s, const Geo& g, const double* pvec, const double* u,
                 const double* ex, const double* ey, const double* ez, double* part, double* out) {
  k_e_diel<<<kParts, kB, 0, s>>>(g, pvec, u, ex, ey, ez, part);
  k_fin<<<1, kB, 0, s>>>(part, out);
}

void energy_qm(hipStream_t s, size_t n, int nonlin, const double* pvec, const double* kappa,
               const double* u, double zks2, double zkappa2, const IonBulk& ions, double* part,
               double* out, int* nchop) {
  k_zero_int<<<1, 1, 0, s>>>(nchop);
  k_e_qm<<<kParts, kB, 0, s>>>(n, nonlin, pvec, kappa, u, zks2, zkappa2, ions, part, nchop);
  k_fin<<<1, kB, 0, s>>>(part, out);
}

void energy_qf_point(hipStream_t s, const Geo& g, const AtomsDev& A, const double* u, double* part,
                     double* out) {
  k_e_qf_point<<<(A.n + kB - 1) / kB, kB, 0, s>>>(g, A, u, part);
  k_seq_sum<<<1, 64, 0, s>>>(part, A.n, out);
}

void force_qf(hipStream_t s, const Geo& g, const AtomsDev& A, int mode, const double* u,
              double* out) {
  k_force_qf<<<(A.n + 63) / 64, 64, 0, s>>>(g, A, mode, u, out);
}

void force_ib(hipStream_t s, const Geo& g, const AtomsDev& A, double win, double irad,
              double zkappa2, double izmagic, int nonlin, double ionstr, const IonBulk& ions,
              const double* kappa, const double* u, double* out, int order, const double* coef) {
  k_force_ib<<<(A.n + 63) / 64, 64, 0, s>>>(g, A, win, irad, zkappa2, izmagic, nonlin, ionstr, ions,
                                            kappa, u, out, order, coef);
}

void force_db(hipStream_t s, const Geo& g, const AtomsDev& A, double win, double srad, double epsp,
              double epsw, double izmagic, const double* epsx, const double* epsy,
              const double* epsz, const double* u, double* out, int* rc, int order,
              const double* coef) {
  k_force_db<<<(A.n + 63) / 64, 64, 0, s>>>(g, A, win, srad, epsp, epsw, izmagic, epsx, epsy, epsz,
                                            u, out, rc, order, coef);
}

// Window coefficients of one (b, e) = (r - win, r + win) pair: fillcoCoefSpline3
// (vpmg.c:7597-7619) for order 3, fillcoCoefSpline4 (vpmg.c:7124-7147) for 4.
static void window_coef(int order, double b, double e, double* c) {
  for (int n = 0; n < 8; n++) c[n] = 0.0;
  if (order == 3) {
    const double e2 = e * e, e3 = e2 * e, e4 = e3 * e, e5 = e4 * e;
    const double b2 = b * b, b3 = b2 * b, b4 = b3 * b, b5 = b4 * b;
    (void)e4;
    (void)e5;
    const double denom = pow((e - b), 5.0);
    c[0] = -10.0 * e2 * b3 + 5.0 * e * b4 - b5;
    c[1] = 30.0 * e2 * b2;
    c[2] = -30.0 * (e2 * b + e * b2);
    c[3] = 10.0 * (e2 + 4.0 * e * b + b2);
    c[4] = -15.0 * (e + b);
    c[5] = 6;
    for (int n = 0; n < 6; n++) c[n] = c[n] / denom;
  } else {
    const double e2 = e * e, e3 = e2 * e, e4 = e3 * e, e5 = e4 * e, e6 = e5 * e, e7 = e6 * e;
    const double b2 = b * b, b3 = b2 * b, b4 = b3 * b, b5 = b4 * b, b6 = b5 * b, b7 = b6 * b;
    const double denom = e7 - 7.0 * b * e6 + 21.0 * b2 * e5 - 35.0 * e4 * b3 + 35.0 * e3 * b4 -
                         21.0 * b5 * e2 + 7.0 * e * b6 - b7;
    c[0] = b4 * (35.0 * e3 - 21.0 * b * e2 + 7 * e * b2 - b3) / denom;
    c[1] = -140.0 * b3 * e3 / denom;
    c[2] = 210.0 * e2 * b2 * (e + b) / denom;
    c[3] = -140.0 * e * b * (e2 + 3.0 * b * e + b2) / denom;
    c[4] = 35.0 * (e3 + 9.0 * b * e2 + +9.0 * e * b2 + b3) / denom;
    c[5] = -84.0 * (e2 + 3.0 * b * e + b2) / denom;
    c[6] = 70.0 * (e + b) / denom;
    c[7] = -20.0 / denom;
  }
}

void spline_coefficients(int order, int natoms, const double* radius, double win, double irad,
                         double* out) {
  for (int a = 0; a < natoms; a++) {
    const double arad = radius[a];
    double* o = out + (size_t)kSplCoefHost * a;
    window_coef(order, arad - win, arad + win, o);
    // the ion window of the kappa map: b, e = irad + arad -+ win
    if (order == 3) {
      const double b = irad + arad - win, e = irad + arad + win;
      const double denom = pow((e - b), 5.0);
      for (int n = 0; n < 8; n++) o[8 + n] = n < 6 ? o[n] / denom : 0.0;
    } else {
      window_coef(order, irad + arad - win, irad + arad + win, o + 8);
    }
    // the ion-boundary force gradient: (arad + irad) -+ win
    const double ar = arad + irad;
    window_coef(order, ar - win, ar + win, o + 16);
  }
}

size_t map_sample_parts(const Geo& g) {
  const dim3 gr = grid_all(g);
  return (size_t)gr.x * gr.y;
}

void map_sample(hipStream_t s, const Geo& g, const MapDev& m, const double* xf, const double* yf,
                const double* zf, double ox, double oy, double oz, int mode, double scale,
                double* out, double* part, int* bad) {
  k_map_sample<<<grid_all(g), kB, 0, s>>>(g, m, xf, yf, zf, ox, oy, oz, mode, scale, out, part,
                                          bad);
}

double* scratch_address() {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static double* addr[kMaxDevices] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      fprintf(stderr, "scratch_address: hipGetSymbolAddress failed on device %d: %s\n", dev,
              hipGetErrorString(e));
      return nullptr;
    }
    addr[dev] = (double*)p;
  }
  return addr[dev];
}

void kappa_scale(hipStream_t s, size_t n, double kappamax, double* kappa) {
  k_kappa_scale<<<nb(n), kB, 0, s>>>(n, kappamax, kappa);
}

}  // namespace pmg
