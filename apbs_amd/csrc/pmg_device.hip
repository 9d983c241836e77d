// pmg_device.hip -- PMG multigrid kernels for CDNA4 (gfx950).
//
// The fine 7-point kernels evaluate exactly the reference's expressions in the
// reference's order (src/pmgc/gsd.c:95-165, matvecd.c:461-496) and are built
// without FMA contraction, so one sweep is bitwise identical to the reference
// arithmetic.  Coarse 27-point levels are smoothed with an 8-colour
// Gauss-Seidel ordering (parity of i, j, k) instead of the reference's
// in-place lexicographic sweep (gsd.c:230-327), whose dependencies are serial.
#include "pmg_device.h"

#include <type_traits>

#include <cstdio>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <utility>

namespace pmg {

namespace {

constexpr int kBlock = 256;

// Kernel-argument container of a batched launch: one argument set per
// calculation (at most kMaxBatch).
template <class P>
struct Batch {
  P v[kMaxBatch];
};

__host__ __device__ inline size_t IX(int i, int j, int k, int nx, int ny) {
  return ((size_t)k * ny + j) * nx + i;
}

inline dim3 grid_interior(int nx, int ny, int nz) {
  long cnt = (long)(nx - 2) * (ny - 2);
  return dim3((unsigned)((cnt + kBlock - 1) / kBlock), (unsigned)(nz - 2), 1);
}

// ------------------------------------------------------------------ 7-point
// fc + north + south + east + west + down + up, left to right.
template <class T>
__device__ inline T acc7(T init, const T* __restrict__ oE, const T* __restrict__ oN,
                         const T* __restrict__ uC, const T* __restrict__ x, size_t p, int nx,
                         size_t nxy) {
  T s = init;
  s += oN[p] * x[p + nx];
  s += oN[p - nx] * x[p - nx];
  s += oE[p] * x[p + 1];
  s += oE[p - 1] * x[p - 1];
  s += uC[p - nxy] * x[p - nxy];
  s += uC[p] * x[p + nxy];
  return s;
}

// One colour of the red-black sweep: points with (i+j+k)&1 == par (0-based).
// The reference's "red" points (1-based i+j+k even) have par == 1.
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_gsrb7(int nx, int ny, int nz, const T* __restrict__ oE, const T* __restrict__ oN,
            const T* __restrict__ uC, const T* __restrict__ D, const T* __restrict__ fc, T* x,
            int par) {
  const int npair = (nx - 1) / 2;
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const int k = blockIdx.y + 1;
  const int ip = t % npair;
  const int j = 1 + t / npair;
  if (j > ny - 2) return;
  const int i = 1 + (((1 + j + k) ^ par) & 1) + 2 * ip;
  if (i > nx - 2) return;
  const size_t nxy = (size_t)nx * ny;
  const size_t p = IX(i, j, k, nx, ny);
  x[p] = acc7(fc[p], oE, oN, uC, x, p, nx, nxy) / D[p];
}

// One complete red-black sweep, out of place: xo = sweep(xi).  In the 7-point
// stencil every neighbour of a point has the other colour, so the first colour
// F reads only old values of the second colour S, and S reads only new F
// values.  A workgroup owns a TX x TY tile of (i, j) and marches over a chunk
// of k planes: the new F values of plane k+1 (tile plus a one-point halo) go
// to a three-plane LDS ring, then the S points of plane k are updated from the
// ring and the whole tile row of plane k is written.  Each value is the
// in-place sweep's value bit for bit (same terms, same order, no FMA).
// Points outside the interior are not written.
constexpr int kFT = 256, kXCD = 8;

// kTX columns (even), kTY rows and kKC planes per workgroup of kNT threads.
// The arrays of one calculation in a fused-sweep launch.
template <class T>
struct SweepArgs {
  const T *oE, *oN, *uC, *D, *fc, *xi;
  T* xo;
};

// kZeroX: xi is zero everywhere (a solve of the correction equation from a
// zero start); its values are not read, and the sweep's values are those of
// reading the zeros.
template <class T, int kTY, int kKC, int kTX, int kNT, bool kZeroX>
__device__ __forceinline__ void fused_body(int nx, int ny, int nz, int tile,
                                           const T* __restrict__ oE, const T* __restrict__ oN,
                                           const T* __restrict__ uC, const T* __restrict__ D,
                                           const T* __restrict__ fc, const T* __restrict__ xi,
                                           T* __restrict__ xo, int first) {
  constexpr int RX = kTX + 2, RY = kTY + 2, RN = RX * RY;
  constexpr int FPR = RX / 2, FN = FPR * RY;  // F points per region row / per plane
  __shared__ T ring[3][RN];
  const int tid = threadIdx.x;
  const int nbx = (nx - 2 + kTX - 1) / kTX, nby = (ny - 2 + kTY - 1) / kTY;
  const int i0 = 1 + (tile % nbx) * kTX, j0 = 1 + ((tile / nbx) % nby) * kTY;
  const int k0 = 1 + (tile / (nbx * nby)) * kKC;
  const int k1 = min(k0 + kKC, nz - 1);  // planes k0 .. k1-1
  const size_t nxy = (size_t)nx * ny;
  const int F = first;                   // parity of the first colour

  // New F values of plane kk over the tile and its halo, one thread per F
  // point in kFR rounds.  loadF issues every round's loads at once into
  // registers; storeF computes the values (acc7's terms and order) into the
  // plane's LDS slot.
  constexpr int kFR = (FN + kNT - 1) / kNT;
  struct FRegs {
    T v[kFR][14];
    int slot[kFR];  // LDS index in the ring plane
    int kind[kFR];  // 0 nothing, 1 boundary copy, 2 interior update
  };
  auto loadF = [&](int kk, FRegs& R) {
#pragma unroll
    for (int u = 0; u < kFR; u++) {
      R.kind[u] = 0;
      const int f = tid + u * kNT;
      if (f >= FN) continue;
      const int row = f / FPR, c = f % FPR;
      const int j = j0 - 1 + row;
      const int ib = i0 - 1;
      const int i = ib + ((F - ib - j - kk) & 1) + 2 * c;
      if (i > nx - 1 || j > ny - 1) continue;
      const size_t p = ((size_t)kk * ny + j) * nx + i;
      R.slot[u] = row * RX + (i - ib);
      if (i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2 && kk >= 1 && kk <= nz - 2) {
        R.kind[u] = 2;
        R.v[u][0] = fc[p];
        R.v[u][1] = oN[p];
        R.v[u][2] = kZeroX ? T(0) : xi[p + nx];
        R.v[u][3] = oN[p - nx];
        R.v[u][4] = kZeroX ? T(0) : xi[p - nx];
        R.v[u][5] = oE[p];
        R.v[u][6] = kZeroX ? T(0) : xi[p + 1];
        R.v[u][7] = oE[p - 1];
        R.v[u][8] = kZeroX ? T(0) : xi[p - 1];
        R.v[u][9] = uC[p - nxy];
        R.v[u][10] = kZeroX ? T(0) : xi[p - nxy];
        R.v[u][11] = uC[p];
        R.v[u][12] = kZeroX ? T(0) : xi[p + nxy];
        R.v[u][13] = D[p];
      } else {
        R.kind[u] = 1;
        R.v[u][0] = kZeroX ? T(0) : xi[p];
      }
    }
  };
  auto storeF = [&](int kk, const FRegs& R) {
    T* r = ring[kk % 3];
#pragma unroll
    for (int u = 0; u < kFR; u++) {
      if (R.kind[u] == 2) {
        T acc = R.v[u][0];
        acc += R.v[u][1] * R.v[u][2];
        acc += R.v[u][3] * R.v[u][4];
        acc += R.v[u][5] * R.v[u][6];
        acc += R.v[u][7] * R.v[u][8];
        acc += R.v[u][9] * R.v[u][10];
        acc += R.v[u][11] * R.v[u][12];
        r[R.slot[u]] = acc / R.v[u][13];
      } else if (R.kind[u] == 1) {
        r[R.slot[u]] = R.v[u][0];
      }
    }
  };

  // The S step: thread u owns the (i, i+1) pair u % 32 of tile row u / 32 (one
  // F and one S point).  The eight operands of the S point of plane kk are
  // loaded one plane ahead, together with stageF(kk), so each coefficient
  // line is fetched once for both colours while it is still cached.
  constexpr int kSP = (kTX / 2 * kTY + kNT - 1) / kNT;
  T cur[kSP][8], nxt[kSP][8];
  auto loadS = [&](int kk, T (&r)[kSP][8]) {
#pragma unroll
    for (int u = 0; u < kSP; u++) {
      const int sidx = tid + u * kNT;
      if (sidx >= kTX / 2 * kTY) continue;
      const int j = j0 + sidx / (kTX / 2);
      const int is = i0 + 2 * (sidx % (kTX / 2)) + ((1 - F - i0 - j - kk) & 1);
      if (is > nx - 2 || j > ny - 2) continue;
      const size_t p = ((size_t)kk * ny + j) * nx + is;
      r[u][0] = fc[p];
      r[u][1] = oN[p];
      r[u][2] = oN[p - nx];
      r[u][3] = oE[p];
      r[u][4] = oE[p - 1];
      r[u][5] = uC[p - nxy];
      r[u][6] = uC[p];
      r[u][7] = D[p];
    }
  };

  FRegs fr;
  loadS(k0, cur);
  loadF(k0 - 1, fr);
  storeF(k0 - 1, fr);
  loadF(k0, fr);
  storeF(k0, fr);
  loadF(k0 + 1, fr);
  for (int k = k0; k < k1; k++) {
    // F(k+1) from registers loaded one step earlier; the loads of F(k+2) and
    // of S(k+1)'s operands are in flight while S(k) is computed
    if (k + 1 < k1) loadS(k + 1, nxt);
    storeF(k + 1, fr);
    __syncthreads();
    if (k + 1 < k1) loadF(k + 2, fr);
    const T* rm = ring[(k - 1) % 3];
    const T* r0 = ring[k % 3];
    const T* rp = ring[(k + 1) % 3];
#pragma unroll
    for (int u = 0; u < kSP; u++) {
      const int sidx = tid + u * kNT;
      if (sidx >= kTX / 2 * kTY) continue;
      const int jj = sidx / (kTX / 2);
      const int j = j0 + jj;
      if (j > ny - 2) continue;
      const int ia = i0 + 2 * (sidx % (kTX / 2));
      const int is = ia + ((1 - F - i0 - j - k) & 1);
      const int iF = ia + ((F - i0 - j - k) & 1);
      const size_t row = ((size_t)k * ny + j) * nx;
      if (iF <= nx - 2) xo[row + iF] = r0[(jj + 1) * RX + (iF - i0 + 1)];
      if (is <= nx - 2) {
        const int e = (jj + 1) * RX + (is - i0 + 1);
        T acc = cur[u][0];
        acc += cur[u][1] * r0[e + RX];
        acc += cur[u][2] * r0[e - RX];
        acc += cur[u][3] * r0[e + 1];
        acc += cur[u][4] * r0[e - 1];
        acc += cur[u][5] * rm[e];
        acc += cur[u][6] * rp[e];
        xo[row + is] = acc / cur[u][7];
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kSP; u++)
#pragma unroll
      for (int w = 0; w < 8; w++) cur[u][w] = nxt[u][w];
  }
}

// Workgroups are dealt round-robin to the 8 XCDs: XCD x gets the (calculation,
// tile) items [x*per, (x+1)*per) in calculation-major, z-major order, so tiles
// that share halo rows and planes share an L2.
template <class T, int kTY, int kKC, int kTX, int kNT, bool kZeroX = false>
__global__ void __launch_bounds__(kNT)
    k_gsrb7_fused(int nx, int ny, int nz, Batch<SweepArgs<T>> B, int nreq, int first) {
  const int nbx = (nx - 2 + kTX - 1) / kTX, nby = (ny - 2 + kTY - 1) / kTY;
  const int nb = nbx * nby * ((nz - 2 + kKC - 1) / kKC);
  const int total = nb * nreq;
  const int per = (total + kXCD - 1) / kXCD;
  const int item = (blockIdx.x % kXCD) * per + blockIdx.x / kXCD;
  if (item >= total) return;
  const SweepArgs<T>& a = B.v[item / nb];
  fused_body<T, kTY, kKC, kTX, kNT, kZeroX>(nx, ny, nz, item % nb, a.oE, a.oN, a.uC, a.D, a.fc,
                                            a.xi, a.xo, first);
}

// The vector form of the fused sweep.  The scalar form above loads one
// 4- or 8-byte word per lane and operand with the lanes two words apart, so
// the texture addresser (four lanes per clock) and the L1 limit it long
// before memory does.  Here a lane owns a run of kV consecutive elements of
// one row (16 bytes: 4 floats or 2 doubles, both colours) and loads every
// operand as a 16-byte vector; the coefficient vectors that the F step of
// plane k loads are the S step's operands of plane k one step later, so the
// S step loads nothing from memory.  Values and their order are the scalar
// form's (acc7's terms, no FMA).
//
// Lanes: region row r (0 .. kTY+1, row j = j0-1+r) and run c of that row
// (columns i0-1+kV*c ..); a run is moved inwards to [1, nx-1-kV] when it would
// cross the first or last interior column, so every load stays inside the
// row and runs near the grid edge may overlap (they then write the same
// values).  The boundary columns 0 and nx-1 come from the x(p-1) / x(p+1)
// vectors of the runs beside them.
template <class T, int N>
__device__ __forceinline__ void vld16(T (&v)[N], const T* p) {
  static_assert(sizeof(T) * N == 16 || sizeof(T) * N == 8, "8- or 16-byte vectors");
  typedef T vT __attribute__((ext_vector_type(N)));
  vT a;
  __builtin_memcpy(&a, p, sizeof(T) * N);
#pragma unroll
  for (int i = 0; i < N; i++) v[i] = a[i];
}

template <class T, int N>
__device__ __forceinline__ void vst16(T* p, const T (&v)[N]) {
  static_assert(sizeof(T) * N == 16 || sizeof(T) * N == 8, "8- or 16-byte vectors");
  typedef T vT __attribute__((ext_vector_type(N)));
  vT a;
#pragma unroll
  for (int i = 0; i < N; i++) a[i] = v[i];
  __builtin_memcpy(p, &a, sizeof(T) * N);
}

template <class T, int kTY, int kKC, int kTX, int kNT, bool kZeroX>
__device__ __forceinline__ void vfused_body(int nx, int ny, int nz, int tile,
                                            const T* __restrict__ oE, const T* __restrict__ oN,
                                            const T* __restrict__ uC, const T* __restrict__ D,
                                            const T* __restrict__ fc, const T* __restrict__ xi,
                                            T* __restrict__ xo, int first) {
  constexpr int kV = 16 / (int)sizeof(T), kH = kV / 2;
  constexpr int RX = kTX + 2, RY = kTY + 2, RN = RX * RY;
  constexpr int NSEG = (RX + kV - 1) / kV;
  static_assert(RY * NSEG <= kNT, "one lane per run");
  __shared__ T ring[3][RN];
  const int tid = threadIdx.x;
  const int nbx = (nx - 2 + kTX - 1) / kTX, nby = (ny - 2 + kTY - 1) / kTY;
  const int i0 = 1 + (tile % nbx) * kTX, j0 = 1 + ((tile / nbx) % nby) * kTY;
  const int k0 = 1 + (tile / (nbx * nby)) * kKC;
  const int k1 = min(k0 + kKC, nz - 1);  // planes k0 .. k1-1
  const size_t nxy = (size_t)nx * ny;
  const int F = first;

  const bool lane = tid < RY * NSEG;
  const int r = lane ? tid / NSEG : 0, c = lane ? tid % NSEG : 0;
  const int j = j0 - 1 + r;
  const bool rowin = lane && j <= ny - 1;
  const bool rowint = j >= 1 && j <= ny - 2;
  const bool srow = lane && r >= 1 && r <= kTY && j <= ny - 2;
  const int jr = min(j, ny - 1), jc = min(max(j, 1), ny - 2);
  const int is = i0 - 1 + kV * c;
  const int ib = min(max(is, 1), nx - 1 - kV);
  const int chi = min(i0 + kTX - 1, nx - 2);  // last column this tile writes
  // a lane holds, per plane: fc, oN, oN(-nx), oE, oE(-1), uC(-nxy), uC, D and
  // x(+nx), x(-nx), x(+1), x(-1), x(+nxy), x(-nxy), each at columns ib ..
  struct Set {
    T a[8][kV];
    T x[6][kV];
  };
  auto load = [&](int kk, Set& S) {
    const int kc = min(max(kk, 1), nz - 2);
    const size_t q = ((size_t)kc * ny + jc) * nx + ib;   // interior row / plane
    const size_t pr = ((size_t)kk * ny + jr) * nx + ib;  // the lane's own row and plane
    vld16(S.a[0], fc + q);
    vld16(S.a[1], oN + q);
    vld16(S.a[2], oN + q - nx);
    vld16(S.a[3], oE + q);
    vld16(S.a[4], oE + q - 1);
    vld16(S.a[5], uC + q - nxy);
    vld16(S.a[6], uC + q);
    vld16(S.a[7], D + q);
    if (!kZeroX) {
      vld16(S.x[0], xi + q + nx);
      vld16(S.x[1], xi + q - nx);
      vld16(S.x[2], xi + pr + 1);
      vld16(S.x[3], xi + pr - 1);
      vld16(S.x[4], xi + q + nxy);
      vld16(S.x[5], xi + q - nxy);
    } else {
#pragma unroll
      for (int v = 0; v < 6; v++)
#pragma unroll
        for (int e = 0; e < kV; e++) S.x[v][e] = T(0);
    }
  };
  // New F values of plane kk into the ring (and fv: the lane's F elements
  // e0, e0+2, ..., with e0 the first element of F colour)
  auto stageF = [&](int kk, const Set& S, T (&fv)[kH]) {
    if (!rowin) return;
    T* rg = ring[kk % 3] + r * RX - (i0 - 1);
    const bool inner = rowint && kk >= 1 && kk <= nz - 2;
    const int e0 = (F - ib - j - kk) & 1;
#pragma unroll
    for (int m = 0; m < kH; m++) {
      // (selects between values, not addresses, so the vectors stay in registers)
      auto pk = [&](const T(&v)[kV]) {
        const T lo = v[2 * m], hi = v[2 * m + 1];
        return e0 ? hi : lo;
      };
      T val;
      if (inner) {
        T acc = pk(S.a[0]);
        acc += pk(S.a[1]) * pk(S.x[0]);
        acc += pk(S.a[2]) * pk(S.x[1]);
        acc += pk(S.a[3]) * pk(S.x[2]);
        acc += pk(S.a[4]) * pk(S.x[3]);
        acc += pk(S.a[5]) * pk(S.x[5]);
        acc += pk(S.a[6]) * pk(S.x[4]);
        val = acc / pk(S.a[7]);
      } else {
        // a boundary row or plane keeps its values: x at element e is
        // x(p-1)[e+1], or x(p+1)[e-1] for the last element
        const T xe = 2 * m + 2 < kV ? S.x[3][(2 * m + 2) % kV] : S.x[2][2 * m];
        const T xo1 = S.x[3][2 * m + 1];
        val = e0 ? xe : xo1;
      }
      fv[m] = val;
      const int col = ib + e0 + 2 * m;
      if (col >= i0 - 1 && col <= i0 + kTX) rg[col] = val;
    }
    // boundary columns 0 and nx-1 when they are in the region and of F colour
    if (is <= 0 && ((j + kk) & 1) == F) rg[0] = S.x[3][0];
    if (ib == nx - 1 - kV && nx - 1 <= i0 + kTX && ((nx - 1 + j + kk) & 1) == F)
      rg[nx - 1] = S.x[2][kV - 1];
  };
  // S values of plane k from the coefficient vectors of plane k and the ring;
  // both colours of the lane's tile columns are written
  auto stageS = [&](int k, const Set& S, const T(&fv)[kH]) {
    if (!srow) return;
    const int off = r * RX - (i0 - 1);
    const T* rm = ring[(k - 1) % 3] + off;
    const T* r0 = ring[k % 3] + off;
    const T* rp = ring[(k + 1) % 3] + off;
    const int e0 = (F - ib - j - k) & 1;
    T* row = xo + ((size_t)k * ny + j) * nx;
    // a run wholly inside the tile's columns is written with one 16-byte
    // store; the runs at the tile's edges element by element
    const bool full = ib >= i0 && ib + kV - 1 <= chi;
    T out[kV];
#pragma unroll
    for (int m = 0; m < kH; m++) {
      const int cf = ib + e0 + 2 * m;
      const int cs = ib + (1 - e0) + 2 * m;
      T sv = T(0);
      if (full || (cs >= i0 && cs <= chi)) {
        auto pk = [&](const T(&v)[kV]) {
          const T lo = v[2 * m], hi = v[2 * m + 1];
          return e0 ? lo : hi;
        };
        T acc = pk(S.a[0]);
        acc += pk(S.a[1]) * r0[cs + RX];
        acc += pk(S.a[2]) * r0[cs - RX];
        acc += pk(S.a[3]) * r0[cs + 1];
        acc += pk(S.a[4]) * r0[cs - 1];
        acc += pk(S.a[5]) * rm[cs];
        acc += pk(S.a[6]) * rp[cs];
        sv = acc / pk(S.a[7]);
      }
      const T f = fv[m];
      out[2 * m] = e0 ? sv : f;
      out[2 * m + 1] = e0 ? f : sv;
      if (!full) {
        if (cf >= i0 && cf <= chi) row[cf] = f;
        if (cs >= i0 && cs <= chi) row[cs] = sv;
      }
    }
    if (full) vst16(row + ib, out);
  };
  // one plane step: F(k+1) from n (loaded one step earlier), then S(k) from
  // c, whose registers then take the loads of plane k+2
  auto step = [&](int k, Set& cur, T(&fc_)[kH], const Set& nxt, T(&fn)[kH]) {
    stageF(k + 1, nxt, fn);
    __syncthreads();
    stageS(k, cur, fc_);
    if (k + 2 <= k1) load(k + 2, cur);
    __syncthreads();
  };

  Set A, B;
  T fa[kH], fb[kH];
  load(k0 - 1, A);
  stageF(k0 - 1, A, fa);
  load(k0, A);
  stageF(k0, A, fa);
  load(k0 + 1, B);
  int k = k0;
  for (; k + 1 < k1; k += 2) {
    step(k, A, fa, B, fb);
    step(k + 1, B, fb, A, fa);
  }
  if (k < k1) step(k, A, fa, B, fb);
}

template <class T, int kTY, int kKC, int kTX, int kNT, bool kZeroX = false>
__global__ void __launch_bounds__(kNT)
    k_gsrb7_vfused(int nx, int ny, int nz, Batch<SweepArgs<T>> B, int nreq, int first) {
  const int nbx = (nx - 2 + kTX - 1) / kTX, nby = (ny - 2 + kTY - 1) / kTY;
  const int nb = nbx * nby * ((nz - 2 + kKC - 1) / kKC);
  const int total = nb * nreq;
  const int per = (total + kXCD - 1) / kXCD;
  const int item = (blockIdx.x % kXCD) * per + blockIdx.x / kXCD;
  if (item >= total) return;
  const SweepArgs<T>& a = B.v[item / nb];
  vfused_body<T, kTY, kKC, kTX, kNT, kZeroX>(nx, ny, nz, item % nb, a.oE, a.oN, a.uC, a.D, a.fc,
                                             a.xi, a.xo, first);
}

// ----------------------------------------------------------------- 27-point
template <class T>
struct P27 {
  T o, u, d;
};

template <class T>
__device__ inline P27<T> part27(const T* const* a, const T* __restrict__ x, size_t p, int nx,
                                size_t nxy) {
  const T *oE = a[1], *oN = a[2], *uC = a[3], *oNE = a[4], *oNW = a[5], *uE = a[6], *uW = a[7],
          *uN = a[8], *uS = a[9], *uNE = a[10], *uNW = a[11], *uSE = a[12], *uSW = a[13];
  P27<T> r;
  T o = oN[p] * x[p + nx];
  o += oN[p - nx] * x[p - nx];
  o += oE[p] * x[p + 1];
  o += oE[p - 1] * x[p - 1];
  o += oNE[p] * x[p + nx + 1];
  o += oNW[p] * x[p + nx - 1];
  o += oNW[p - nx + 1] * x[p - nx + 1];
  o += oNE[p - nx - 1] * x[p - nx - 1];
  const size_t q = p + nxy;
  T u = uC[p] * x[q];
  u += uN[p] * x[q + nx];
  u += uS[p] * x[q - nx];
  u += uE[p] * x[q + 1];
  u += uW[p] * x[q - 1];
  u += uNE[p] * x[q + nx + 1];
  u += uNW[p] * x[q + nx - 1];
  u += uSE[p] * x[q - nx + 1];
  u += uSW[p] * x[q - nx - 1];
  const size_t m = p - nxy;
  T d = uC[m] * x[m];
  d += uS[m + nx] * x[m + nx];
  d += uN[m - nx] * x[m - nx];
  d += uW[m + 1] * x[m + 1];
  d += uE[m - 1] * x[m - 1];
  d += uSW[m + nx + 1] * x[m + nx + 1];
  d += uSE[m + nx - 1] * x[m + nx - 1];
  d += uNW[m - nx + 1] * x[m - nx + 1];
  d += uNE[m - nx - 1] * x[m - nx - 1];
  r.o = o;
  r.u = u;
  r.d = d;
  return r;
}

template <class T>
struct Ops27 {
  const T* a[14];
};

// One colour of the 8-colour sweep: 0-based parities (c&1, (c>>1)&1, (c>>2)&1).
// Latency-bound on the small coarse levels: up to 128 VGPRs, so that all 54
// operands of a point are in flight at once.
template <class T>
__device__ __forceinline__ void gsrb27_body(int nx, int ny, int nz, const Ops27<T>& A,
                                            const T* __restrict__ D, const T* __restrict__ fc,
                                            T* x, int c) {
  const int ci = c & 1, cj = (c >> 1) & 1, ck = (c >> 2) & 1;
  const int i0 = ci ? 1 : 2, j0 = cj ? 1 : 2, k0 = ck ? 1 : 2;
  const int ni = (nx - 2 - i0) / 2 + 1, nj = (ny - 2 - j0) / 2 + 1;
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const int k = k0 + 2 * (int)blockIdx.y;
  if (k > nz - 2 || ni <= 0 || nj <= 0) return;
  const int ti = t % ni, tj = t / ni;
  if (tj >= nj) return;
  const int i = i0 + 2 * ti, j = j0 + 2 * tj;
  if (i > nx - 2 || j > ny - 2) return;
  const size_t nxy = (size_t)nx * ny;
  const size_t p = IX(i, j, k, nx, ny);
  P27<T> s = part27<T>(A.a, x, p, nx, nxy);
  x[p] = (fc[p] + (s.o + s.u + s.d)) / D[p];
}

// One calculation of a 27-point smoothing, residual or damping launch.
template <class T>
struct Args27 {
  Ops27<T> A;
  const T* D;
  const T* f;
  T* x;
  T* y;  // plane form: the neighbour planes' values; residual: the output
  T* z;  // plane form: the output
};

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_gsrb27(int nx, int ny, int nz, Batch<Args27<T>> B, int c) {
  const Args27<T>& a = B.v[blockIdx.z];
  gsrb27_body<T>(nx, ny, nz, a.A, a.D, a.f, a.x, c);
}

// Two colours of that sweep in one launch: A = c and B = c ^ 1, which differ
// only in the parity of i and follow each other in both sweep orders.  An
// A point and a B point are neighbours only along i within one row, and the
// rows of one (j, k) parity do not touch each other, so a wavefront owns a
// row: lane t updates the A point iA0 + 2t, writes it, and takes the new A
// values beside its B point iB0 + 2t from the neighbouring lanes.  Rows of up
// to 128 interior points.  The arithmetic is part27's, term for term.
template <class T>
__device__ inline T lane_value(T v, int src) {
  return __shfl(v, src < 0 ? 0 : (src > 63 ? 63 : src), 64);
}

template <class T>
__device__ __forceinline__ void pair_body(int nx, int ny, int nz, const Ops27<T>& A,
                                          const T* __restrict__ D, const T* __restrict__ fc, T* x,
                                          int pa, int cj, int ck) {
  const T *oE = A.a[1], *oN = A.a[2], *uC = A.a[3], *oNE = A.a[4], *oNW = A.a[5], *uE = A.a[6],
          *uW = A.a[7], *uN = A.a[8], *uS = A.a[9], *uNE = A.a[10], *uNW = A.a[11],
          *uSE = A.a[12], *uSW = A.a[13];
  const int j0 = cj ? 1 : 2, k0 = ck ? 1 : 2;
  const int nj = (ny - 2 - j0) / 2 + 1;
  const int row = (int)((blockIdx.x * (unsigned)kBlock + threadIdx.x) / 64);
  const int lane = (int)(threadIdx.x % 64);
  const int nk = (nz - 2 - k0) / 2 + 1;
  if (row >= nj * nk) return;  // uniform per wavefront
  const int j = j0 + 2 * (row % nj), k = k0 + 2 * (row / nj);
  const size_t nxy = (size_t)nx * ny;
  const size_t rowp = (size_t)k * nxy + (size_t)j * nx;
  const int iA = (pa ? 1 : 2) + 2 * lane, iB = (pa ? 2 : 1) + 2 * lane;
  T a = 0;
  if (iA <= nx - 2) {
    const size_t p = rowp + iA;
    P27<T> r = part27<T>(A.a, x, p, nx, nxy);
    a = (fc[p] + (r.o + r.u + r.d)) / D[p];
    x[p] = a;
  }
  // new A values at iB - 1 and iB + 1: lanes t - 1 and t when A is even,
  // t and t + 1 when A is odd
  const T am = lane_value(a, pa ? lane : lane - 1);
  const T ap = lane_value(a, pa ? lane + 1 : lane);
  if (iB > nx - 2) return;
  const size_t p = rowp + iB;
  const T xm = (iB - 1 >= 1) ? am : x[p - 1];
  const T xp = (iB + 1 <= nx - 2) ? ap : x[p + 1];
  T o = oN[p] * x[p + nx];
  o += oN[p - nx] * x[p - nx];
  o += oE[p] * xp;
  o += oE[p - 1] * xm;
  o += oNE[p] * x[p + nx + 1];
  o += oNW[p] * x[p + nx - 1];
  o += oNW[p - nx + 1] * x[p - nx + 1];
  o += oNE[p - nx - 1] * x[p - nx - 1];
  const size_t q = p + nxy;
  T u = uC[p] * x[q];
  u += uN[p] * x[q + nx];
  u += uS[p] * x[q - nx];
  u += uE[p] * x[q + 1];
  u += uW[p] * x[q - 1];
  u += uNE[p] * x[q + nx + 1];
  u += uNW[p] * x[q + nx - 1];
  u += uSE[p] * x[q - nx + 1];
  u += uSW[p] * x[q - nx - 1];
  const size_t m = p - nxy;
  T d = uC[m] * x[m];
  d += uS[m + nx] * x[m + nx];
  d += uN[m - nx] * x[m - nx];
  d += uW[m + 1] * x[m + 1];
  d += uE[m - 1] * x[m - 1];
  d += uSW[m + nx + 1] * x[m + nx + 1];
  d += uSE[m + nx - 1] * x[m + nx - 1];
  d += uNW[m - nx + 1] * x[m - nx + 1];
  d += uNE[m - nx - 1] * x[m - nx - 1];
  x[p] = (fc[p] + (o + u + d)) / D[p];
}

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_gsrb27_pair(int nx, int ny, int nz, Batch<Args27<T>> B, int pa, int cj, int ck) {
  const Args27<T>& a = B.v[blockIdx.z];
  pair_body<T>(nx, ny, nz, a.A, a.D, a.f, a.x, pa, cj, ck);
}

// The vector form of the colour-pair launch.  In the form above a lane loads
// one word per operand, with the lanes two words apart, so its 54 loads per
// point keep the texture addresser (four lanes per clock) busy for most of
// the launch.  Here a lane owns a run of kV consecutive points of a row (16
// bytes: 4 floats or 2 doubles, both colours): each of the 26 coefficient
// terms, f and D is one 16-byte load, and each of the 9 neighbouring rows of
// x one window of kV + 2 values.  The A points of the run are updated first
// (their row neighbours are B points, still old) and written to a row buffer
// in LDS; after a barrier the B points take their new A neighbours from it.
// The last run of a row is moved inwards so that every load stays inside its
// row; overlapping runs compute and write the same values.  part27's terms in
// its order, so the results are the scalar form's bit for bit.
template <class T, int N>
__device__ __forceinline__ void vldw(T (&w)[N + 2], const T* p) {
  static_assert(sizeof(T) * N == 16 || sizeof(T) * N == 8, "8- or 16-byte vectors");
  typedef T vT __attribute__((ext_vector_type(N)));
  typedef T v2 __attribute__((ext_vector_type(2)));
  vT a;
  v2 b;
  __builtin_memcpy(&a, p, sizeof(T) * N);
  __builtin_memcpy(&b, p + N, 2 * sizeof(T));
#pragma unroll
  for (int i = 0; i < N; i++) w[i] = a[i];
  w[N] = b[0];
  w[N + 1] = b[1];
}

// R runs per row, RB rows per workgroup; LDS holds RB rows of new A values.
template <class T, int kV>
__global__ void __launch_bounds__(kBlock)
    k_gsrb27_vpair(int nx, int ny, int nz, Batch<Args27<T>> B, int pa, int cj, int ck, int R,
                   int RB) {
  constexpr int kH = kV / 2;
  extern __shared__ unsigned char vpair_smem[];
  T* na = reinterpret_cast<T*>(vpair_smem);
  const Args27<T>& ar = B.v[blockIdx.z];
  const T *oE = ar.A.a[1], *oN = ar.A.a[2], *uC = ar.A.a[3], *oNE = ar.A.a[4],
          *oNW = ar.A.a[5], *uE = ar.A.a[6], *uW = ar.A.a[7], *uN = ar.A.a[8], *uS = ar.A.a[9],
          *uNE = ar.A.a[10], *uNW = ar.A.a[11], *uSE = ar.A.a[12], *uSW = ar.A.a[13];
  const T *D = ar.D, *fc = ar.f;
  T* x = ar.x;
  const int j0 = cj ? 1 : 2, k0 = ck ? 1 : 2;
  const int nj = (ny - 2 - j0) / 2 + 1, nk = (nz - 2 - k0) / 2 + 1;
  const int lr = (int)threadIdx.x / R, run = (int)threadIdx.x % R;
  const int row = (int)blockIdx.x * RB + lr;
  const bool act = lr < RB && row < nj * nk;
  const int rr = act ? row : 0;  // idle lanes load row 0 of the class (unused)
  const int j = j0 + 2 * (rr % nj), k = k0 + 2 * (rr / nj);
  const int ib = min(1 + kV * run, nx - 1 - kV);
  const size_t nxy = (size_t)nx * ny;
  const size_t p = (size_t)k * nxy + (size_t)j * nx + ib;
  const size_t q = p + nxy, m = p - nxy;
  T* nrow = na + (size_t)(act ? lr : 0) * nx;

  // coefficient vectors (columns ib .. ib+kV-1 of each term)
  T c[28][kV];
  vld16(c[0], oN + p);
  vld16(c[1], oN + p - nx);
  vld16(c[2], oE + p);
  vld16(c[3], oE + p - 1);
  vld16(c[4], oNE + p);
  vld16(c[5], oNW + p);
  vld16(c[6], oNW + p - nx + 1);
  vld16(c[7], oNE + p - nx - 1);
  vld16(c[8], uC + p);
  vld16(c[9], uN + p);
  vld16(c[10], uS + p);
  vld16(c[11], uE + p);
  vld16(c[12], uW + p);
  vld16(c[13], uNE + p);
  vld16(c[14], uNW + p);
  vld16(c[15], uSE + p);
  vld16(c[16], uSW + p);
  vld16(c[17], uC + m);
  vld16(c[18], uS + m + nx);
  vld16(c[19], uN + m - nx);
  vld16(c[20], uW + m + 1);
  vld16(c[21], uE + m - 1);
  vld16(c[22], uSW + m + nx + 1);
  vld16(c[23], uSE + m + nx - 1);
  vld16(c[24], uNW + m - nx + 1);
  vld16(c[25], uNE + m - nx - 1);
  vld16(c[26], fc + p);
  vld16(c[27], D + p);
  // x windows (columns ib-1 .. ib+kV) of the 9 rows: 0 (j,k), 1 (j+1,k),
  // 2 (j-1,k), 3 (j,k+1), 4 (j+1,k+1), 5 (j-1,k+1), 6 (j,k-1), 7 (j+1,k-1),
  // 8 (j-1,k-1)
  T w[9][kV + 2];
  vldw<T, kV>(w[0], x + p - 1);
  vldw<T, kV>(w[1], x + p + nx - 1);
  vldw<T, kV>(w[2], x + p - nx - 1);
  vldw<T, kV>(w[3], x + q - 1);
  vldw<T, kV>(w[4], x + q + nx - 1);
  vldw<T, kV>(w[5], x + q - nx - 1);
  vldw<T, kV>(w[6], x + m - 1);
  vldw<T, kV>(w[7], x + m + nx - 1);
  vldw<T, kV>(w[8], x + m - nx - 1);

  // the value of element e0 + 2*mm (e0 = 0 or 1): selects between values
  auto point = [&](int e0, int mm, T xm, T xp) {
    auto pc = [&](int t) {
      const T lo = c[t][2 * mm], hi = c[t][2 * mm + 1];
      return e0 ? hi : lo;
    };
    auto pw = [&](int r, int d) {
      const T lo = w[r][2 * mm + d], hi = w[r][2 * mm + 1 + d];
      return e0 ? hi : lo;
    };
    T o = pc(0) * pw(1, 1);
    o += pc(1) * pw(2, 1);
    o += pc(2) * xp;
    o += pc(3) * xm;
    o += pc(4) * pw(1, 2);
    o += pc(5) * pw(1, 0);
    o += pc(6) * pw(2, 2);
    o += pc(7) * pw(2, 0);
    T u = pc(8) * pw(3, 1);
    u += pc(9) * pw(4, 1);
    u += pc(10) * pw(5, 1);
    u += pc(11) * pw(3, 2);
    u += pc(12) * pw(3, 0);
    u += pc(13) * pw(4, 2);
    u += pc(14) * pw(4, 0);
    u += pc(15) * pw(5, 2);
    u += pc(16) * pw(5, 0);
    T d = pc(17) * pw(6, 1);
    d += pc(18) * pw(7, 1);
    d += pc(19) * pw(8, 1);
    d += pc(20) * pw(6, 2);
    d += pc(21) * pw(6, 0);
    d += pc(22) * pw(7, 2);
    d += pc(23) * pw(7, 0);
    d += pc(24) * pw(8, 2);
    d += pc(25) * pw(8, 0);
    return (pc(26) + (o + u + d)) / pc(27);
  };

  // A points (column parity pa): row neighbours are old B values.  The
  // element pairs are handled with constant indices (the calls below, not a
  // loop), so every selected value is a register, never an address.
  const int eA = (pa - ib) & 1, eB = eA ^ 1;
  T av[kH];
  auto stageA = [&](auto mmc) {
    constexpr int mm = decltype(mmc)::value;
    const T xm = eA ? w[0][2 * mm + 1] : w[0][2 * mm];
    const T xp = eA ? w[0][2 * mm + 3] : w[0][2 * mm + 2];
    av[mm] = point(eA, mm, xm, xp);
    if (act) nrow[ib + eA + 2 * mm] = av[mm];
  };
  stageA(std::integral_constant<int, 0>{});
  if constexpr (kH > 1) stageA(std::integral_constant<int, 1>{});
  __syncthreads();
  // B points: new A neighbours from the row buffer, boundary columns old
  T out[kV];
  auto stageB = [&](auto mmc) {
    constexpr int mm = decltype(mmc)::value;
    const int col = ib + eB + 2 * mm;
    const T om = eB ? w[0][2 * mm + 1] : w[0][2 * mm];
    const T op = eB ? w[0][2 * mm + 3] : w[0][2 * mm + 2];
    const T xm = col - 1 >= 1 ? nrow[col - 1] : om;
    const T xp = col + 1 <= nx - 2 ? nrow[col + 1] : op;
    const T bv = point(eB, mm, xm, xp);
    out[2 * mm] = eA ? bv : av[mm];
    out[2 * mm + 1] = eA ? av[mm] : bv;
  };
  stageB(std::integral_constant<int, 0>{});
  if constexpr (kH > 1) stageB(std::integral_constant<int, 1>{});
  if (act) vst16(x + p, out);
}

// The four colours of one plane parity of that sweep in one launch: colours
// 4ck .. 4ck+3 (4ck+3 .. 4ck for the adjoint sweep).  Planes of one parity do
// not touch each other, so a workgroup owns a TX x TY tile of one plane.  It
// loads the plane's values over the tile and a 4-point margin into LDS and
// runs the four colours there in order: colour s over the tile grown by 3 - s
// points, so the margin holds every value the tile's later colours read.
// Margin points are computed by several workgroups, identically, and only the
// tile is written.  In-plane values come from src, the neighbouring planes
// from nbr, and the result goes to dst, so no workgroup reads what another
// writes.  The arithmetic is part27's, term for term.
template <class T, int TX, int TY, int NTH>
__device__ __forceinline__ void plane_body(int nx, int ny, int nz, const Ops27<T>& A,
                                           const T* __restrict__ D, const T* __restrict__ fc,
                                           const T* __restrict__ src, const T* __restrict__ nbr,
                                           T* __restrict__ dst, int ck, int adj) {
  constexpr int M = 4, LX = TX + 2 * M, LY = TY + 2 * M;
  __shared__ T sx[LX * LY];
  const T *oE = A.a[1], *oN = A.a[2], *uC = A.a[3], *oNE = A.a[4], *oNW = A.a[5], *uE = A.a[6],
          *uW = A.a[7], *uN = A.a[8], *uS = A.a[9], *uNE = A.a[10], *uNW = A.a[11],
          *uSE = A.a[12], *uSW = A.a[13];
  const int tilesx = (nx - 2 + TX - 1) / TX;
  const int tx0 = 1 + (int)(blockIdx.x % tilesx) * TX;
  const int ty0 = 1 + (int)(blockIdx.x / tilesx) * TY;
  const int k = (ck ? 1 : 2) + 2 * (int)blockIdx.y;
  const size_t nxy = (size_t)nx * ny;
  const size_t kp = (size_t)k * nxy;
  const int bx = tx0 - M, by = ty0 - M;  // grid coordinates of sx[0]
  for (int t = threadIdx.x; t < LX * LY; t += NTH) {
    const int i = bx + t % LX, j = by + t / LX;
    sx[t] = (i >= 0 && i < nx && j >= 0 && j < ny) ? src[kp + (size_t)j * nx + i] : T(0);
  }
  __syncthreads();
  for (int s = 0; s < 4; s++) {
    const int q = adj ? 3 - s : s;  // in-plane colour: i parity (q & 1), j parity (q >> 1)
    const int pi = q & 1, pj = q >> 1;
    const int g = 3 - s;
    int ilo = max(1, tx0 - g), jlo = max(1, ty0 - g);
    const int ihi = min(nx - 2, tx0 + TX - 1 + g), jhi = min(ny - 2, ty0 + TY - 1 + g);
    // colour parity 0 is the even indices (first interior point 2), 1 the odd
    ilo += (ilo & 1) != pi;
    jlo += (jlo & 1) != pj;
    const int ni = ilo <= ihi ? (ihi - ilo) / 2 + 1 : 0;
    const int nj = jlo <= jhi ? (jhi - jlo) / 2 + 1 : 0;
    for (int t = threadIdx.x; t < ni * nj; t += NTH) {
      const int i = ilo + 2 * (t % ni), j = jlo + 2 * (t / ni);
      const int e = (j - by) * LX + (i - bx);
      const size_t p = kp + (size_t)j * nx + i;
      T o = oN[p] * sx[e + LX];
      o += oN[p - nx] * sx[e - LX];
      o += oE[p] * sx[e + 1];
      o += oE[p - 1] * sx[e - 1];
      o += oNE[p] * sx[e + LX + 1];
      o += oNW[p] * sx[e + LX - 1];
      o += oNW[p - nx + 1] * sx[e - LX + 1];
      o += oNE[p - nx - 1] * sx[e - LX - 1];
      const size_t qa = p + nxy;
      T u = uC[p] * nbr[qa];
      u += uN[p] * nbr[qa + nx];
      u += uS[p] * nbr[qa - nx];
      u += uE[p] * nbr[qa + 1];
      u += uW[p] * nbr[qa - 1];
      u += uNE[p] * nbr[qa + nx + 1];
      u += uNW[p] * nbr[qa + nx - 1];
      u += uSE[p] * nbr[qa - nx + 1];
      u += uSW[p] * nbr[qa - nx - 1];
      const size_t m = p - nxy;
      T d = uC[m] * nbr[m];
      d += uS[m + nx] * nbr[m + nx];
      d += uN[m - nx] * nbr[m - nx];
      d += uW[m + 1] * nbr[m + 1];
      d += uE[m - 1] * nbr[m - 1];
      d += uSW[m + nx + 1] * nbr[m + nx + 1];
      d += uSE[m + nx - 1] * nbr[m + nx - 1];
      d += uNW[m - nx + 1] * nbr[m - nx + 1];
      d += uNE[m - nx - 1] * nbr[m - nx - 1];
      sx[e] = (fc[p] + (o + u + d)) / D[p];
    }
    __syncthreads();
  }
  for (int t = threadIdx.x; t < TX * TY; t += NTH) {
    const int i = tx0 + t % TX, j = ty0 + t / TX;
    if (i <= nx - 2 && j <= ny - 2) dst[kp + (size_t)j * nx + i] = sx[(j - by) * LX + (i - bx)];
  }
}

// (x, y, z) of a calculation: the plane's own values, the neighbour planes'
// values and the output.
template <class T, int TX, int TY, int NTH>
__global__ void __launch_bounds__(NTH)
    k_gs27_plane(int nx, int ny, int nz, Batch<Args27<T>> B, int ck, int adj) {
  const Args27<T>& a = B.v[blockIdx.z];
  plane_body<T, TX, TY, NTH>(nx, ny, nz, a.A, a.D, a.f, a.x, a.y, a.z, ck, adj);
}

template <class T>
__device__ __forceinline__ void resid7_body(int nx, int ny, int nz, const T* __restrict__ oE,
                                            const T* __restrict__ oN, const T* __restrict__ uC,
                                            const T* __restrict__ D, const T* __restrict__ fc,
                                            const T* __restrict__ x, T* __restrict__ r) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const int k = blockIdx.y + 1;
  const int i = 1 + t % (nx - 2), j = 1 + t / (nx - 2);
  if (j > ny - 2) return;
  const size_t nxy = (size_t)nx * ny;
  const size_t p = IX(i, j, k, nx, ny);
  r[p] = acc7(fc[p], oE, oN, uC, x, p, nx, nxy) - D[p] * x[p];
}

template <class T>
__global__ void __launch_bounds__(kBlock) k_resid7(int nx, int ny, int nz, Batch<Args27<T>> B) {
  const Args27<T>& a = B.v[blockIdx.z];
  resid7_body<T>(nx, ny, nz, a.A.a[1], a.A.a[2], a.A.a[3], a.D, a.f, a.x, a.y);
}

// The 7-point residual with one lane per 16-byte run of a row (4 floats or
// 2 doubles): every operand is one 16-byte load, x of the row a window of
// kV + 2 values, r one 16-byte store.  The last run of a row is moved inwards
// so that every load stays inside its row (overlapping runs write the same
// values).  acc7's terms and order, so the values are k_resid7's bit for bit.
template <class T>
__global__ void __launch_bounds__(kBlock) k_resid7v(int nx, int ny, int nz, Batch<Args27<T>> B) {
  constexpr int kV = 16 / (int)sizeof(T);
  const Args27<T>& a = B.v[blockIdx.z];
  const T *oE = a.A.a[1], *oN = a.A.a[2], *uC = a.A.a[3], *D = a.D, *fc = a.f, *x = a.x;
  T* r = a.y;
  const int R = (nx - 2 + kV - 1) / kV;
  const int t = (int)(blockIdx.x * kBlock + threadIdx.x);
  const int run = t % R, row = t / R;
  if (row >= (ny - 2) * (nz - 2)) return;
  const int j = 1 + row % (ny - 2), k = 1 + row / (ny - 2);
  const int ib = min(1 + kV * run, nx - 1 - kV);
  const size_t nxy = (size_t)nx * ny;
  const size_t p = (size_t)k * nxy + (size_t)j * nx + ib;
  T vf[kV], vN[kV], vNm[kV], vE[kV], vEm[kV], vUm[kV], vU[kV], vD[kV];
  T xn[kV], xs[kV], xu[kV], xd[kV], xw[kV + 2];
  vld16(vf, fc + p);
  vld16(vN, oN + p);
  vld16(vNm, oN + p - nx);
  vld16(vE, oE + p);
  vld16(vEm, oE + p - 1);
  vld16(vUm, uC + p - nxy);
  vld16(vU, uC + p);
  vld16(vD, D + p);
  vld16(xn, x + p + nx);
  vld16(xs, x + p - nx);
  vld16(xu, x + p + nxy);
  vld16(xd, x + p - nxy);
  vldw<T, kV>(xw, x + p - 1);
  T out[kV];
#pragma unroll
  for (int e = 0; e < kV; e++) {
    T sv = vf[e];
    sv += vN[e] * xn[e];
    sv += vNm[e] * xs[e];
    sv += vE[e] * xw[e + 2];
    sv += vEm[e] * xw[e];
    sv += vUm[e] * xd[e];
    sv += vU[e] * xu[e];
    out[e] = sv - vD[e] * xw[e + 1];
  }
  vst16(r + p, out);
}

template <class T>
__device__ __forceinline__ void resid27_body(int nx, int ny, int nz, const Ops27<T>& A,
                                             const T* __restrict__ D, const T* __restrict__ fc,
                                             const T* __restrict__ x, T* __restrict__ r) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const int k = blockIdx.y + 1;
  const int i = 1 + t % (nx - 2), j = 1 + t / (nx - 2);
  if (j > ny - 2) return;
  const size_t nxy = (size_t)nx * ny;
  const size_t p = IX(i, j, k, nx, ny);
  P27<T> s = part27<T>(A.a, x, p, nx, nxy);
  r[p] = fc[p] + s.o + s.u + s.d - D[p] * x[p];
}

template <class T>
__global__ void __launch_bounds__(kBlock) k_resid27(int nx, int ny, int nz, Batch<Args27<T>> B) {
  const Args27<T>& a = B.v[blockIdx.z];
  resid27_body<T>(nx, ny, nz, a.A, a.D, a.f, a.x, a.y);
}

// --------------------------------------------------------------- reductions
__device__ inline double block_sum(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

// The interior points [begin, end) (x fastest) a reduction block owns: kParts
// contiguous chunks, dealt so that XCD x (blocks x, x + 8, ...) owns the
// chunks [x * kParts/8, (x + 1) * kParts/8), a contiguous slab of the grid.
__device__ inline void part_range(int n, int& begin, int& end) {
  const int b = blockIdx.x;
  const int c = (b % 8) * (kParts / 8) + b / 8;
  const int chunk = (n + kParts - 1) / kParts;
  begin = c * chunk;
  end = begin + chunk < n ? begin + chunk : n;
}

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_nrm1_part(int nx, int ny, int nz, const T* __restrict__ x, double* part) {
  __shared__ double sh[kBlock];
  const int ni = nx - 2, nij = ni * (ny - 2), n = nij * (nz - 2);
  int t0, t1;
  part_range(n, t0, t1);
  double acc = 0.0;
#pragma unroll 4
  for (int t = t0 + threadIdx.x; t < t1; t += kBlock) {
    const int k = 1 + t / nij, rem = t - (k - 1) * nij;
    const int j = 1 + rem / ni, i = 1 + rem - (j - 1) * ni;
    acc += fabs((double)x[IX(i, j, k, nx, ny)]);
  }
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// The 7-point residual and the partial sums of its L1 norm in one pass: the
// points are visited in k_nrm1_part's order (part_range), so the norm is
// bitwise that of storing r and calling k_nrm1_part.  r is stored when
// non-null.
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_resid7_nrm(int nx, int ny, int nz, const T* __restrict__ oE, const T* __restrict__ oN,
                 const T* __restrict__ uC, const T* __restrict__ D, const T* __restrict__ fc,
                 const T* __restrict__ x, T* __restrict__ r, float* __restrict__ r32,
                 double* part) {
  __shared__ double sh[kBlock];
  const int ni = nx - 2, nij = ni * (ny - 2), n = nij * (nz - 2);
  const size_t nxy = (size_t)nx * ny;
  int t0, t1;
  part_range(n, t0, t1);
  double acc = 0.0;
  for (int t = t0 + threadIdx.x; t < t1; t += kBlock) {
    const int k = 1 + t / nij, rem = t - (k - 1) * nij;
    const int j = 1 + rem / ni, i = 1 + rem - (j - 1) * ni;
    const size_t p = IX(i, j, k, nx, ny);
    // every operand first; acc7's terms and order
    const T x0 = x[p], xjp = x[p + nx], xjm = x[p - nx], xip = x[p + 1], xim = x[p - 1];
    const T xkm = x[p - nxy], xkp = x[p + nxy];
    const T vN = oN[p], vNm = oN[p - nx], vE = oE[p], vEm = oE[p - 1];
    const T vUm = uC[p - nxy], vU = uC[p], vD = D[p];
    T sv = fc[p];
    sv += vN * xjp;
    sv += vNm * xjm;
    sv += vE * xip;
    sv += vEm * xim;
    sv += vUm * xkm;
    sv += vU * xkp;
    const T v = sv - vD * x0;
    if (r) r[p] = v;
    if (r32) r32[p] = (float)v;
    acc += fabs((double)v);
  }
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// The L1 norm partials of x (k_nrm1_part's order) and y = (float) x on the
// interior, in one pass.
__global__ void __launch_bounds__(kBlock)
    k_nrm1_to32(int nx, int ny, int nz, const double* __restrict__ x, float* __restrict__ y,
                double* part) {
  __shared__ double sh[kBlock];
  const int ni = nx - 2, nij = ni * (ny - 2), n = nij * (nz - 2);
  int t0, t1;
  part_range(n, t0, t1);
  double acc = 0.0;
  for (int t = t0 + threadIdx.x; t < t1; t += kBlock) {
    const int k = 1 + t / nij, rem = t - (k - 1) * nij;
    const int j = 1 + rem / ni, i = 1 + rem - (j - 1) * ni;
    const size_t p = IX(i, j, k, nx, ny);
    const double v = x[p];
    y[p] = (float)v;
    acc += fabs(v);
  }
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// One calculation of a reduction's finishing launch.
struct FinishArgs {
  const double* part;
  double* out;
  double* hout;
};

// out[0] = sum(part[0..kParts-1]); ratio: out[0] = sum(a)/sum(b)
// (and hout[0] = out[0] when hout is given); calculation blockIdx.x.  Only
// the first nparts partial sums are read: the others are taken as +0.0, which
// leaves the sum's bits as they are (no partial sum is -0.0).
__global__ void __launch_bounds__(kBlock) k_finish(Batch<FinishArgs> B, int ratio, int nparts) {
  __shared__ double sh[kBlock];
  const double* part = B.v[blockIdx.x].part;
  double* out = B.v[blockIdx.x].out;
  double* hout = B.v[blockIdx.x].hout;
  double a = 0.0, b = 0.0;
  for (int t = threadIdx.x; t < nparts; t += kBlock) {
    a += part[t];
    if (ratio) b += part[kParts + t];
  }
  double sa = block_sum(a, sh);
  double sb = ratio ? block_sum(b, sh) : 0.0;
  if (threadIdx.x == 0) {
    if (ratio) {
      out[0] = sa / sb;
      out[1] = sa;
      out[2] = sb;
    } else {
      out[0] = sa;
    }
    // a system-scope store: the value reaches host memory past the GPU caches
    if (hout) __hip_atomic_store(hout, ratio ? sa / sb : sa, __ATOMIC_RELEASE,
                                 __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// One calculation of a damping launch: the level's operator, the coarse
// correction x and right-hand side w0, and its partial sums.
template <class T>
struct DampPart {
  Ops27<T> A;
  const T* D;
  const T* x;
  const T* w0;
  double* part;
};

template <class T>
__device__ __forceinline__ void damp_body(int nx, int ny, int nz, const Ops27<T>& A,
                                          const T* __restrict__ D, const T* __restrict__ x,
                                          const T* __restrict__ w0, double* part) {
  __shared__ double sh[kBlock];
  const long ni = nx - 2, nij = ni * (ny - 2), n = nij * (nz - 2);
  const size_t nxy = (size_t)nx * ny;
  double num = 0.0, den = 0.0;
  for (long t = (long)blockIdx.x * kBlock + threadIdx.x; t < n; t += (long)kParts * kBlock) {
    const int i = 1 + (int)(t % ni), j = 1 + (int)((t / ni) % (ny - 2)), k = 1 + (int)(t / nij);
    const size_t p = IX(i, j, k, nx, ny);
    P27<T> s = part27<T>(A.a, x, p, nx, nxy);
    const T ax = (-s.o) + (-s.u) + (-s.d) + D[p] * x[p];
    num += (double)x[p] * (double)w0[p];
    den += (double)x[p] * (double)ax;
  }
  double sn = block_sum(num, sh);
  double sd = block_sum(den, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = sn;
    part[kParts + blockIdx.x] = sd;
  }
}

template <class T>
__global__ void __launch_bounds__(kBlock) k_damp_part(int nx, int ny, int nz, Batch<DampPart<T>> B) {
  const DampPart<T>& a = B.v[blockIdx.z];
  damp_body<T>(nx, ny, nz, a.A, a.D, a.x, a.w0, a.part);
}

// ---------------------------------------------------------------- transfers
// Restriction P^T (matvecd.c:801-913) with weights 1, 1/2, 1/4, 1/8.
template <class T>
__device__ __forceinline__ void restrict_body(int nxf, int nyf, int nzf, int nxc, int nyc, int nzc,
                                              const T* __restrict__ in, T* __restrict__ out,
                                              T* __restrict__ zero) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const int K = blockIdx.y + 1;
  const int I = 1 + t % (nxc - 2), J = 1 + t / (nxc - 2);
  if (J > nyc - 2) return;
  const T W1 = T(1.0), W2 = T(0.5), W4 = T(0.25), W8 = T(0.125);
  const int ii = 2 * I, jj = 2 * J, kk = 2 * K;
  const size_t sx = 1, sy = nxf, sz = (size_t)nxf * nyf;
  const size_t c = IX(ii, jj, kk, nxf, nyf);
  T o = W1 * in[c];
  o += W2 * in[c + sy];
  o += W2 * in[c - sy];
  o += W2 * in[c + sx];
  o += W2 * in[c - sx];
  o += W4 * in[c + sx + sy];
  o += W4 * in[c - sx + sy];
  o += W4 * in[c + sx - sy];
  o += W4 * in[c - sx - sy];
  const size_t u0 = c + sz;
  T u = W2 * in[u0];
  u += W4 * in[u0 + sy];
  u += W4 * in[u0 - sy];
  u += W4 * in[u0 + sx];
  u += W4 * in[u0 - sx];
  u += W8 * in[u0 + sx + sy];
  u += W8 * in[u0 - sx + sy];
  u += W8 * in[u0 + sx - sy];
  u += W8 * in[u0 - sx - sy];
  const size_t d0 = c - sz;
  T d = W2 * in[d0];
  d += W4 * in[d0 + sy];
  d += W4 * in[d0 - sy];
  d += W4 * in[d0 + sx];
  d += W4 * in[d0 - sx];
  d += W8 * in[d0 + sx + sy];
  d += W8 * in[d0 - sx + sy];
  d += W8 * in[d0 + sx - sy];
  d += W8 * in[d0 - sx - sy];
  out[IX(I, J, K, nxc, nyc)] = o + u + d;
  if (zero) zero[IX(I, J, K, nxc, nyc)] = T(0);
}

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_restrict(int nxf, int nyf, int nzf, int nxc, int nyc, int nzc, Batch<TransferArgs<T>> B) {
  const TransferArgs<T>& a = B.v[blockIdx.z];
  restrict_body<T>(nxf, nyf, nzf, nxc, nyc, nzc, a.in, a.out, a.zero);
}

// Trilinear prolongation and correction x_f += damp * P x_c (matvecd.c:935-1076
// followed by Vxaxpy, mgcsd.c:485-487).  The reference gives coarse boundary
// points weight 0; the coarse solution's boundary layer is zero (never
// written), so here they are read and add +0: the same terms in the same
// order, W * x_c with W = 1/2^(number of odd fine indices), without the
// per-term boundary tests.
template <class T>
__device__ __forceinline__ void prolong_body(int nxc, int nyc, int nzc, int nxf, int nyf, int nzf,
                                             const T* __restrict__ xc, T* __restrict__ xf,
                                             const double* __restrict__ damp) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const int k = blockIdx.y + 1;
  const int i = 1 + t % (nxf - 2), j = 1 + t / (nxf - 2);
  if (j > nyf - 2) return;
  const int bx = i & 1, by = j & 1, bz = k & 1;
  const T W = (bx ? T(0.5) : T(1.0)) * (by ? T(0.5) : T(1.0)) * (bz ? T(0.5) : T(1.0));
  const size_t sy = nxc, sz = (size_t)nxc * nyc;
  const size_t c0 = IX(i >> 1, j >> 1, k >> 1, nxc, nyc);
  // terms in the order (0,0,0), (1,0,0), (0,1,0), (1,1,0), then the same at kk + 1
  T v = W * xc[c0];
  if (bx) v += W * xc[c0 + 1];
  if (by) {
    v += W * xc[c0 + sy];
    if (bx) v += W * xc[c0 + sy + 1];
  }
  if (bz) {
    v += W * xc[c0 + sz];
    if (bx) v += W * xc[c0 + sz + 1];
    if (by) {
      v += W * xc[c0 + sz + sy];
      if (bx) v += W * xc[c0 + sz + sy + 1];
    }
  }
  const size_t p = IX(i, j, k, nxf, nyf);
  xf[p] = xf[p] + T(damp[0]) * v;
}

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_prolong_add(int nxc, int nyc, int nzc, int nxf, int nyf, int nzf, Batch<TransferArgs<T>> B) {
  const TransferArgs<T>& a = B.v[blockIdx.z];
  prolong_body<T>(nxc, nyc, nzc, nxf, nyf, nzf, a.in, a.out, a.damp);
}

// ------------------------------------------------------------- vector kit
template <class T>
__global__ void k_fill(T* x, size_t n, T v) {
  for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < n; t += (size_t)gridDim.x * kBlock)
    x[t] = v;
}

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_axpy_int(int nx, int ny, int nz, double a, const T* __restrict__ x, T* __restrict__ y) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const int k = blockIdx.y + 1;
  const int i = 1 + t % (nx - 2), j = 1 + t / (nx - 2);
  if (j > ny - 2) return;
  const size_t p = IX(i, j, k, nx, ny);
  y[p] += T(a) * x[p];
}

__global__ void __launch_bounds__(kBlock)
    k_xpay_int(int nx, int ny, int nz, const double* __restrict__ x, double a,
               const double* __restrict__ z, double* __restrict__ y) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const int k = blockIdx.y + 1;
  const int i = 1 + t % (nx - 2), j = 1 + t / (nx - 2);
  if (j > ny - 2) return;
  const size_t p = IX(i, j, k, nx, ny);
  double v = x[p];
  v += a * z[p];
  y[p] = v;
}

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_copy_int(int nx, int ny, int nz, const T* __restrict__ x, T* __restrict__ y) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const int k = blockIdx.y + 1;
  const int i = 1 + t % (nx - 2), j = 1 + t / (nx - 2);
  if (j > ny - 2) return;
  const size_t p = IX(i, j, k, nx, ny);
  y[p] = x[p];
}

template <class T>
__global__ void k_add_diag(size_t n, const T* __restrict__ oC, const T* __restrict__ cc,
                           T* __restrict__ D, float* __restrict__ Df) {
  for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < n; t += (size_t)gridDim.x * kBlock) {
    const T v = oC[t] + cc[t];
    D[t] = v;
    if (Df) Df[t] = (float)v;
  }
}

// Zero up to kZeroMax arrays in one launch.  Each array owns a contiguous
// range of blocks (found once per block).  A whole array is cleared in 16-byte
// stores, kZeroChunk words per block; a shell array one outer-layer element
// per thread: the planes k = 0 and nz-1, then per inner plane the rows j = 0
// and ny-1 and the ends i = 0 and nx-1 of the other rows.
constexpr unsigned kZeroChunk = 4 * 4 * kBlock;

__device__ inline size_t shell_index(int nx, int ny, int nz, size_t e) {
  const size_t plane = (size_t)nx * ny;
  if (e < plane) return e;
  if (e < 2 * plane) return (size_t)(nz - 1) * plane + (e - plane);
  const size_t r0 = e - 2 * plane;
  const size_t per = 2 * (size_t)nx + 2 * (size_t)(ny - 2);
  const size_t k = 1 + r0 / per;
  const size_t r = r0 % per;
  size_t i, j;
  if (r < (size_t)nx) {
    i = r;
    j = 0;
  } else if (r < 2 * (size_t)nx) {
    i = r - nx;
    j = ny - 1;
  } else {
    const size_t q = r - 2 * (size_t)nx;
    j = 1 + q / 2;
    i = (q & 1) ? nx - 1 : 0;
  }
  return (k * ny + j) * nx + i;
}

__global__ void __launch_bounds__(kBlock) k_zero_many(ZeroArgs Z) {
  const unsigned b = blockIdx.x;
  int a = 0;
  while (a + 1 < Z.n && b >= Z.blk[a + 1]) a++;
  const unsigned lb = b - Z.blk[a];
  unsigned* base = Z.p[a];
  if (Z.nx[a] == 0) {
    const size_t w0 = (size_t)lb * kZeroChunk;
    const size_t w1 = min(w0 + kZeroChunk, Z.words[a]);
    const size_t v1 = w0 + ((w1 - w0) & ~(size_t)3);
    uint4* v = reinterpret_cast<uint4*>(base + w0);
    for (size_t t = threadIdx.x; w0 + 4 * t < v1; t += kBlock) v[t] = make_uint4(0u, 0u, 0u, 0u);
    for (size_t w = v1 + threadIdx.x; w < w1; w += kBlock) base[w] = 0u;
  } else {
    const size_t e = (size_t)lb * kBlock + threadIdx.x;
    if (e >= Z.words[a]) return;
    const size_t p = shell_index(Z.nx[a], Z.ny[a], Z.nz[a], e) * Z.wpe[a];
    for (int w = 0; w < Z.wpe[a]; w++) base[p + w] = 0u;
  }
}

// D_l = oC_l + c_l on every level in one launch (and Df_l = (float) D_l
// where Df_l is given): the same values as one add_diag per level.
__global__ void __launch_bounds__(kBlock) k_add_diag_levels(DiagLevels L) {
  const size_t tot = L.off[L.nl];
  for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < tot;
       t += (size_t)gridDim.x * kBlock) {
    int l = 0;
    while (t >= L.off[l + 1]) l++;
    const size_t i = t - L.off[l];
    const double v = L.oC[l][i] + L.c[l][i];
    L.D[l][i] = v;
    if (L.Df[l]) L.Df[l][i] = (float)v;
  }
}

// y += a * (double) x on the interior: an FP32 correction added to an FP64
// solution, the same bits as converting x first and calling axpy_interior
__global__ void __launch_bounds__(kBlock)
    k_axpy_int_f32(int nx, int ny, int nz, double a, const float* __restrict__ x,
                   double* __restrict__ y, int assign) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const int k = blockIdx.y + 1;
  const int i = 1 + t % (nx - 2), j = 1 + t / (nx - 2);
  if (j > ny - 2) return;
  const size_t p = IX(i, j, k, nx, ny);
  // assign: y starts at zero, and 0.0 + v rounds like the update of a zero y
  double v = assign ? 0.0 : y[p];
  v += a * (double)x[p];
  y[p] = v;
}

__global__ void k_scale(size_t n, double a, const double* __restrict__ x, double* __restrict__ y) {
  for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < n; t += (size_t)gridDim.x * kBlock)
    y[t] = a * x[t];
}

template <class Tin, class Tout>
__global__ void k_convert(size_t n, const Tin* __restrict__ in, Tout* __restrict__ out) {
  for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < n; t += (size_t)gridDim.x * kBlock)
    out[t] = (Tout)in[t];
}

// array blockIdx.y of the list
__global__ void __launch_bounds__(kBlock) k_convert_many(ConvArgs C) {
  const double* __restrict__ in = C.in[blockIdx.y];
  float* __restrict__ out = C.out[blockIdx.y];
  for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < C.n;
       t += (size_t)gridDim.x * kBlock)
    out[t] = (float)in[t];
}

// VfboundPMG (mikpckd.c:209-256): x faces, then y faces, then z faces, so
// edge points take the value of the last face written.
__global__ void k_faces(int nx, int ny, int nz, double* u, const double* gx, const double* gy,
                        const double* gz, int pass) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (pass == 0) {
    if (t >= ny * nz) return;
    const int j = t % ny, k = t / ny;
    u[IX(0, j, k, nx, ny)] = gx[j + ny * k];
    u[IX(nx - 1, j, k, nx, ny)] = gx[j + ny * k + (size_t)ny * nz];
  } else if (pass == 1) {
    if (t >= nx * nz) return;
    const int i = t % nx, k = t / nx;
    u[IX(i, 0, k, nx, ny)] = gy[i + nx * k];
    u[IX(i, ny - 1, k, nx, ny)] = gy[i + nx * k + (size_t)nx * nz];
  } else {
    if (t >= nx * ny) return;
    const int i = t % nx, j = t / nx;
    u[IX(i, j, 0, nx, ny)] = gz[i + nx * j];
    u[IX(i, j, nz - 1, nx, ny)] = gz[i + nx * j + (size_t)nx * ny];
  }
}

// ---------------------------------------------------------------- nonlinear
// sum over species of (-c_s q_s) cc exp(clamp(-q_s u)) (mypdec.c:145-223)
__device__ inline double ncvec(double coef, double u, const Ions& I) {
  double out = 0;
  for (int s = 0; s < I.nion; s++) {
    const double zcf2 = -1.0 * I.sconc[s] * I.q[s];
    const double zu2 = -1.0 * I.q[s];
    double az = fabs(zcf2 * coef);
    const double am_zero = (1.0e-20 <= az ? 1.0e-20 : az) * 1.0e20;
    double a = zu2 * u;
    double mn = (a <= 0.0) ? a : 0.0;
    const double am_neg = (mn >= -85.0) ? mn : -85.0;
    double mx = (a >= 0.0) ? a : 0.0;
    const double am_pos = (mx <= 85.0) ? mx : 85.0;
    const double arg = am_zero * (am_neg + am_pos);
    out = out + zcf2 * coef * exp(arg);
  }
  return out;
}

__global__ void __launch_bounds__(kBlock)
    k_nresid7(int nx, int ny, int nz, const double* __restrict__ oC, const double* __restrict__ oE,
              const double* __restrict__ oN, const double* __restrict__ uC,
              const double* __restrict__ cc, const double* __restrict__ fc,
              const double* __restrict__ x, double* __restrict__ r, Ions ions) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const int k = blockIdx.y + 1;
  const int i = 1 + t % (nx - 2), j = 1 + t / (nx - 2);
  if (j > ny - 2) return;
  const size_t nxy = (size_t)nx * ny;
  const size_t p = IX(i, j, k, nx, ny);
  const double w = ncvec(cc[p], x[p], ions);
  r[p] = acc7(fc[p], oE, oN, uC, x, p, nx, nxy) - oC[p] * x[p] - w;
}

// k_nresid7 with the L1-norm partials of the residual (order of k_nrm1_part).
__global__ void __launch_bounds__(kBlock)
    k_nresid7_nrm(int nx, int ny, int nz, const double* __restrict__ oC,
                  const double* __restrict__ oE, const double* __restrict__ oN,
                  const double* __restrict__ uC, const double* __restrict__ cc,
                  const double* __restrict__ fc, const double* __restrict__ x,
                  double* __restrict__ r, Ions ions, double* part) {
  __shared__ double sh[kBlock];
  const int ni = nx - 2, nij = ni * (ny - 2), n = nij * (nz - 2);
  const size_t nxy = (size_t)nx * ny;
  int t0, t1;
  part_range(n, t0, t1);
  double acc = 0.0;
  for (int t = t0 + threadIdx.x; t < t1; t += kBlock) {
    const int k = 1 + t / nij, rem = t - (k - 1) * nij;
    const int j = 1 + rem / ni, i = 1 + rem - (j - 1) * ni;
    const size_t p = IX(i, j, k, nx, ny);
    // every operand first; acc7's terms and order
    const double x0 = x[p], xjp = x[p + nx], xjm = x[p - nx], xip = x[p + 1], xim = x[p - 1];
    const double xkm = x[p - nxy], xkp = x[p + nxy];
    const double vN = oN[p], vNm = oN[p - nx], vE = oE[p], vEm = oE[p - 1];
    const double vUm = uC[p - nxy], vU = uC[p], vC = oC[p], vcc = cc[p], vfc = fc[p];
    const double w = ncvec(vcc, x0, ions);
    double s7 = vfc;
    s7 += vN * xjp;
    s7 += vNm * xjm;
    s7 += vE * xip;
    s7 += vEm * xim;
    s7 += vUm * xkm;
    s7 += vU * xkp;
    const double v = s7 - vC * x0 - w;
    r[p] = v;
    acc += fabs(v);
  }
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// k_nresid7_nrm at the trial point y = x + a z (k_xpay_int's arithmetic,
// evaluated for every stencil point it reads), storing y and the residual:
// the same bits as k_xpay_int then k_nresid7_nrm.  The boundary layers of x
// and z are zero, as y's is.
__global__ void __launch_bounds__(kBlock)
    k_nresid7_xpay(int nx, int ny, int nz, const double* __restrict__ oC,
                   const double* __restrict__ oE, const double* __restrict__ oN,
                   const double* __restrict__ uC, const double* __restrict__ cc,
                   const double* __restrict__ fc, const double* __restrict__ x, double a,
                   const double* __restrict__ z, double* __restrict__ y,
                   double* __restrict__ r, Ions ions, double* part) {
  __shared__ double sh[kBlock];
  const int ni = nx - 2, nij = ni * (ny - 2), n = nij * (nz - 2);
  const size_t nxy = (size_t)nx * ny;
  int t0, t1;
  part_range(n, t0, t1);
  double acc = 0.0;
  for (int t = t0 + threadIdx.x; t < t1; t += kBlock) {
    const int k = 1 + t / nij, rem = t - (k - 1) * nij;
    const int j = 1 + rem / ni, i = 1 + rem - (j - 1) * ni;
    const size_t p = IX(i, j, k, nx, ny);
    // every operand first: the point and its neighbours p, +j, -j, +i, -i, -k, +k
    const size_t q[7] = {p, p + nx, p - nx, p + 1, p - 1, p - nxy, p + nxy};
    double xc[7], zc[7];
#pragma unroll
    for (int u = 0; u < 7; u++) {
      xc[u] = x[q[u]];
      zc[u] = z[q[u]];
    }
    const double vN = oN[p], vNm = oN[p - nx], vE = oE[p], vEm = oE[p - 1];
    const double vUm = uC[p - nxy], vU = uC[p], vC = oC[p], vcc = cc[p], vfc = fc[p];
    double yc[7];
#pragma unroll
    for (int u = 0; u < 7; u++) {
      double v = xc[u];
      v += a * zc[u];
      yc[u] = v;
    }
    const double w = ncvec(vcc, yc[0], ions);
    double s7 = vfc;
    s7 += vN * yc[1];
    s7 += vNm * yc[2];
    s7 += vE * yc[3];
    s7 += vEm * yc[4];
    s7 += vUm * yc[5];
    s7 += vU * yc[6];
    const double v = s7 - vC * yc[0] - w;
    acc += fabs(v);
    y[p] = yc[0];
    r[p] = v;
  }
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// The end of a Newton step's inner cycle and the step's first trial point in
// one pass.  With z the correction the cycle produced (its boundary layer
// zero; read as (double) zf when zf is given, the values an FP32 correction
// assigned to z would have): the inner residual fz + J z of the Jacobian
// system (J's diagonal D), its FP32 image in r32 and its L1 partials in
// part[0, kParts) (k_resid7_nrm<double>'s values and order); and y = x + a z
// with the nonlinear residual at y in rn (its FP32 image in rn32 when given),
// its L1 partials in part[kParts, 2 kParts) (k_nresid7_xpay's values and
// order).  The two stencils share their operator loads.
// Every operand of a point is loaded before the arithmetic (the two stencils
// read z and x at the same seven points); kZF selects z from zf.
template <bool kZF>
__global__ void __launch_bounds__(kBlock)
    k_jres_trial(int nx, int ny, int nz, const double* __restrict__ oC,
                 const double* __restrict__ oE, const double* __restrict__ oN,
                 const double* __restrict__ uC, const double* __restrict__ D,
                 const double* __restrict__ cc, const double* __restrict__ fc,
                 const double* __restrict__ fz, const double* __restrict__ x, double a,
                 const double* __restrict__ z, const float* __restrict__ zf,
                 double* __restrict__ y, double* __restrict__ rn, float* __restrict__ r32,
                 float* __restrict__ rn32, Ions ions, double* part) {
  __shared__ double sh[kBlock];
  const int ni = nx - 2, nij = ni * (ny - 2), n = nij * (nz - 2);
  const size_t nxy = (size_t)nx * ny;
  int t0, t1;
  part_range(n, t0, t1);
  double acc_j = 0.0, acc_n = 0.0;
  for (int t = t0 + threadIdx.x; t < t1; t += kBlock) {
    const int k = 1 + t / nij, rem = t - (k - 1) * nij;
    const int j = 1 + rem / ni, i = 1 + rem - (j - 1) * ni;
    const size_t p = IX(i, j, k, nx, ny);
    // the point and its neighbours in acc7's order: p, +j, -j, +i, -i, -k, +k
    const size_t q[7] = {p, p + nx, p - nx, p + 1, p - 1, p - nxy, p + nxy};
    double zc[7], xc[7];
#pragma unroll
    for (int u = 0; u < 7; u++) {
      zc[u] = kZF ? (double)zf[q[u]] : z[q[u]];
      xc[u] = x[q[u]];
    }
    const double vfz = fz[p], vN = oN[p], vNm = oN[p - nx], vE = oE[p], vEm = oE[p - 1];
    const double vUm = uC[p - nxy], vU = uC[p], vD = D[p], vC = oC[p], vcc = cc[p], vfc = fc[p];
    // acc7's terms and order at z
    double sj = vfz;
    sj += vN * zc[1];
    sj += vNm * zc[2];
    sj += vE * zc[3];
    sj += vEm * zc[4];
    sj += vUm * zc[5];
    sj += vU * zc[6];
    const double vj = sj - vD * zc[0];
    acc_j += fabs(vj);
    // y = x + a z at the seven points (k_xpay_int's arithmetic)
    double yc[7];
#pragma unroll
    for (int u = 0; u < 7; u++) {
      double v = xc[u];
      v += a * zc[u];
      yc[u] = v;
    }
    const double w = ncvec(vcc, yc[0], ions);
    double s7 = vfc;
    s7 += vN * yc[1];
    s7 += vNm * yc[2];
    s7 += vE * yc[3];
    s7 += vEm * yc[4];
    s7 += vUm * yc[5];
    s7 += vU * yc[6];
    const double v = s7 - vC * yc[0] - w;
    acc_n += fabs(v);
    if (r32) r32[p] = (float)vj;
    y[p] = yc[0];
    rn[p] = v;
    if (rn32) rn32[p] = (float)v;
  }
  const double bj = block_sum(acc_j, sh);
  const double bn = block_sum(acc_n, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = bj;
    part[kParts + blockIdx.x] = bn;
  }
}

// Vdc_vecpmg (mypdec.c:366-433) at every point; with oC given also the
// diagonal D = oC + N'(x) and, with Df, its FP32 image (k_add_diag_levels'
// values for this level) in the same pass.
__global__ void k_dcvec(size_t n, const double* __restrict__ cc, const double* __restrict__ x,
                        double* __restrict__ out, Ions I, const double* __restrict__ oC,
                        double* __restrict__ D, float* __restrict__ Df) {
  for (size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x; t < n; t += (size_t)gridDim.x * kBlock) {
    double o = 0.0;
    const double coef = cc[t], u = x[t];
    for (int s = 0; s < I.nion; s++) {
      const double zcf2 = I.sconc[s] * I.q[s] * I.q[s];
      const double zu2 = -1.0 * I.q[s];
      double az = fabs(zcf2 * coef);
      const double am_zero = (1.0e-20 <= az ? 1.0e-20 : az) * 1.0e20;
      double a = zu2 * u;
      double mn = (a <= 0.0) ? a : 0.0;
      const double am_neg = (mn >= -85.0) ? mn : -85.0;
      double mx = (a >= 0.0) ? a : 0.0;
      const double am_pos = (mx <= 85.0) ? mx : 85.0;
      const double arg = am_zero * (am_neg + am_pos);
      o += zcf2 * coef * exp(arg);
    }
    out[t] = o;
    if (oC) {
      const double v = oC[t] + o;
      D[t] = v;
      if (Df) Df[t] = (float)v;
    }
  }
}

// ---------------------------------------------------------------- operators
// VbuildA_fv (buildAd.c:105-281), face terms folded into fc in the order
// east, north, up, west, south, down.
__global__ void __launch_bounds__(kBlock)
    k_buildA(int nx, int ny, int nz, const double* xf, const double* yf, const double* zf,
             const double* gx, const double* gy, const double* gz, const double* a1,
             const double* a2, const double* a3, const double* ccf, const double* fcf,
             double* oC, double* cc, double* fc, double* oE, double* oN, double* uC) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const int k = blockIdx.y + 1;
  const int i = 1 + t % (nx - 2), j = 1 + t / (nx - 2);
  if (j > ny - 2) return;
  const double diag = 1.0;
  const double hzm1 = zf[k] - zf[k - 1], hz = zf[k + 1] - zf[k];
  const double hym1 = yf[j] - yf[j - 1], hy = yf[j + 1] - yf[j];
  const double hxm1 = xf[i] - xf[i - 1], hx = xf[i + 1] - xf[i];
  const double c_oE = diag * (hym1 + hy) * (hzm1 + hz) / (4.0 * hx);
  const double c_oEm1 = diag * (hym1 + hy) * (hzm1 + hz) / (4.0 * hxm1);
  const double c_oN = diag * (hxm1 + hx) * (hzm1 + hz) / (4.0 * hy);
  const double c_oNm1 = diag * (hxm1 + hx) * (hzm1 + hz) / (4.0 * hym1);
  const double c_uC = diag * (hxm1 + hx) * (hym1 + hy) / (4.0 * hz);
  const double c_uCm1 = diag * (hxm1 + hx) * (hym1 + hy) / (4.0 * hzm1);
  const double c_fc = diag * (hxm1 + hx) * (hym1 + hy) * (hzm1 + hz) / 8.0;
  const size_t nxy = (size_t)nx * ny;
  const size_t p = IX(i, j, k, nx, ny);
  double f = c_fc * fcf[p];
  cc[p] = c_fc * ccf[p];
  double d = c_oE * a1[p];
  d += c_oEm1 * a1[p - 1];
  d += c_oN * a2[p];
  d += c_oNm1 * a2[p - nx];
  d += c_uC * a3[p];
  d += c_uCm1 * a3[p - nxy];
  oC[p] = d;
  if (i != nx - 2) {
    oE[p] = c_oE * a1[p];
  } else {
    oE[p] = 0.0;
    f += c_oE * a1[p] * gx[j + ny * k + (size_t)ny * nz];
  }
  if (j != ny - 2) {
    oN[p] = c_oN * a2[p];
  } else {
    oN[p] = 0.0;
    f += c_oN * a2[p] * gy[i + nx * k + (size_t)nx * nz];
  }
  if (k != nz - 2) {
    uC[p] = c_uC * a3[p];
  } else {
    uC[p] = 0.0;
    f += c_uC * a3[p] * gz[i + nx * j + (size_t)nx * ny];
  }
  if (i == 1) f += c_oEm1 * a1[p - 1] * gx[j + ny * k];
  if (j == 1) f += c_oNm1 * a2[p - nx] * gy[i + nx * k];
  if (k == 1) f += c_uCm1 * a3[p - nxy] * gz[i + nx * j];
  fc[p] = f;
}

// forward offsets of the 14 stored coefficients
__constant__ int c_fwd[14][3] = {{0, 0, 0},  {1, 0, 0},  {0, 1, 0},  {0, 0, 1}, {1, 1, 0},
                                 {-1, 1, 0}, {1, 0, 1},  {-1, 0, 1}, {0, 1, 1}, {0, -1, 1},
                                 {1, 1, 1},  {-1, 1, 1}, {1, -1, 1}, {-1, -1, 1}};

// stored direction of each offset (sx,sy,sz) at (sz+1)*9+(sy+1)*3+(sx+1): +s forward, -s backward
__constant__ int c_ent[27] = {-10, -8, -11, -6, -3, -7, -12, -9, -13, -4, -2, -5, -1, 0, 1, 5, 2, 4, 13, 9, 12, 7, 3, 6, 11, 8, 10};

__device__ inline double pweight(int dx, int dy, int dz) {
  if (dx < -1 || dx > 1 || dy < -1 || dy > 1 || dz < -1 || dz > 1) return 0.0;
  const int m = abs(dx) + abs(dy) + abs(dz);
  return m == 0 ? 1.0 : (m == 1 ? 0.5 : (m == 2 ? 0.25 : 0.125));
}

// A(f, f+s) of the fine operator without cc (same search order as the
// oracle's restatement: first matching stored coefficient, forward then back).
__device__ inline double fine_entry(int st, const Ops27<double>& A, int nxf, int nyf, int fi,
                                    int fj, int fk, int sx, int sy, int sz) {
  if (sx == 0 && sy == 0 && sz == 0) return A.a[0][IX(fi, fj, fk, nxf, nyf)];
  // each offset is exactly one stored direction, forward (+) or backward (-)
  const int e = c_ent[(sz + 1) * 9 + (sy + 1) * 3 + (sx + 1)];
  const int s = e > 0 ? e : -e;
  if (s >= (st == 7 ? 4 : 14)) return 0.0;
  if (e > 0) return -A.a[s][IX(fi, fj, fk, nxf, nyf)];
  return -A.a[s][IX(fi + sx, fj + sy, fk + sz, nxf, nyf)];
}

struct OutPtrs {
  double* a[14];
};

// Inner sum of the Galerkin product for coarse point (I, J, K), stored
// direction to (I2, J2, K2) and fine neighbour d = 9(dz+1) + 3(dy+1) + (dx+1)
// of the fine point 2(I, J, K) (the reference's loop order).  The loops visit
// only the terms with a non-zero second weight pweight(f + s - g0), i.e.
// |f + s - g0| <= 1 per axis, in the full scan's order, so the sum is the
// full scan's.
__device__ __forceinline__ double galerkin_tt(int st, const Ops27<double>& Af, int nxf, int nyf,
                                              int I, int J, int K, int I2, int J2, int K2,
                                              int d) {
  const int dz = d / 9 - 1, dy = (d / 3) % 3 - 1, dx = d % 3 - 1;
  const int fx = 2 * I + dx, fy = 2 * J + dy, fz = 2 * K + dz;
  const int g0x = 2 * I2, g0y = 2 * J2, g0z = 2 * K2;
  const int sx0 = max(-1, g0x - 1 - fx), sx1 = min(1, g0x + 1 - fx);
  const int sy0 = max(-1, g0y - 1 - fy), sy1 = min(1, g0y + 1 - fy);
  const int sz0 = max(-1, g0z - 1 - fz), sz1 = min(1, g0z + 1 - fz);
  double tt = 0.0;
  for (int sz = sz0; sz <= sz1; sz++)
    for (int sy = sy0; sy <= sy1; sy++)
      for (int sx = sx0; sx <= sx1; sx++) {
        if (st == 7 && abs(sx) + abs(sy) + abs(sz) > 1) continue;
        const double p2 = pweight(fx + sx - g0x, fy + sy - g0y, fz + sz - g0z);
        tt += fine_entry(st, Af, nxf, nyf, fx, fy, fz, sx, sy, sz) * p2;
      }
  return tt;
}

// Large levels: one thread per (coarse interior point, stored coefficient),
// 64 consecutive coarse points per wavefront and one direction s per
// wavefront, so the pruned loop bounds (which depend on s only) never
// diverge within a wavefront.
__global__ void __launch_bounds__(kBlock)
    k_galerkin_point(int st, int nxf, int nyf, int nzf, Ops27<double> Af, int nxc, int nyc,
                     int nzc, OutPtrs out) {
  const long ni = nxc - 2, nj = nyc - 2, nk = nzc - 2;
  const long nq = ni * nj * nk;
  const long t = (long)blockIdx.x * kBlock + threadIdx.x;
  const long w = t / 64;
  const int s = (int)(w % 14);
  const long q = (w / 14) * 64 + (t % 64);
  if (q >= nq) return;
  const int I = 1 + (int)(q % ni), J = 1 + (int)((q / ni) % nj), K = 1 + (int)(q / (ni * nj));
  const int I2 = I + c_fwd[s][0], J2 = J + c_fwd[s][1], K2 = K + c_fwd[s][2];
  const size_t pc = IX(I, J, K, nxc, nyc);
  if (I2 < 1 || I2 > nxc - 2 || J2 < 1 || J2 > nyc - 2 || K2 < 1 || K2 > nzc - 2) {
    out.a[s][pc] = 0.0;
    return;
  }
  double sum = 0.0;
  for (int d = 0; d < 27; d++)
    sum += pweight(d % 3 - 1, (d / 3) % 3 - 1, d / 9 - 1) *
           galerkin_tt(st, Af, nxf, nyf, I, J, K, I2, J2, K2, d);
  out.a[s][pc] = (s == 0) ? sum : -sum;
}

// The same sum with every loop unrolled at compile time for one stencil
// width ST and one stored direction S: the pruned loop bounds, both
// prolongation weights and the stored array of each fine offset become
// constants, so a term is one load and one multiply-add.  Terms whose fine
// entry is a structural zero (the diagonal offsets of a 7-point operator) add
// +0.0 to a partial sum that starts at +0.0 and so cannot be -0.0; they are
// left out without changing a bit.
namespace gal {
constexpr int kFwd[14][3] = {{0, 0, 0},  {1, 0, 0},  {0, 1, 0},  {0, 0, 1}, {1, 1, 0},
                             {-1, 1, 0}, {1, 0, 1},  {-1, 0, 1}, {0, 1, 1}, {0, -1, 1},
                             {1, 1, 1},  {-1, 1, 1}, {1, -1, 1}, {-1, -1, 1}};
__host__ __device__ constexpr int iabs(int v) { return v < 0 ? -v : v; }
__host__ __device__ constexpr int imax(int a, int b) { return a > b ? a : b; }
__host__ __device__ constexpr int imin(int a, int b) { return a < b ? a : b; }
// stored direction of offset (sx, sy, sz): +s forward, -s backward
__host__ __device__ constexpr int ent(int sx, int sy, int sz) {
  for (int s = 0; s < 14; s++) {
    if (kFwd[s][0] == sx && kFwd[s][1] == sy && kFwd[s][2] == sz) return s;
    if (kFwd[s][0] == -sx && kFwd[s][1] == -sy && kFwd[s][2] == -sz) return -s;
  }
  return 99;
}
__host__ __device__ constexpr double pw(int dx, int dy, int dz) {
  return (dx < -1 || dx > 1 || dy < -1 || dy > 1 || dz < -1 || dz > 1)
             ? 0.0
             : (iabs(dx) + iabs(dy) + iabs(dz) == 0
                    ? 1.0
                    : (iabs(dx) + iabs(dy) + iabs(dz) == 1
                           ? 0.5
                           : (iabs(dx) + iabs(dy) + iabs(dz) == 2 ? 0.25 : 0.125)));
}

template <int ST, int S>
__device__ __forceinline__ double coef(const Ops27<double>& A, size_t pf, int nx, size_t nxy) {
  constexpr int FX = kFwd[S][0], FY = kFwd[S][1], FZ = kFwd[S][2];
  double sum = 0.0;
#pragma unroll
  for (int d = 0; d < 27; d++) {
    const int dz = d / 9 - 1, dy = (d / 3) % 3 - 1, dx = d % 3 - 1;
    const int ox = 2 * FX - dx, oy = 2 * FY - dy, oz = 2 * FZ - dz;
    const long f = (long)dx + (long)dy * nx + (long)dz * (long)nxy;
    double tt = 0.0;
#pragma unroll
    for (int sz = -1; sz <= 1; sz++)
#pragma unroll
      for (int sy = -1; sy <= 1; sy++)
#pragma unroll
        for (int sx = -1; sx <= 1; sx++) {
          if (sx < imax(-1, ox - 1) || sx > imin(1, ox + 1) || sy < imax(-1, oy - 1) ||
              sy > imin(1, oy + 1) || sz < imax(-1, oz - 1) || sz > imin(1, oz + 1))
            continue;
          if (ST == 7 && iabs(sx) + iabs(sy) + iabs(sz) > 1) continue;
          const double p2 = pw(sx - ox, sy - oy, sz - oz);
          const int e = ent(sx, sy, sz);
          double v;
          if (e == 0)
            v = A.a[0][pf + f];
          else if (e > 0)
            v = -A.a[e][pf + f];
          else
            v = -A.a[-e][pf + f + (long)sx + (long)sy * nx + (long)sz * (long)nxy];
          tt += v * p2;
        }
    sum += pw(dx, dy, dz) * tt;
  }
  return sum;
}

template <int ST, int... S>
__device__ __forceinline__ double coef_switch(int s, const Ops27<double>& A, size_t pf, int nx,
                                              size_t nxy, std::integer_sequence<int, S...>) {
  double r = 0.0;
  ((s == S ? (r = coef<ST, S>(A, pf, nx, nxy), 0) : 0), ...);
  return r;
}
}  // namespace gal

template <int ST, int WAVES>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WAVES)))
    k_galerkin_unrolled(int nxf, int nyf, int nzf, Ops27<double> Af, int nxc, int nyc, int nzc,
                        OutPtrs out) {
  const long ni = nxc - 2, nj = nyc - 2, nk = nzc - 2;
  const long nq = ni * nj * nk;
  const long t = (long)blockIdx.x * kBlock + threadIdx.x;
  const long w = t / 64;
  const int s = (int)(w % 14);
  const long q = (w / 14) * 64 + (t % 64);
  if (q >= nq) return;
  const int I = 1 + (int)(q % ni), J = 1 + (int)((q / ni) % nj), K = 1 + (int)(q / (ni * nj));
  const int I2 = I + c_fwd[s][0], J2 = J + c_fwd[s][1], K2 = K + c_fwd[s][2];
  const size_t pc = IX(I, J, K, nxc, nyc);
  if (I2 < 1 || I2 > nxc - 2 || J2 < 1 || J2 > nyc - 2 || K2 < 1 || K2 > nzc - 2) {
    out.a[s][pc] = 0.0;
    return;
  }
  const size_t nxy = (size_t)nxf * nyf;
  const size_t pf = IX(2 * I, 2 * J, 2 * K, nxf, nyf);
  const double sum =
      gal::coef_switch<ST>(s, Af, pf, nxf, nxy, std::make_integer_sequence<int, 14>{});
  out.a[s][pc] = (s == 0) ? sum : -sum;
}

// Small levels (too few coarse points to fill the chip one thread each): one
// 32-lane group per (coarse interior point, stored coefficient); lane d < 27
// computes tt_d, lane 0 then adds p1_d * tt_d for d = 0..26 in order, the
// same sum as k_galerkin_point's.
constexpr int kGalGroups = kBlock / 32;

__global__ void __launch_bounds__(kBlock)
    k_galerkin_group(int st, int nxf, int nyf, int nzf, Ops27<double> Af, int nxc, int nyc,
                     int nzc, OutPtrs out) {
  __shared__ double tt_sh[kGalGroups][27];
  const long ni = nxc - 2, nj = nyc - 2, nk = nzc - 2;
  const long nq = ni * nj * nk;
  const int grp = threadIdx.x / 32, d = threadIdx.x % 32;
  const long g = (long)blockIdx.x * kGalGroups + grp;
  const int s = (int)(g % 14);
  const long q = g / 14;
  const bool live = q < nq;
  int I = 0, J = 0, K = 0, I2 = 0, J2 = 0, K2 = 0;
  if (live) {
    I = 1 + (int)(q % ni);
    J = 1 + (int)((q / ni) % nj);
    K = 1 + (int)(q / (ni * nj));
    I2 = I + c_fwd[s][0];
    J2 = J + c_fwd[s][1];
    K2 = K + c_fwd[s][2];
  }
  const bool inside = live && !(I2 < 1 || I2 > nxc - 2 || J2 < 1 || J2 > nyc - 2 || K2 < 1 ||
                                K2 > nzc - 2);
  if (inside && d < 27) tt_sh[grp][d] = galerkin_tt(st, Af, nxf, nyf, I, J, K, I2, J2, K2, d);
  __syncthreads();
  if (!live || d != 0) return;
  const size_t pc = IX(I, J, K, nxc, nyc);
  if (!inside) {
    out.a[s][pc] = 0.0;
    return;
  }
  double sum = 0.0;
  for (int e = 0; e < 27; e++) sum += pweight(e % 3 - 1, (e / 3) % 3 - 1, e / 9 - 1) * tt_sh[grp][e];
  out.a[s][pc] = (s == 0) ? sum : -sum;
}

// ------------------------------------------------------------ dense coarse
__global__ void k_dense_assemble(int nx, int ny, int nz, Batch<InvertArgs> B, int N) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  const Ops<double>& A = B.v[blockIdx.y].A;
  const double* D = B.v[blockIdx.y].D;
  double* M = B.v[blockIdx.y].M;
  // the workgroup's rows are cleared with consecutive stores first
  const int r0 = blockIdx.x * kBlock, rows = min(kBlock, N - r0);
  double* Mb = M + (size_t)r0 * N;
  for (int t = threadIdx.x; t < rows * N; t += kBlock) Mb[t] = 0.0;
  __syncthreads();
  if (r >= N) return;
  const int ni = nx - 2, nj = ny - 2;
  const int i = r % ni + 1, j = (r / ni) % nj + 1, k = r / (ni * nj) + 1;
  M[(size_t)r * N + r] = D[IX(i, j, k, nx, ny)];
  // couplings to all 26 neighbours: forward ones from this point's
  // coefficients, backward ones from the neighbour's
  for (int s = 1; s < 14; s++) {
    for (int dir = -1; dir <= 1; dir += 2) {
      const int i2 = i + dir * c_fwd[s][0], j2 = j + dir * c_fwd[s][1], k2 = k + dir * c_fwd[s][2];
      if (i2 < 1 || i2 > nx - 2 || j2 < 1 || j2 > ny - 2 || k2 < 1 || k2 > nz - 2) continue;
      const int c = (i2 - 1) + ni * ((j2 - 1) + nj * (k2 - 1));
      const double v = dir > 0 ? -A.a[s][IX(i, j, k, nx, ny)] : -A.a[s][IX(i2, j2, k2, nx, ny)];
      M[(size_t)r * N + c] = v;
    }
  }
}

// In-place Gauss-Jordan inverse of the (SPD) coarsest operator.  For large N
// the elimination is spread over the whole GPU, one pivot at a time: save
// column k, scale the pivot row, eliminate the other rows.
__global__ void __launch_bounds__(kBlock) k_gj_col(const double* M, double* col, int N, int k) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < N) col[i] = M[(size_t)i * N + k];
}

// A pivot that is not positive and finite means the operator is not positive
// definite (Vdpbfa's leading-minor test, mlinpckd.c:194-233): its 1-based
// index goes to *flag (host-mapped memory) with a system-scope store.
__device__ inline void flag_pivot(double piv, int k, double* flag) {
  if (flag && !(piv > 0.0 && piv <= DBL_MAX))
    __hip_atomic_store(flag, (double)(k + 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

__global__ void __launch_bounds__(kBlock) k_gj_row(double* M, const double* col, int N, int k,
                                                   double* flag) {
  const int j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= N) return;
  if (j == k) flag_pivot(col[k], k, flag);
  const double v = (j == k) ? 1.0 : M[(size_t)k * N + j];
  M[(size_t)k * N + j] = v / col[k];
}

__global__ void __launch_bounds__(kBlock) k_gj_elim(double* M, const double* col, int N, int k) {
  const size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= (size_t)N * N) return;
  const int i = (int)(e / N), j = (int)(e % N);
  if (i == k) return;
  const double f = col[i];
  const double base = (j == k) ? 0.0 : M[(size_t)i * N + j];
  M[(size_t)i * N + j] = base - f * M[(size_t)k * N + j];
}

// x = Minv f reading Minv by columns (coalesced across rows); Minv is the
// inverse of a symmetric operator.
template <class T>
__global__ void k_dense_solve_cols(int nx, int ny, int nz, Batch<DenseArgs<T>> B, int N) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= N) return;
  const double* Minv = B.v[blockIdx.y].Minv;
  const T* f = B.v[blockIdx.y].f;
  T* x = B.v[blockIdx.y].x;
  const int ni = nx - 2, nj = ny - 2;
  double s = 0.0;
  for (int c = 0; c < N; c++) {
    const int i = c % ni + 1, j = (c / ni) % nj + 1, k = c / (ni * nj) + 1;
    s += Minv[(size_t)c * N + r] * (double)f[IX(i, j, k, nx, ny)];
  }
  const int i = r % ni + 1, j = (r / ni) % nj + 1, k = r / (ni * nj) + 1;
  x[IX(i, j, k, nx, ny)] = (T)s;
}

// The same elimination in one workgroup with the matrix held in LDS
// (N*N + 2N doubles, N <= 89).  Per pivot: stage column k and the scaled
// pivot row (v / piv, as k_gj_row computes it), then update every element
// with k_gj_elim's expression; each thread's (i, j) pairs are decoded once.
constexpr int kInvPer = 8;  // elements per thread: 89 * 89 <= 8 * 1024

__global__ void __launch_bounds__(1024) k_dense_invert_lds(Batch<InvertArgs> B, int N) {
  double* Mg = B.v[blockIdx.x].M;
  double* flag = B.v[blockIdx.x].flag;
  extern __shared__ double sm[];
  double* M = sm;
  double* col = sm + (size_t)N * N;
  double* prow = col + N;
  const int tid = threadIdx.x;
  int ei[kInvPer], ej[kInvPer];
  int ne = 0;
  for (int e = tid; e < N * N && ne < kInvPer; e += 1024, ne++) {
    ei[ne] = e / N;
    ej[ne] = e - ei[ne] * N;
  }
  for (int e = tid; e < N * N; e += 1024) M[e] = Mg[e];
  __syncthreads();
  int bad = -1;
  for (int k = 0; k < N; k++) {
    if (tid == 0 && bad < 0 && !(M[k * N + k] > 0.0 && M[k * N + k] <= DBL_MAX)) bad = k;
    if (tid < N) {
      col[tid] = M[tid * N + k];
      const double v = (tid == k) ? 1.0 : M[k * N + tid];
      prow[tid] = v / M[k * N + k];
    }
    __syncthreads();
    for (int q = 0; q < ne; q++) {
      const int i = ei[q], j = ej[q];
      if (i == k) {
        M[i * N + j] = prow[j];
      } else {
        const double base = (j == k) ? 0.0 : M[i * N + j];
        M[i * N + j] = base - col[i] * prow[j];
      }
    }
    __syncthreads();
  }
  for (int e = tid; e < N * N; e += 1024) Mg[e] = M[e];
  if (tid == 0 && bad >= 0) flag_pivot(-1.0, bad, flag);
}

// The same Gauss-Jordan inverse for N <= 64 with the matrix in registers:
// thread t holds 16 entries of row t / 4.  Each pivot step publishes the
// pivot row and column through double-buffered LDS (one barrier per pivot)
// and scales the pivot row by the pivot's reciprocal.
__global__ void __launch_bounds__(256) k_dense_invert_reg(Batch<InvertArgs> B, int N) {
  double* Mg = B.v[blockIdx.x].M;
  double* flag = B.v[blockIdx.x].flag;
  __shared__ double colb[2][64], rowb[2][64];
  const int t = threadIdx.x, r = t >> 2, cb = t & 3, c0 = cb * 16;
  double m[16];
#pragma unroll
  for (int q = 0; q < 16; q++)
    m[q] = (r < N && c0 + q < N) ? Mg[(size_t)r * N + c0 + q] : 0.0;
  int bad = -1;
  for (int k = 0; k < N; k++) {
    const int b = k & 1;
    if (r < N && (k >> 4) == cb) {
      double v = 0.0;
#pragma unroll
      for (int q = 0; q < 16; q++)
        if (q == (k & 15)) v = m[q];
      colb[b][r] = v;
    }
    if (r == k) {
#pragma unroll
      for (int q = 0; q < 16; q++) rowb[b][c0 + q] = m[q];
    }
    __syncthreads();
    if (t == 0 && bad < 0 && !(rowb[b][k] > 0.0 && rowb[b][k] <= DBL_MAX)) bad = k;
    const double inv = 1.0 / rowb[b][k];
    const double ci = r < N ? colb[b][r] : 0.0;
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int j = c0 + q;
      const double pj = ((j == k) ? 1.0 : rowb[b][j]) * inv;
      m[q] = (r == k) ? pj : ((j == k) ? 0.0 : m[q]) - ci * pj;
    }
  }
#pragma unroll
  for (int q = 0; q < 16; q++)
    if (r < N && c0 + q < N) Mg[(size_t)r * N + c0 + q] = m[q];
  if (t == 0 && bad >= 0) flag_pivot(-1.0, bad, flag);
}

// x_r = sum_c Minv[r][c] f_c over the interior unknowns, N <= kDenseSmall:
// one workgroup gathers f into LDS, then each row sums in column order.
constexpr int kDenseSmall = 96;

template <class T>
__global__ void __launch_bounds__(128) k_dense_solve(int nx, int ny, int nz, Batch<DenseArgs<T>> B,
                                                     int N) {
  __shared__ double fv[kDenseSmall];
  const double* Minv = B.v[blockIdx.x].Minv;
  const T* f = B.v[blockIdx.x].f;
  T* x = B.v[blockIdx.x].x;
  const int ni = nx - 2, nj = ny - 2;
  for (int c = threadIdx.x; c < N; c += blockDim.x) {
    const int i = c % ni + 1, j = (c / ni) % nj + 1, k = c / (ni * nj) + 1;
    fv[c] = (double)f[IX(i, j, k, nx, ny)];
  }
  __syncthreads();
  const int r = threadIdx.x;
  if (r >= N) return;
  const double* row = Minv + (size_t)r * N;
  double s = 0.0;
#pragma unroll 8
  for (int c = 0; c < N; c++) s += row[c] * fv[c];
  const int i = r % ni + 1, j = (r / ni) % nj + 1, k = r / (ni * nj) + 1;
  x[IX(i, j, k, nx, ny)] = (T)s;
}

inline unsigned nblocks(size_t n) {
  size_t b = (n + kBlock - 1) / kBlock;
  if (b > 65535) b = 65535;
  if (b == 0) b = 1;
  return (unsigned)b;
}

template <class T>
Ops27<T> to27(const Ops<T>& A) {
  Ops27<T> o;
  for (int s = 0; s < 14; s++) o.a[s] = A.a[s];
  return o;
}

// A batched launch takes at most kMaxBatch calculations.
inline void check_batch(int nb) {
  if (nb < 1 || nb > kMaxBatch) {
    fprintf(stderr, "pmg: a batched launch takes 1 to %d calculations, not %d\n", kMaxBatch, nb);
    abort();
  }
}

}  // namespace

// =================================================================== wrappers

template <class T, int TY, int KC, int TX, int NT = kFT, bool ZX = false>
void fused_sweep(hipStream_t s, int nx, int ny, int nz, const Batch<SweepArgs<T>>& B, int nb,
                 int first) {
  const int tiles = ((nx - 2 + TX - 1) / TX) * ((ny - 2 + TY - 1) / TY) * ((nz - 2 + KC - 1) / KC);
  const long total = (long)tiles * nb;
  const unsigned g = (unsigned)(((total + kXCD - 1) / kXCD) * kXCD);
  k_gsrb7_fused<T, TY, KC, TX, NT, ZX><<<g, NT, 0, s>>>(nx, ny, nz, B, nb, first);
}

// The vector form: one lane per 16-byte run of a region row, the workgroup
// rounded up to whole wavefronts.
template <class T, int TY, int KC, int TX, bool ZX = false>
void vfused_sweep(hipStream_t s, int nx, int ny, int nz, const Batch<SweepArgs<T>>& B, int nb,
                  int first) {
  constexpr int kV = 16 / (int)sizeof(T);
  constexpr int NT = (((TY + 2) * ((TX + 2 + kV - 1) / kV) + 63) / 64) * 64;
  const int tiles = ((nx - 2 + TX - 1) / TX) * ((ny - 2 + TY - 1) / TY) * ((nz - 2 + KC - 1) / KC);
  const long total = (long)tiles * nb;
  const unsigned g = (unsigned)(((total + kXCD - 1) / kXCD) * kXCD);
  k_gsrb7_vfused<T, TY, KC, TX, NT, ZX><<<g, NT, 0, s>>>(nx, ny, nz, B, nb, first);
}

// The vector-form shapes (70TP or 7TPP: TY rows x P planes) need rows of at
// least 16 bytes plus the two boundary points.
static bool vector_shape(int rp) {
  switch (rp) {
    case 7048: case 7088: case 7410: case 7412: case 7416: case 7816:
      return true;
    default:
      return false;
  }
}

// Launches the vector form of shape rp (vector_shape(rp) holds).
template <class T, int TX, bool ZX>
void vfused_dispatch(hipStream_t s, int rp, int nx, int ny, int nz, const Batch<SweepArgs<T>>& B,
                     int nb, int first) {
  switch (rp) {
    case 7048: vfused_sweep<T, 4, 8, TX, ZX>(s, nx, ny, nz, B, nb, first); break;
    case 7088: vfused_sweep<T, 8, 8, TX, ZX>(s, nx, ny, nz, B, nb, first); break;
    case 7410: vfused_sweep<T, 4, 10, TX, ZX>(s, nx, ny, nz, B, nb, first); break;
    case 7412: vfused_sweep<T, 4, 12, TX, ZX>(s, nx, ny, nz, B, nb, first); break;
    case 7416: vfused_sweep<T, 4, 16, TX, ZX>(s, nx, ny, nz, B, nb, first); break;
    default: vfused_sweep<T, 8, 16, TX, ZX>(s, nx, ny, nz, B, nb, first); break;
  }
}

// The scalar shapes 4 rows x 8 planes and 8 rows x 8 planes with 512
// threads and every vector shape also come in the zero-start form; other
// shapes start from a cleared x.
static bool zero_start_shape(int rp) { return rp == 48 || rp == 5088 || vector_shape(rp); }

static bool valid_tile(int v) {
  const int w = v / 10000, rp = v % 10000;
  const bool shape = rp == 416 || rp == 48 || rp == 816 || rp == 88 || rp == 84 || rp == 1616 ||
                     rp == 832 || rp == 44 || rp == 68 || rp == 5088 || rp == 5084 ||
                     rp == 5816 || vector_shape(rp);
  return shape && (w == 0 || w == 64 || w == 80 || w == 96);
}

// tile shape of the fused sweep: rows x planes per workgroup
// An APBS_PMG_FUSED_TILE shape, or 0.
static int fused_tile_env() {
  static int t = [] {
    const char* e = getenv("APBS_PMG_FUSED_TILE");
    if (!e) return 0;
    const int v = atoi(e);
    return valid_tile(v) ? v : 0;
  }();
  return t;
}

// Default shape for a grid and word size: the vector form, with the rows x
// planes per workgroup that were fastest in standalone sweeps (us per sweep;
// the best shape moves with how the workgroups fill the chip):
//   FP32: up to 2.5M points 4 x 8 (65^3 12.2, 97^3 15.0, 129^3 20.3); up to
//         5.5M 4 x 10 (161^3 29.7 against 36.5 for 4 x 8 and 50.2 for the
//         scalar form); up to 20M 8 x 16 (193^3 49.3 against 70.8 for the
//         scalar 8 x 8 with 512 threads); above 8 x 16 in 96-column tiles
//         (353^3 403, against 421 in 64-column tiles and 573 scalar).
//   FP64: up to 0.5M points 4 x 8 (65^3 12.9); up to 1M 4 x 10 (97^3 18.7);
//         up to 3M 4 x 12 (129^3 31.8); up to 5.5M 4 x 10 (161^3 48.9 against
//         65.4 scalar); up to 20M 4 x 12 (193^3 125.4 against 137.9); above
//         4 x 12 in 80-column tiles (353^3 786, against 822 for 4 x 16 in
//         64-column tiles and 931 scalar).
// The column width otherwise follows fused_width; at 353^3 the widths above
// leave fewer idle lanes (a run is 16 bytes, the workgroup whole wavefronts).
int fused_tile(int nx, int ny, int nz, int word) {
  if (const int t = fused_tile_env()) return t;
  const long n = (long)nx * ny * nz;
  if (word == 4) {
    if (n <= 2500000L) return 7048;
    if (n <= 5500000L) return 7410;
    return n <= 20000000L ? 7816 : 967816;
  }
  if (n <= 500000L) return 7048;
  if (n <= 1000000L) return 7410;
  if (n <= 3000000L) return 7412;
  if (n <= 5500000L) return 7410;
  return n <= 20000000L ? 7412 : 807412;
}

// Whether the default fused sweep of a grid runs in the vector form.
bool fused_vector(int nx, int ny, int nz, int word) {
  return vector_shape(fused_tile(nx, ny, nz, word) % 10000) && nx >= 16 / word + 2;
}

// Tile width for an x extent: the one of 64 / 80 / 96 columns that leaves
// the fewest idle columns in the last tile (64 on a tie).
int fused_width(int nx) {
  int best = 64, waste = 1 << 30;
  for (int w : {64, 80, 96}) {
    const int m = ((nx - 2 + w - 1) / w) * w - (nx - 2);
    if (m < waste) {
      waste = m;
      best = w;
    }
  }
  return best;
}

template <class T, int TX>
static void fused_sweep_w(hipStream_t s, int rp, int nx, int ny, int nz,
                          const Batch<SweepArgs<T>>& B, int nb, int first, bool zero_x) {
  if (zero_x) {
    if (rp == 48)
      fused_sweep<T, 4, 8, TX, kFT, true>(s, nx, ny, nz, B, nb, first);
    else if (vector_shape(rp))
      vfused_dispatch<T, TX, true>(s, rp, nx, ny, nz, B, nb, first);
    else
      fused_sweep<T, 8, 8, TX, 512, true>(s, nx, ny, nz, B, nb, first);
    return;
  }
  if (vector_shape(rp)) {
    vfused_dispatch<T, TX, false>(s, rp, nx, ny, nz, B, nb, first);
    return;
  }
  switch (rp) {
    case 416: fused_sweep<T, 4, 16, TX>(s, nx, ny, nz, B, nb, first); break;
    case 48: fused_sweep<T, 4, 8, TX>(s, nx, ny, nz, B, nb, first); break;
    case 44: fused_sweep<T, 4, 4, TX>(s, nx, ny, nz, B, nb, first); break;
    // 512-thread workgroups: 8 or 12 rows with one F round and one S round
    case 5088: fused_sweep<T, 8, 8, TX, 512>(s, nx, ny, nz, B, nb, first); break;
    case 5084: fused_sweep<T, 8, 4, TX, 512>(s, nx, ny, nz, B, nb, first); break;
    case 5816: fused_sweep<T, 8, 16, TX, 512>(s, nx, ny, nz, B, nb, first); break;
    case 68: fused_sweep<T, 6, 8, TX>(s, nx, ny, nz, B, nb, first); break;
    case 88: fused_sweep<T, 8, 8, TX>(s, nx, ny, nz, B, nb, first); break;
    case 84: fused_sweep<T, 8, 4, TX>(s, nx, ny, nz, B, nb, first); break;
    case 1616: fused_sweep<T, 16, 16, TX>(s, nx, ny, nz, B, nb, first); break;
    case 832: fused_sweep<T, 8, 32, TX>(s, nx, ny, nz, B, nb, first); break;
    default: fused_sweep<T, 8, 16, TX>(s, nx, ny, nz, B, nb, first); break;
  }
}

template <class T>
void gsrb7_fused_batch(hipStream_t s, int nx, int ny, int nz, const LevelArgs<T>* c, int nb,
                       int nu, int iadjoint, int tile_shape, bool zero_x) {
  check_batch(nb);
  const int first = iadjoint ? 0 : 1;
  const int tile =
      tile_shape > 0 && valid_tile(tile_shape) ? tile_shape
                                               : fused_tile(nx, ny, nz, (int)sizeof(T));
  int rp = tile % 10000;
  const int w = tile / 10000 ? tile / 10000 : fused_width(nx);
  if (vector_shape(rp) && nx < 16 / (int)sizeof(T) + 2) rp = 48;
  if (zero_x && !zero_start_shape(rp)) {
    for (int b = 0; b < nb; b++) fill<T>(s, c[b].x, (size_t)nx * ny * nz, T(0));
    zero_x = false;
  }
  for (int it = 0; it < nu; it++) {
    // sweep it reads x (even it) or xs and writes the other one
    Batch<SweepArgs<T>> B;
    for (int b = 0; b < nb; b++) {
      const LevelArgs<T>& a = c[b];
      B.v[b] = SweepArgs<T>{a.A.a[1], a.A.a[2], a.A.a[3], a.D, a.f,
                            (it & 1) ? a.xs : a.x, (it & 1) ? a.x : a.xs};
    }
    const bool zx = zero_x && it == 0;
    if (w == 80)
      fused_sweep_w<T, 80>(s, rp, nx, ny, nz, B, nb, first, zx);
    else if (w == 96)
      fused_sweep_w<T, 96>(s, rp, nx, ny, nz, B, nb, first, zx);
    else
      fused_sweep_w<T, 64>(s, rp, nx, ny, nz, B, nb, first, zx);
  }
  if (nu & 1)
    for (int b = 0; b < nb; b++) copy_interior<T>(s, nx, ny, nz, c[b].xs, c[b].x);
}

template <class T>
void gsrb7_fused(hipStream_t s, int nx, int ny, int nz, const Ops<T>& A, const T* D, const T* fc,
                 T* x, T* tmp, int nu, int iadjoint, int tile_shape) {
  const LevelArgs<T> c{A, D, fc, x, tmp, nullptr};
  gsrb7_fused_batch<T>(s, nx, ny, nz, &c, 1, nu, iadjoint, tile_shape);
}

int gsrb7_fused_launches(int nu) { return nu + (nu & 1); }

template <class T>
void gsrb7(hipStream_t s, int nx, int ny, int nz, const Ops<T>& A, const T* D, const T* fc, T* x,
           int nu, int iadjoint) {
  const int npair = (nx - 1) / 2;
  const long cnt = (long)npair * (ny - 2);
  dim3 g((unsigned)((cnt + kBlock - 1) / kBlock), (unsigned)(nz - 2));
  for (int it = 0; it < nu; it++) {
    // red (par 1) then black (par 0); swapped for the adjoint sweep
    const int first = iadjoint ? 0 : 1;
    k_gsrb7<T><<<g, kBlock, 0, s>>>(nx, ny, nz, A.a[1], A.a[2], A.a[3], D, fc, x, first);
    k_gsrb7<T><<<g, kBlock, 0, s>>>(nx, ny, nz, A.a[1], A.a[2], A.a[3], D, fc, x, 1 - first);
  }
}

int gsrb7_launches(int nu) { return 2 * nu; }

template <class T>
static Batch<Args27<T>> args27(const LevelArgs<T>* c, int nb) {
  Batch<Args27<T>> B;
  for (int b = 0; b < nb; b++)
    B.v[b] = Args27<T>{to27(c[b].A), c[b].D, c[b].f, c[b].x, c[b].r, nullptr};
  return B;
}

template <class T>
void gsrb27_batch(hipStream_t s, int nx, int ny, int nz, const LevelArgs<T>* c, int nb, int nu,
                  int iadjoint) {
  check_batch(nb);
  const Batch<Args27<T>> B = args27(c, nb);
  for (int it = 0; it < nu; it++)
    for (int cc = 0; cc < 8; cc++) {
      const int col = iadjoint ? 7 - cc : cc;
      const int i0 = (col & 1) ? 1 : 2, j0 = ((col >> 1) & 1) ? 1 : 2, k0 = ((col >> 2) & 1) ? 1 : 2;
      const int ni = (nx - 2 - i0) / 2 + 1, nj = (ny - 2 - j0) / 2 + 1, nk = (nz - 2 - k0) / 2 + 1;
      if (ni <= 0 || nj <= 0 || nk <= 0) continue;
      dim3 g((unsigned)(((long)ni * nj + kBlock - 1) / kBlock), (unsigned)nk, (unsigned)nb);
      k_gsrb27<T><<<g, kBlock, 0, s>>>(nx, ny, nz, B, col);
    }
}

template <class T>
void gsrb27(hipStream_t s, int nx, int ny, int nz, const Ops<T>& A, const T* D, const T* fc, T* x,
            int nu, int iadjoint) {
  const LevelArgs<T> c{A, D, fc, x, nullptr, nullptr};
  gsrb27_batch<T>(s, nx, ny, nz, &c, 1, nu, iadjoint);
}

// The colour-pair sweeps with the vector kernel of kV-element runs.
template <class T, int kV>
static void vpair_launch(hipStream_t s, int nx, int ny, int nz, const Batch<Args27<T>>& B,
                         int nb, int nu, int iadjoint) {
  const int R = (nx - 2 + kV - 1) / kV;
  // workgroup size: whole rows, fewest idle lanes
  int NT = kBlock, idle = 1 << 30;
  for (int t = 64; t <= kBlock; t += 64) {
    if (t < R) continue;
    const int id = t - (t / R) * R;
    if (id * kBlock < idle * t) {  // compare idle fractions
      idle = id;
      NT = t;
    }
  }
  const int RB = NT / R;
  for (int it = 0; it < nu; it++)
    for (int h = 0; h < 4; h++) {
      // forward: (0,1) (2,3) (4,5) (6,7); adjoint: (7,6) (5,4) (3,2) (1,0)
      const int col = iadjoint ? 7 - 2 * h : 2 * h;
      const int pa = col & 1, cj = (col >> 1) & 1, ck = (col >> 2) & 1;
      const int j0 = cj ? 1 : 2, k0 = ck ? 1 : 2;
      if (j0 > ny - 2 || k0 > nz - 2) continue;
      const long rows = (long)((ny - 2 - j0) / 2 + 1) * ((nz - 2 - k0) / 2 + 1);
      const dim3 g((unsigned)((rows + RB - 1) / RB), 1, (unsigned)nb);
      k_gsrb27_vpair<T, kV><<<g, NT, (size_t)RB * nx * sizeof(T), s>>>(nx, ny, nz, B, pa, cj, ck,
                                                                       R, RB);
    }
}

// The colour-pair form.  Form 0 (the default): the vector kernel where a row
// holds at least one run plus its boundary points and fits one workgroup,
// with 16-byte runs in FP64 and 8-byte runs in FP32 (16-byte runs for rows
// of 80 to 90 points), else the one-point-per-lane kernel; form 1 that
// kernel; form 2 the vector kernel with 8-byte runs (FP32).  Per sweep (4
// launches), us, FP32 8-byte runs / 16-byte runs / one point per lane: 57^3
// 14.9 / - / 16.4, 65^3 15.3 / 19.9 / 16.6, 81^3 23.7 / 23.1 / 28.1, 97^3
// 29.2 / 31.0 / 40.1, 113^3 43.1 / 45.0 / 57.4, 129^3 53.2 / 54.9 / 75.1;
// FP64 16-byte runs / one point per lane: 65^3 20.8 / 23.3, 81^3 35.4 /
// 40.3, 97^3 45.6 / 79.6, 129^3 95.2 / 162.7.
template <class T>
void gsrb27_pairs_form(hipStream_t s, int nx, int ny, int nz, const LevelArgs<T>* c, int nb,
                       int nu, int iadjoint, int form) {
  check_batch(nb);
  const Batch<Args27<T>> B = args27(c, nb);
  constexpr int kV = 16 / (int)sizeof(T), kV8 = 8 / (int)sizeof(T);
  const bool wide = sizeof(T) == 8 || (nx >= 80 && nx <= 90);
  if (form == 0 && wide && nx >= kV + 2 && (nx - 2 + kV - 1) / kV <= kBlock) {
    vpair_launch<T, kV>(s, nx, ny, nz, B, nb, nu, iadjoint);
    return;
  }
  if constexpr (kV8 >= 2) {
    if ((form == 2 || (form == 0 && !wide)) && nx >= kV8 + 2 &&
        (nx - 2 + kV8 - 1) / kV8 <= kBlock) {
      vpair_launch<T, kV8>(s, nx, ny, nz, B, nb, nu, iadjoint);
      return;
    }
  }
  for (int it = 0; it < nu; it++)
    for (int h = 0; h < 4; h++) {
      const int col = iadjoint ? 7 - 2 * h : 2 * h;
      const int pa = col & 1, cj = (col >> 1) & 1, ck = (col >> 2) & 1;
      const int j0 = cj ? 1 : 2, k0 = ck ? 1 : 2;
      if (j0 > ny - 2 || k0 > nz - 2) continue;
      const long rows = (long)((ny - 2 - j0) / 2 + 1) * ((nz - 2 - k0) / 2 + 1);
      const dim3 g((unsigned)((rows * 64 + kBlock - 1) / kBlock), 1, (unsigned)nb);
      k_gsrb27_pair<T><<<g, kBlock, 0, s>>>(nx, ny, nz, B, pa, cj, ck);
    }
}

template <class T>
void gsrb27_pairs_batch(hipStream_t s, int nx, int ny, int nz, const LevelArgs<T>* c, int nb,
                        int nu, int iadjoint) {
  gsrb27_pairs_form<T>(s, nx, ny, nz, c, nb, nu, iadjoint, 0);
}

template <class T>
void gsrb27_pairs(hipStream_t s, int nx, int ny, int nz, const Ops<T>& A, const T* D,
                  const T* fc, T* x, int nu, int iadjoint) {
  const LevelArgs<T> c{A, D, fc, x, nullptr, nullptr};
  gsrb27_pairs_batch<T>(s, nx, ny, nz, &c, 1, nu, iadjoint);
}

// Workgroups of a plane-parity launch with TX x TY tiles.
static long gs27_groups(int nx, int ny, int nz, int tx, int ty) {
  return (long)((nx - 2 + tx - 1) / tx) * ((ny - 2 + ty - 1) / ty) * ((nz - 2) / 2);
}

int gsrb27_planes_tile(int nx, int ny, int nz) {
  if (gs27_groups(nx, ny, nz, 32, 32) >= 256) return 32;
  if (gs27_groups(nx, ny, nz, 16, 16) >= 64) return 16;
  return 8;
}

template <class T>
void gsrb27_planes_batch(hipStream_t s, int nx, int ny, int nz, const LevelArgs<T>* c, int nb,
                         int nu, int iadjoint, int tile) {
  check_batch(nb);
  if (nx < 3 || ny < 3 || nz < 3) return;
  if (tile != 8 && tile != 16 && tile != 32) tile = gsrb27_planes_tile(nx, ny, nz);
  for (int it = 0; it < nu; it++) {
    // sweep it reads x (even it) or xs and writes the other one
    const int first = iadjoint ? 1 : 0;
    for (int h = 0; h < 2; h++) {
      const int ck = h == 0 ? first : 1 - first;
      const int k0 = ck ? 1 : 2;
      if (k0 > nz - 2) continue;
      const int nk = (nz - 2 - k0) / 2 + 1;
      Batch<Args27<T>> B;
      for (int b = 0; b < nb; b++) {
        T* in = (it & 1) ? c[b].xs : c[b].x;
        T* out = (it & 1) ? c[b].x : c[b].xs;
        B.v[b] = Args27<T>{to27(c[b].A), c[b].D, c[b].f, in, h == 0 ? in : out, out};
      }
      const unsigned tiles = (unsigned)(((nx - 2 + tile - 1) / tile) * ((ny - 2 + tile - 1) / tile));
      const dim3 g(tiles, (unsigned)nk, (unsigned)nb);
      if (tile == 32)
        k_gs27_plane<T, 32, 32, kBlock><<<g, kBlock, 0, s>>>(nx, ny, nz, B, ck, iadjoint);
      else if (tile == 16)
        k_gs27_plane<T, 16, 16, kBlock><<<g, kBlock, 0, s>>>(nx, ny, nz, B, ck, iadjoint);
      else
        k_gs27_plane<T, 8, 8, kBlock><<<g, kBlock, 0, s>>>(nx, ny, nz, B, ck, iadjoint);
    }
  }
  if (nu & 1)
    for (int b = 0; b < nb; b++) copy_interior<T>(s, nx, ny, nz, c[b].xs, c[b].x);
}

template <class T>
void gsrb27_planes(hipStream_t s, int nx, int ny, int nz, const Ops<T>& A, const T* D,
                   const T* fc, T* x, T* xs, int nu, int iadjoint, int tile) {
  const LevelArgs<T> c{A, D, fc, x, xs, nullptr};
  gsrb27_planes_batch<T>(s, nx, ny, nz, &c, 1, nu, iadjoint, tile);
}

bool gsrb27_use_planes(int nx, int ny, int nz) { return nx <= 49 && ny <= 49 && nz >= 3; }

bool gsrb27_use_pairs(int nx) { return nx >= 3 && nx <= 130; }


template <class T>
void resid_batch(hipStream_t s, int st, int nx, int ny, int nz, const LevelArgs<T>* c, int nb) {
  check_batch(nb);
  dim3 g = grid_interior(nx, ny, nz);
  g.z = (unsigned)nb;
  const Batch<Args27<T>> B = args27(c, nb);
  constexpr int kV = 16 / (int)sizeof(T);
  if (st == 7 && nx >= kV + 2) {
    const long lanes = (long)((nx - 2 + kV - 1) / kV) * (ny - 2) * (nz - 2);
    const dim3 gv((unsigned)((lanes + kBlock - 1) / kBlock), 1, (unsigned)nb);
    k_resid7v<T><<<gv, kBlock, 0, s>>>(nx, ny, nz, B);
  } else if (st == 7)
    k_resid7<T><<<g, kBlock, 0, s>>>(nx, ny, nz, B);
  else
    k_resid27<T><<<g, kBlock, 0, s>>>(nx, ny, nz, B);
}

template <class T>
void resid(hipStream_t s, int st, int nx, int ny, int nz, const Ops<T>& A, const T* D, const T* fc,
           const T* x, T* r) {
  const LevelArgs<T> c{A, D, fc, const_cast<T*>(x), nullptr, r};
  resid_batch<T>(s, st, nx, ny, nz, &c, 1);
}

void nresid7(hipStream_t s, int nx, int ny, int nz, const Ops<double>& A, const double* cc,
             const double* fc, const double* x, double* r, const Ions& ions) {
  k_nresid7<<<grid_interior(nx, ny, nz), kBlock, 0, s>>>(nx, ny, nz, A.a[0], A.a[1], A.a[2],
                                                         A.a[3], cc, fc, x, r, ions);
}

void dcvec(hipStream_t s, size_t n, const double* cc, const double* x, double* out,
           const Ions& ions, const double* oC, double* D, float* Df) {
  k_dcvec<<<nblocks(n), kBlock, 0, s>>>(n, cc, x, out, ions, oC, D, Df);
}

template <class T>
void restrict3_batch(hipStream_t s, int nxf, int nyf, int nzf, int nxc, int nyc, int nzc,
                     const TransferArgs<T>* c, int nb) {
  check_batch(nb);
  Batch<TransferArgs<T>> B;
  for (int b = 0; b < nb; b++) B.v[b] = c[b];
  dim3 g = grid_interior(nxc, nyc, nzc);
  g.z = (unsigned)nb;
  k_restrict<T><<<g, kBlock, 0, s>>>(nxf, nyf, nzf, nxc, nyc, nzc, B);
}

template <class T>
void restrict3(hipStream_t s, int nxf, int nyf, int nzf, int nxc, int nyc, int nzc, const T* in,
               T* out, T* zero) {
  const TransferArgs<T> c{in, out, zero, nullptr};
  restrict3_batch<T>(s, nxf, nyf, nzf, nxc, nyc, nzc, &c, 1);
}

template <class T>
void prolong_add_batch(hipStream_t s, int nxc, int nyc, int nzc, int nxf, int nyf, int nzf,
                       const TransferArgs<T>* c, int nb) {
  check_batch(nb);
  Batch<TransferArgs<T>> B;
  for (int b = 0; b < nb; b++) B.v[b] = c[b];
  dim3 g = grid_interior(nxf, nyf, nzf);
  g.z = (unsigned)nb;
  k_prolong_add<T><<<g, kBlock, 0, s>>>(nxc, nyc, nzc, nxf, nyf, nzf, B);
}

template <class T>
void prolong_add(hipStream_t s, int nxc, int nyc, int nzc, int nxf, int nyf, int nzf, const T* xc,
                 T* xf, const double* damp) {
  const TransferArgs<T> c{xc, xf, nullptr, damp};
  prolong_add_batch<T>(s, nxc, nyc, nzc, nxf, nyf, nzf, &c, 1);
}

template <class T>
void nrm1(hipStream_t s, int nx, int ny, int nz, const T* x, double* part, double* out,
          double* hout) {
  k_nrm1_part<T><<<kParts, kBlock, 0, s>>>(nx, ny, nz, x, part);
  {
    Batch<FinishArgs> F;
    F.v[0] = FinishArgs{part, out, hout};
    k_finish<<<1, kBlock, 0, s>>>(F, 0, kParts);
  }
}

template <class T>
void resid7_nrm1(hipStream_t s, int nx, int ny, int nz, const Ops<T>& A, const T* D, const T* fc,
                 const T* x, T* r, double* part, double* out, double* hout, float* r32) {
  k_resid7_nrm<T><<<kParts, kBlock, 0, s>>>(nx, ny, nz, A.a[1], A.a[2], A.a[3], D, fc, x, r, r32,
                                            part);
  {
    Batch<FinishArgs> F;
    F.v[0] = FinishArgs{part, out, hout};
    k_finish<<<1, kBlock, 0, s>>>(F, 0, kParts);
  }
}

void nrm1_to32(hipStream_t s, int nx, int ny, int nz, const double* x, float* y, double* part,
               double* out, double* hout) {
  k_nrm1_to32<<<kParts, kBlock, 0, s>>>(nx, ny, nz, x, y, part);
  Batch<FinishArgs> F;
  F.v[0] = FinishArgs{part, out, hout};
  k_finish<<<1, kBlock, 0, s>>>(F, 0, kParts);
}

void nresid7_xpay_nrm1(hipStream_t s, int nx, int ny, int nz, const Ops<double>& A,
                       const double* cc, const double* fc, const double* x, double a,
                       const double* z, double* y, double* r, const Ions& ions, double* part,
                       double* out, double* hout) {
  k_nresid7_xpay<<<kParts, kBlock, 0, s>>>(nx, ny, nz, A.a[0], A.a[1], A.a[2], A.a[3], cc, fc, x,
                                           a, z, y, r, ions, part);
  Batch<FinishArgs> F;
  F.v[0] = FinishArgs{part, out, hout};
  k_finish<<<1, kBlock, 0, s>>>(F, 0, kParts);
}

void jres_trial_nrm1(hipStream_t s, int nx, int ny, int nz, const Ops<double>& A,
                     const double* D, const double* cc, const double* fc, const double* fz,
                     const double* x, double a, const double* z, const float* zf, double* y,
                     double* rn, float* r32, float* rn32, const Ions& ions, double* part,
                     double* out, double* hout, double* out2, double* hout2) {
  if (zf)
    k_jres_trial<true><<<kParts, kBlock, 0, s>>>(nx, ny, nz, A.a[0], A.a[1], A.a[2], A.a[3], D, cc,
                                                 fc, fz, x, a, z, zf, y, rn, r32, rn32, ions,
                                                 part);
  else
    k_jres_trial<false><<<kParts, kBlock, 0, s>>>(nx, ny, nz, A.a[0], A.a[1], A.a[2], A.a[3], D,
                                                  cc, fc, fz, x, a, z, zf, y, rn, r32, rn32, ions,
                                                  part);
  Batch<FinishArgs> F;
  F.v[0] = FinishArgs{part, out, hout};
  F.v[1] = FinishArgs{part + kParts, out2, hout2};
  k_finish<<<2, kBlock, 0, s>>>(F, 0, kParts);
}

void nresid7_nrm1(hipStream_t s, int nx, int ny, int nz, const Ops<double>& A, const double* cc,
                  const double* fc, const double* x, double* r, const Ions& ions, double* part,
                  double* out, double* hout) {
  k_nresid7_nrm<<<kParts, kBlock, 0, s>>>(nx, ny, nz, A.a[0], A.a[1], A.a[2], A.a[3], cc, fc, x, r,
                                          ions, part);
  {
    Batch<FinishArgs> F;
    F.v[0] = FinishArgs{part, out, hout};
    k_finish<<<1, kBlock, 0, s>>>(F, 0, kParts);
  }
}

template <class T>
void damp_factor_batch(hipStream_t s, int nx, int ny, int nz, const DampArgs<T>* c, int nb) {
  check_batch(nb);
  Batch<DampPart<T>> B;
  Batch<FinishArgs> F;
  for (int b = 0; b < nb; b++) {
    B.v[b] = DampPart<T>{to27(c[b].A), c[b].D, c[b].x, c[b].w0, c[b].part};
    F.v[b] = FinishArgs{c[b].part, c[b].out, nullptr};
  }
  // a level of fewer than kParts * kBlock interior points leaves the blocks
  // past its last point empty (partial sums +0.0): they are not launched
  const long n = (long)(nx - 2) * (ny - 2) * (nz - 2);
  const int np = (int)std::min<long>(kParts, std::max<long>(1, (n + kBlock - 1) / kBlock));
  k_damp_part<T><<<dim3(np, 1, nb), kBlock, 0, s>>>(nx, ny, nz, B);
  k_finish<<<nb, kBlock, 0, s>>>(F, 1, np);
}

template <class T>
void damp_factor(hipStream_t s, int nx, int ny, int nz, const Ops<T>& A, const T* D, const T* x,
                 const T* w0, double* part, double* out) {
  const DampArgs<T> c{A, D, x, w0, part, out};
  damp_factor_batch<T>(s, nx, ny, nz, &c, 1);
}

template <class T>
void fill(hipStream_t s, T* x, size_t n, T v) {
  k_fill<T><<<nblocks(n), kBlock, 0, s>>>(x, n, v);
}

template <class T>
void axpy_interior(hipStream_t s, int nx, int ny, int nz, double a, const T* x, T* y) {
  k_axpy_int<T><<<grid_interior(nx, ny, nz), kBlock, 0, s>>>(nx, ny, nz, a, x, y);
}

void xpay_interior(hipStream_t s, int nx, int ny, int nz, const double* x, double a,
                   const double* z, double* y) {
  k_xpay_int<<<grid_interior(nx, ny, nz), kBlock, 0, s>>>(nx, ny, nz, x, a, z, y);
}

template <class T>
void copy_interior(hipStream_t s, int nx, int ny, int nz, const T* x, T* y) {
  k_copy_int<T><<<grid_interior(nx, ny, nz), kBlock, 0, s>>>(nx, ny, nz, x, y);
}

template <class T>
void add_diag(hipStream_t s, size_t n, const T* oC, const T* cc, T* D, float* Df) {
  k_add_diag<T><<<nblocks(n), kBlock, 0, s>>>(n, oC, cc, D, Df);
}

void ZeroList::add(void* ptr, size_t bytes) {
  if (!ptr || bytes < 4) return;
  if (z.n == kZeroMax) flush();
  z.p[z.n] = (unsigned*)ptr;
  z.words[z.n] = bytes / 4;
  z.nx[z.n] = z.ny[z.n] = z.nz[z.n] = 0;
  z.wpe[z.n] = 0;
  z.blk[z.n + 1] = z.blk[z.n] + (unsigned)((bytes / 4 + kZeroChunk - 1) / kZeroChunk);
  z.n++;
}

void ZeroList::add_shell(void* ptr, int nx, int ny, int nz, size_t elem_bytes) {
  if (!ptr || nx < 1 || ny < 1 || nz < 1) return;
  if (nx < 3 || ny < 3 || nz < 3) {  // every point is on the boundary
    add(ptr, (size_t)nx * ny * nz * elem_bytes);
    return;
  }
  if (z.n == kZeroMax) flush();
  const size_t pts = 2 * (size_t)nx * ny + (size_t)(nz - 2) * (2 * (size_t)nx + 2 * (size_t)(ny - 2));
  z.p[z.n] = (unsigned*)ptr;
  z.words[z.n] = pts;
  z.nx[z.n] = nx;
  z.ny[z.n] = ny;
  z.nz[z.n] = nz;
  z.wpe[z.n] = (int)(elem_bytes / 4);
  z.blk[z.n + 1] = z.blk[z.n] + (unsigned)((pts + kBlock - 1) / kBlock);
  z.n++;
}

void ZeroList::flush() {
  if (z.n > 0 && z.blk[z.n] > 0) k_zero_many<<<z.blk[z.n], kBlock, 0, s>>>(z);
  z.n = 0;
  z.blk[0] = 0;
}

void add_diag_levels(hipStream_t s, const DiagLevels& L) {
  k_add_diag_levels<<<nblocks(L.off[L.nl]), kBlock, 0, s>>>(L);
}

void axpy_interior_f32(hipStream_t s, int nx, int ny, int nz, double a, const float* x,
                       double* y, bool assign) {
  k_axpy_int_f32<<<grid_interior(nx, ny, nz), kBlock, 0, s>>>(nx, ny, nz, a, x, y,
                                                               assign ? 1 : 0);
}

void scale(hipStream_t s, size_t n, double a, const double* x, double* y) {
  k_scale<<<nblocks(n), kBlock, 0, s>>>(n, a, x, y);
}

void convert_many(hipStream_t s, const ConvArgs& c) {
  if (c.m <= 0 || c.n == 0) return;
  if (c.m > kConvMax) {
    fprintf(stderr, "convert_many: %d arrays, at most %d\n", c.m, kConvMax);
    abort();
  }
  k_convert_many<<<dim3(nblocks(c.n), (unsigned)c.m), kBlock, 0, s>>>(c);
}

template <class Tin, class Tout>
void convert(hipStream_t s, size_t n, const Tin* in, Tout* out) {
  k_convert<Tin, Tout><<<nblocks(n), kBlock, 0, s>>>(n, in, out);
}

void set_faces(hipStream_t s, int nx, int ny, int nz, double* u, const double* gx,
               const double* gy, const double* gz) {
  k_faces<<<nblocks((size_t)ny * nz), kBlock, 0, s>>>(nx, ny, nz, u, gx, gy, gz, 0);
  k_faces<<<nblocks((size_t)nx * nz), kBlock, 0, s>>>(nx, ny, nz, u, gx, gy, gz, 1);
  k_faces<<<nblocks((size_t)nx * ny), kBlock, 0, s>>>(nx, ny, nz, u, gx, gy, gz, 2);
}

void buildA_fv(hipStream_t s, int nx, int ny, int nz, const double* xf, const double* yf,
               const double* zf, const double* gx, const double* gy, const double* gz,
               const double* a1, const double* a2, const double* a3, const double* ccf,
               const double* fcf, double* oC, double* cc, double* fc, double* oE, double* oN,
               double* uC) {
  k_buildA<<<grid_interior(nx, ny, nz), kBlock, 0, s>>>(nx, ny, nz, xf, yf, zf, gx, gy, gz, a1,
                                                        a2, a3, ccf, fcf, oC, cc, fc, oE, oN, uC);
}

void galerkin(hipStream_t s, int st, int nxf, int nyf, int nzf, const Ops<double>& Af, int nxc,
              int nyc, int nzc, double* const ac[14], int form) {
  OutPtrs o;
  for (int k = 0; k < 14; k++) o.a[k] = ac[k];
  const long nq = (long)(nxc - 2) * (nyc - 2) * (nzc - 2);
  // Form 0 is the unrolled kernel on every level (measured on MI355X, per
  // call: 161^3 7-point fine level 118 us unrolled vs 1055 us per thread;
  // 81^3 27-point 127 vs 408; 41^3 21 vs 171 (147 grouped); 21^3 17.5 vs
  // 21 grouped).  Forms 1 and 2 stay for the tests.
  if (form == 0) form = 3;
  if (form == 3 || form == 4) {
    const long tot = ((nq + 63) / 64) * 64 * 14;
    const unsigned nb = (unsigned)((tot + kBlock - 1) / kBlock);
    if (st == 7 && form == 3)
      k_galerkin_unrolled<7, 1><<<nb, kBlock, 0, s>>>(nxf, nyf, nzf, to27(Af), nxc, nyc, nzc, o);
    else if (st == 7)
      k_galerkin_unrolled<7, 4><<<nb, kBlock, 0, s>>>(nxf, nyf, nzf, to27(Af), nxc, nyc, nzc, o);
    else if (form == 3)
      k_galerkin_unrolled<27, 1><<<nb, kBlock, 0, s>>>(nxf, nyf, nzf, to27(Af), nxc, nyc, nzc, o);
    else
      k_galerkin_unrolled<27, 4><<<nb, kBlock, 0, s>>>(nxf, nyf, nzf, to27(Af), nxc, nyc, nzc, o);
  } else if (form == 2) {
    const long groups = nq * 14;
    k_galerkin_group<<<(unsigned)((groups + kGalGroups - 1) / kGalGroups), kBlock, 0, s>>>(
        st, nxf, nyf, nzf, to27(Af), nxc, nyc, nzc, o);
  } else {
    const long tot = ((nq + 63) / 64) * 64 * 14;
    k_galerkin_point<<<(unsigned)((tot + kBlock - 1) / kBlock), kBlock, 0, s>>>(
        st, nxf, nyf, nzf, to27(Af), nxc, nyc, nzc, o);
  }
}

void dense_assemble_invert_batch(hipStream_t s, int nx, int ny, int nz, int N,
                                 const InvertArgs* c, int nb, int form) {
  check_batch(nb);
  Batch<InvertArgs> B;
  for (int b = 0; b < nb; b++) B.v[b] = c[b];
  k_dense_assemble<<<dim3((N + kBlock - 1) / kBlock, nb), kBlock, 0, s>>>(nx, ny, nz, B, N);
  if (N <= 64 && form != 1) {
    k_dense_invert_reg<<<nb, 256, 0, s>>>(B, N);
    return;
  }
  const size_t lds = ((size_t)N * N + 2 * N) * sizeof(double);
  if (lds <= 64 * 1024 && (size_t)N * N <= (size_t)kInvPer * 1024) {
    k_dense_invert_lds<<<nb, 1024, lds, s>>>(B, N);
    return;
  }
  const unsigned gn = (unsigned)((N + kBlock - 1) / kBlock);
  const unsigned gnn = (unsigned)(((size_t)N * N + kBlock - 1) / kBlock);
  for (int b = 0; b < nb; b++)
    for (int k = 0; k < N; k++) {
      k_gj_col<<<gn, kBlock, 0, s>>>(c[b].M, c[b].work, N, k);
      k_gj_row<<<gn, kBlock, 0, s>>>(c[b].M, c[b].work, N, k, c[b].flag);
      k_gj_elim<<<gnn, kBlock, 0, s>>>(c[b].M, c[b].work, N, k);
    }
}

void dense_assemble(hipStream_t s, int nx, int ny, int nz, const Ops<double>& A, const double* D,
                    double* M, int N) {
  Batch<InvertArgs> B;
  B.v[0] = InvertArgs{A, D, M, nullptr, nullptr};
  k_dense_assemble<<<dim3((N + kBlock - 1) / kBlock, 1), kBlock, 0, s>>>(nx, ny, nz, B, N);
}

void dense_invert(hipStream_t s, double* M, double* work, int N, int form, double* flag) {
  Batch<InvertArgs> B;
  B.v[0] = InvertArgs{Ops<double>{}, nullptr, M, work, flag};
  if (N <= 64 && form != 1) {
    k_dense_invert_reg<<<1, 256, 0, s>>>(B, N);
    return;
  }
  const size_t lds = ((size_t)N * N + 2 * N) * sizeof(double);
  if (lds <= 64 * 1024 && (size_t)N * N <= (size_t)kInvPer * 1024) {
    k_dense_invert_lds<<<1, 1024, lds, s>>>(B, N);
    return;
  }
  const unsigned gn = (unsigned)((N + kBlock - 1) / kBlock);
  const unsigned gnn = (unsigned)(((size_t)N * N + kBlock - 1) / kBlock);
  for (int k = 0; k < N; k++) {
    k_gj_col<<<gn, kBlock, 0, s>>>(M, work, N, k);
    k_gj_row<<<gn, kBlock, 0, s>>>(M, work, N, k, flag);
    k_gj_elim<<<gnn, kBlock, 0, s>>>(M, work, N, k);
  }
}

template <class T>
void dense_solve_batch(hipStream_t s, int nx, int ny, int nz, int N, const DenseArgs<T>* c,
                       int nb) {
  check_batch(nb);
  Batch<DenseArgs<T>> B;
  for (int b = 0; b < nb; b++) B.v[b] = c[b];
  if (((size_t)N * N + N) * sizeof(double) <= 64 * 1024 && N <= kDenseSmall)
    k_dense_solve<T><<<nb, 128, 0, s>>>(nx, ny, nz, B, N);
  else
    k_dense_solve_cols<T><<<dim3((N + kBlock - 1) / kBlock, nb), kBlock, 0, s>>>(nx, ny, nz, B, N);
}

template <class T>
void dense_solve(hipStream_t s, int nx, int ny, int nz, const double* Minv, int N, const T* f,
                 T* x) {
  const DenseArgs<T> c{Minv, f, x};
  dense_solve_batch<T>(s, nx, ny, nz, N, &c, 1);
}

// ------------------------------------------------------ explicit instances
#define INST(T)                                                                                 \
  template void gsrb7<T>(hipStream_t, int, int, int, const Ops<T>&, const T*, const T*, T*, int, \
                         int);                                                                  \
  template void gsrb27<T>(hipStream_t, int, int, int, const Ops<T>&, const T*, const T*, T*, int, \
                          int);                                                                 \
  template void gsrb27_planes<T>(hipStream_t, int, int, int, const Ops<T>&, const T*, const T*,  \
                                 T*, T*, int, int, int);                                        \
  template void gsrb27_pairs<T>(hipStream_t, int, int, int, const Ops<T>&, const T*, const T*,   \
                                T*, int, int);                                                  \
  template void resid<T>(hipStream_t, int, int, int, int, const Ops<T>&, const T*, const T*,     \
                         const T*, T*);                                                         \
  template void restrict3<T>(hipStream_t, int, int, int, int, int, int, const T*, T*, T*);      \
  template void prolong_add<T>(hipStream_t, int, int, int, int, int, int, const T*, T*,         \
                               const double*);                                                  \
  template void nrm1<T>(hipStream_t, int, int, int, const T*, double*, double*, double*);       \
  template void resid7_nrm1<T>(hipStream_t, int, int, int, const Ops<T>&, const T*, const T*,   \
                               const T*, T*, double*, double*, double*, float*);                \
  template void damp_factor<T>(hipStream_t, int, int, int, const Ops<T>&, const T*, const T*,   \
                               const T*, double*, double*);                                     \
  template void fill<T>(hipStream_t, T*, size_t, T);                                            \
  template void axpy_interior<T>(hipStream_t, int, int, int, double, const T*, T*);             \
  template void copy_interior<T>(hipStream_t, int, int, int, const T*, T*);                     \
  template void add_diag<T>(hipStream_t, size_t, const T*, const T*, T*, float*);               \
  template void dense_solve<T>(hipStream_t, int, int, int, const double*, int, const T*, T*);  \
  template void gsrb7_fused<T>(hipStream_t, int, int, int, const Ops<T>&, const T*, const T*, T*, \
                               T*, int, int, int);                                              \
  template void gsrb7_fused_batch<T>(hipStream_t, int, int, int, const LevelArgs<T>*, int, int,  \
                                     int, int, bool);                                           \
  template void gsrb27_batch<T>(hipStream_t, int, int, int, const LevelArgs<T>*, int, int, int); \
  template void gsrb27_pairs_batch<T>(hipStream_t, int, int, int, const LevelArgs<T>*, int, int, \
                                      int);                                                     \
  template void gsrb27_pairs_form<T>(hipStream_t, int, int, int, const LevelArgs<T>*, int, int,  \
                                     int, int);                                                 \
  template void gsrb27_planes_batch<T>(hipStream_t, int, int, int, const LevelArgs<T>*, int,     \
                                       int, int, int);                                          \
  template void resid_batch<T>(hipStream_t, int, int, int, int, const LevelArgs<T>*, int);       \
  template void restrict3_batch<T>(hipStream_t, int, int, int, int, int, int,                     \
                                   const TransferArgs<T>*, int);                                \
  template void prolong_add_batch<T>(hipStream_t, int, int, int, int, int, int,                   \
                                     const TransferArgs<T>*, int);                              \
  template void damp_factor_batch<T>(hipStream_t, int, int, int, const DampArgs<T>*, int);       \
  template void dense_solve_batch<T>(hipStream_t, int, int, int, int, const DenseArgs<T>*, int);
INST(double)
INST(float)
#undef INST
template void convert<double, float>(hipStream_t, size_t, const double*, float*);
template void convert<float, double>(hipStream_t, size_t, const float*, double*);

}  // namespace pmg
