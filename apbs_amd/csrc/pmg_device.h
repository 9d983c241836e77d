// pmg_device.h -- launch wrappers for the PMG multigrid kernels (HIP, gfx950).
//
// Grids are (nx, ny, nz), x fastest, index p = (k*ny + j)*nx + i, 0-based;
// the Dirichlet boundary is the outer layer, interior is 1..n-2 per axis.
// Operator arrays follow the reference layout of src/pmgc/gsd.c:74-87:
//   0 oC, 1 oE, 2 oN, 3 uC, 4 oNE, 5 oNW, 6 uE, 7 uW, 8 uN, 9 uS,
//   10 uNE, 11 uNW, 12 uSE, 13 uSW.
// The fine level holds 0..3 (7-point), coarse levels all 14 (27-point).
// D is the smoothing/residual diagonal oC + cc, precomputed once per operator
// so every use rounds exactly like the reference's inline (oC + cc).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>

namespace pmg {

constexpr int kMaxIon = 10;

template <class T>
struct Ops {
  const T* a[14];  // nullptr beyond index 3 on 7-point levels
};

struct Ions {
  int nion;
  double q[kMaxIon];
  double sconc[kMaxIon];  // bulk concentrations scaled by 1/(2 I)
};

// Reduction scratch: kParts partial sums per slot, results in dev scalars.
constexpr int kParts = 1024;

// ---------------------------------------------------------------- smoothers
template <class T>
void gsrb7(hipStream_t s, int nx, int ny, int nz, const Ops<T>& A, const T* D, const T* fc,
           T* x, int nu, int iadjoint);
// kernel launches one gsrb7 call of nu sweeps makes
int gsrb7_launches(int nu);
// The same nu sweeps with one fused red-black launch per sweep, ping-ponging
// between x and tmp (fine-size scratch).  Only interior points are written,
// so x and tmp must hold the same boundary values (zero in the V-cycle).
// The result is bitwise that of gsrb7 and ends in x.
// tile_shape: width*10000 + rows*100 + planes of a workgroup's block for the
// scalar form (rows and planes 416, 48, 44, 68, 816, 88, 84, 1616, 832 with
// 256 threads, or 5088, 5084, 5816 for 8 rows with 512 threads), or
// width*10000 + a vector-form code (7048, 7088: 4 or 8 rows x 8 planes;
// 7410, 7412, 7416: 4 rows x 10, 12, 16 planes; 7816: 8 rows x 16 planes;
// one lane per 16-byte run); width 64, 80, 96, or 0 for the width that fits
// nx best; <= 0 takes APBS_PMG_FUSED_TILE, else the vector shape measured
// fastest for the grid's size and the word size.
template <class T>
void gsrb7_fused(hipStream_t s, int nx, int ny, int nz, const Ops<T>& A, const T* D,
                 const T* fc, T* x, T* tmp, int nu, int iadjoint, int tile_shape = 0);
int gsrb7_fused_launches(int nu);
// whether the default fused sweep of a grid runs in the vector form (word 4 or 8)
bool fused_vector(int nx, int ny, int nz, int word);
template <class T>
void gsrb27(hipStream_t s, int nx, int ny, int nz, const Ops<T>& A, const T* D, const T* fc,
            T* x, int nu, int iadjoint);
// The same 8-colour sweeps with one launch per plane parity (the four colours
// of the even or odd planes), ping-ponging between x and xs (level-size
// scratch).  Only interior points are written, so x and xs must hold the same
// boundary values (zero on the coarse levels).  Bitwise the result of gsrb27,
// ending in x.  tile: 8, 16 or 32 (square tiles of a plane); other values
// take gsrb27_planes_tile.
template <class T>
void gsrb27_planes(hipStream_t s, int nx, int ny, int nz, const Ops<T>& A, const T* D,
                   const T* fc, T* x, T* xs, int nu, int iadjoint, int tile = 0);
int gsrb27_planes_tile(int nx, int ny, int nz);
// The same 8-colour sweeps in place with the two colours of each (j, k)
// parity (c, c ^ 1) in one launch, one wavefront per row: 4 launches per
// sweep.  Rows of up to 128 interior points (nx <= 130).  Bitwise gsrb27.
template <class T>
void gsrb27_pairs(hipStream_t s, int nx, int ny, int nz, const Ops<T>& A, const T* D,
                  const T* fc, T* x, int nu, int iadjoint);
// Whether the plane-parity form is the one for a level: it trades a quarter
// of the launches for four dependent colour passes per workgroup, which pays
// on levels up to about 49^2 points per plane (measured on MI355X, FP64 us
// per sweep: 41^3 22 against 29 per colour, 21^3 16 against 22, 11^3 14
// against 22; at 81^3 the per-colour launches are faster, 63 against 75).
bool gsrb27_use_planes(int nx, int ny, int nz);
// Whether the colour-pair form is the one for a level that is not
// plane-parity: rows of up to 128 interior points (81^3: 40 against 62 us per
// FP64 sweep, 28 against 40 FP32; 97^3: 79 against 87; 65^3: 23 against 34).
bool gsrb27_use_pairs(int nx);

// ------------------------------------------------------- residual / matvec
template <class T>
void resid(hipStream_t s, int st, int nx, int ny, int nz, const Ops<T>& A, const T* D,
           const T* fc, const T* x, T* r);
// nonlinear residual on the fine 7-point level: f + sum a x - oC x - N(x)
void nresid7(hipStream_t s, int nx, int ny, int nz, const Ops<double>& A, const double* cc,
             const double* fc, const double* x, double* r, const Ions& ions);
// cprime = N'(x) (Vdc_vecpmg), full array; with oC also the diagonal
// D = oC + cprime and, with Df, its FP32 image (add_diag's values), in one pass
void dcvec(hipStream_t s, size_t n, const double* cc, const double* x, double* out,
           const Ions& ions, const double* oC = nullptr, double* D = nullptr,
           float* Df = nullptr);

// ------------------------------------------------------------- transfers
// out = P^T in on the coarse interior; zero (optional, same coarse size) is
// set to 0 on the same points, for the coarse solution the cycle starts from.
template <class T>
void restrict3(hipStream_t s, int nxf, int nyf, int nzf, int nxc, int nyc, int nzc,
               const T* in, T* out, T* zero = nullptr);
// xf += damp[0] * P xc  (interior of the fine grid)
template <class T>
void prolong_add(hipStream_t s, int nxc, int nyc, int nzc, int nxf, int nyf, int nzf,
                 const T* xc, T* xf, const double* damp);

// ------------------------------------------------------------ reductions
// Each writes one double to out (device), deterministically.
template <class T>
void nrm1(hipStream_t s, int nx, int ny, int nz, const T* x, double* part, double* out,
          double* hout = nullptr);
// r = f + sum a x - D x on the 7-point level (stored when r != nullptr;
// r32, when given, receives (float) r on the interior) and out[0] = ||r||_1
// over the interior, bitwise equal to resid + nrm1.
template <class T>
void resid7_nrm1(hipStream_t s, int nx, int ny, int nz, const Ops<T>& A, const T* D,
                 const T* fc, const T* x, T* r, double* part, double* out, double* hout = nullptr,
                 float* r32 = nullptr);
// nrm1 of x and y = (float) x on the interior, in one pass
void nrm1_to32(hipStream_t s, int nx, int ny, int nz, const double* x, float* y, double* part,
               double* out, double* hout = nullptr);
// y = x + a z (xpay_interior) and nresid7_nrm1 at y, in one pass, bitwise the
// two calls (the boundary layers of x, z and y are zero)
void nresid7_xpay_nrm1(hipStream_t s, int nx, int ny, int nz, const Ops<double>& A,
                       const double* cc, const double* fc, const double* x, double a,
                       const double* z, double* y, double* r, const Ions& ions, double* part,
                       double* out, double* hout = nullptr);
// resid7_nrm1 of the Jacobian system at z (r stored only as its FP32 image
// r32, when given) and nresid7_xpay_nrm1 at y = x + a z, in one pass, bitwise
// the two calls; with zf given, z is read as (double) zf (z is not read);
// rn32, when given, receives (float) rn; part holds 2 * kParts partial sums,
// out / hout receive the first norm, out2 / hout2 the second
void jres_trial_nrm1(hipStream_t s, int nx, int ny, int nz, const Ops<double>& A,
                     const double* D, const double* cc, const double* fc, const double* fz,
                     const double* x, double a, const double* z, const float* zf, double* y,
                     double* rn, float* r32, float* rn32, const Ions& ions, double* part,
                     double* out, double* hout, double* out2, double* hout2);
// nresid7 followed by nrm1 of its result, in one pass.
void nresid7_nrm1(hipStream_t s, int nx, int ny, int nz, const Ops<double>& A, const double* cc,
                  const double* fc, const double* x, double* r, const Ions& ions, double* part,
                  double* out, double* hout = nullptr);
// (hout: host-mapped pinned memory that also receives out[0], so the host
// reads the norm after a stream synchronization without a copy)
// out[0] = <x, w0> / <x, A x>   (Hackbusch-Reusken damping, 27-point A)
template <class T>
void damp_factor(hipStream_t s, int nx, int ny, int nz, const Ops<T>& A, const T* D,
                 const T* x, const T* w0, double* part, double* out);

// ------------------------------------------------------------ vector kit
template <class T>
void fill(hipStream_t s, T* x, size_t n, T v);
template <class T>
void axpy_interior(hipStream_t s, int nx, int ny, int nz, double a, const T* x, T* y);
// y = x + a*z on the interior (boundary of y untouched)
void xpay_interior(hipStream_t s, int nx, int ny, int nz, const double* x, double a,
                   const double* z, double* y);
template <class T>
void copy_interior(hipStream_t s, int nx, int ny, int nz, const T* x, T* y);
// y = a * x over the whole array
void scale(hipStream_t s, size_t n, double a, const double* x, double* y);
template <class T>
void add_diag(hipStream_t s, size_t n, const T* oC, const T* cc, T* D, float* Df = nullptr);
// Arrays cleared together on one stream (sizes multiples of 4 bytes): add()
// queues a whole array, add_shell() only the outer layer of a grid array (the
// points with i, j or k on the boundary), flush() clears the queued ones in
// one launch; a full queue and the destructor flush.
constexpr int kZeroMax = 48;
struct ZeroArgs {
  int n;
  unsigned blk[kZeroMax + 1];  // first block of each array; blk[n] = blocks
  unsigned* p[kZeroMax];
  size_t words[kZeroMax];      // whole arrays: 4-byte words to clear
  int nx[kZeroMax], ny[kZeroMax], nz[kZeroMax];  // shells: grid shape (nx = 0: whole)
  int wpe[kZeroMax];           // shells: 4-byte words per element
};
struct ZeroList {
  hipStream_t s;
  ZeroArgs z{};
  explicit ZeroList(hipStream_t st) : s(st) {}
  ~ZeroList() { flush(); }
  ZeroList(const ZeroList&) = delete;
  ZeroList& operator=(const ZeroList&) = delete;
  void add(void* ptr, size_t bytes);
  void add_shell(void* ptr, int nx, int ny, int nz, size_t elem_bytes);
  void flush();
};
// add_diag of up to 24 levels in one launch; off[l] is the first element of
// level l in the concatenation, off[nl] the total; Df[l] may be null.
struct DiagLevels {
  int nl;
  size_t off[25];
  const double* oC[24];
  const double* c[24];
  double* D[24];
  float* Df[24];
};
void add_diag_levels(hipStream_t s, const DiagLevels& L);
// y += a * (double) x on the interior (x FP32, y FP64); assign: y starts at
// zero (its interior is not read)
void axpy_interior_f32(hipStream_t s, int nx, int ny, int nz, double a, const float* x,
                       double* y, bool assign = false);
template <class Tin, class Tout>
void convert(hipStream_t s, size_t n, const Tin* in, Tout* out);
// FP32 images of up to kConvMax FP64 arrays of one length in one launch.
constexpr int kConvMax = 16;
struct ConvArgs {
  int m;
  size_t n;
  const double* in[kConvMax];
  float* out[kConvMax];
};
void convert_many(hipStream_t s, const ConvArgs& c);
void set_faces(hipStream_t s, int nx, int ny, int nz, double* u, const double* gx,
               const double* gy, const double* gz);

// --------------------------------------------------------------- operators
void buildA_fv(hipStream_t s, int nx, int ny, int nz, const double* xf, const double* yf,
               const double* zf, const double* gx, const double* gy, const double* gz,
               const double* a1, const double* a2, const double* a3, const double* ccf,
               const double* fcf, double* oC, double* cc, double* fc, double* oE, double* oN,
               double* uC);
// Galerkin A_c = P^T A P (without the Helmholtz term); ac holds 14 arrays.
// form 0 picks by level size; 1 = one thread per coefficient, 2 = one 32-lane
// group per coefficient.  Both give the same bits.
void galerkin(hipStream_t s, int st, int nxf, int nyf, int nzf, const Ops<double>& Af, int nxc,
              int nyc, int nzc, double* const ac[14], int form = 0);

// ------------------------------------------------------------- batches
// Up to kMaxBatch calculations of one grid shape share each launch of the
// V-cycle kernels: a launch takes one argument set per calculation, and a
// workgroup's calculation is its blockIdx.z (a flattened workgroup index for
// the fused fine sweep, blockIdx.x or .y for the one-workgroup-per-matrix
// kernels).  Every calculation's arithmetic is the same in a batch as alone;
// the single-calculation wrappers above are batches of one.
constexpr int kMaxBatch = 8;

// One calculation's arrays on one level.  x is the solution (smoothers update
// it), xs the partner buffer of the ping-pong sweeps (fused fine sweep and
// plane-parity form), r the output of resid_batch.
template <class T>
struct LevelArgs {
  Ops<T> A;
  const T* D;
  const T* f;
  T* x;
  T* xs;
  T* r;
};
// zero_x: every x starts at zero (its contents need not be cleared: the first
// sweep does not read them)
template <class T>
void gsrb7_fused_batch(hipStream_t s, int nx, int ny, int nz, const LevelArgs<T>* c, int nb,
                       int nu, int iadjoint, int tile_shape = 0, bool zero_x = false);
template <class T>
void gsrb27_batch(hipStream_t s, int nx, int ny, int nz, const LevelArgs<T>* c, int nb, int nu,
                  int iadjoint);
template <class T>
void gsrb27_pairs_batch(hipStream_t s, int nx, int ny, int nz, const LevelArgs<T>* c, int nb,
                        int nu, int iadjoint);
// The same with the kernel chosen: form 0 the vector kernel (16-byte runs per
// lane) where the row allows it, 1 the one-point-per-lane kernel.
template <class T>
void gsrb27_pairs_form(hipStream_t s, int nx, int ny, int nz, const LevelArgs<T>* c, int nb,
                       int nu, int iadjoint, int form);
template <class T>
void gsrb27_planes_batch(hipStream_t s, int nx, int ny, int nz, const LevelArgs<T>* c, int nb,
                         int nu, int iadjoint, int tile = 0);
template <class T>
void resid_batch(hipStream_t s, int st, int nx, int ny, int nz, const LevelArgs<T>* c, int nb);
// restriction: out = P^T in, zero cleared (optional); prolongation: out += damp[0] P in
template <class T>
struct TransferArgs {
  const T* in;
  T* out;
  T* zero;
  const double* damp;
};
template <class T>
void restrict3_batch(hipStream_t s, int nxf, int nyf, int nzf, int nxc, int nyc, int nzc,
                     const TransferArgs<T>* c, int nb);
template <class T>
void prolong_add_batch(hipStream_t s, int nxc, int nyc, int nzc, int nxf, int nyf, int nzf,
                       const TransferArgs<T>* c, int nb);
// damp_factor per calculation (part: its 2 * kParts partial sums)
template <class T>
struct DampArgs {
  Ops<T> A;
  const T* D;
  const T* x;
  const T* w0;
  double* part;
  double* out;
};
template <class T>
void damp_factor_batch(hipStream_t s, int nx, int ny, int nz, const DampArgs<T>* c, int nb);
// dense_solve per calculation (N interior unknowns, the same for all)
template <class T>
struct DenseArgs {
  const double* Minv;
  const T* f;
  T* x;
};
template <class T>
void dense_solve_batch(hipStream_t s, int nx, int ny, int nz, int N, const DenseArgs<T>* c,
                       int nb);
// dense_assemble then dense_invert per calculation (form and flag as there)
struct InvertArgs {
  Ops<double> A;
  const double* D;
  double* M;
  double* work;
  double* flag;
};
void dense_assemble_invert_batch(hipStream_t s, int nx, int ny, int nz, int N,
                                 const InvertArgs* c, int nb, int form = 0);

// ----------------------------------------------------- coarsest level (dense)
// Assemble the dense (N x N) operator of the interior of a 27-point level with
// diagonal D, then invert it in place (Gauss-Jordan, SPD, one workgroup).
void dense_assemble(hipStream_t s, int nx, int ny, int nz, const Ops<double>& A, const double* D,
                    double* M, int N);
// form 0: registers for N <= 64 (one barrier per pivot), else LDS / global;
// form 1: never the register form.  flag (host-mapped, may be null) receives
// the 1-based index of a pivot that is not positive and finite (the operator
// is not positive definite); it is left alone otherwise.
void dense_invert(hipStream_t s, double* M, double* work, int N, int form = 0,
                  double* flag = nullptr);
// x(interior) = Minv * f(interior); boundary untouched
template <class T>
void dense_solve(hipStream_t s, int nx, int ny, int nz, const double* Minv, int N, const T* f,
                 T* x);

}  // namespace pmg
