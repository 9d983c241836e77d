// fill_device.h -- coefficient maps, boundary values, energies and forces on
// the device (Vpmg_fillco, bcCalc, focusFillBound, Vpmg_*Energy, Vpmg_*Force).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>

namespace pmg {

// Grid geometry of one Vpmg (pmgp->xmin.. as set by Vpmg_fillco).
struct Geo {
  int nx, ny, nz;
  double hx, hy, hz;
  double xmin, ymin, zmin, xmax, ymax, zmax;
};

// Uniform cell hash over points (atoms or SAS points), indices ascending per
// cell.
struct Cells {
  double lo[3];
  double cs;       // cell edge (A)
  int dim[3];
  const int* start;  // dim0*dim1*dim2 + 1
  const int* list;
};

// Per-tile atom lists for the order-sensitive gathers: the grid is cut into
// kTile^3-point tiles (tdim per axis, x fastest); tile t lists, in ascending
// atom index, every atom that can reach a point of the tile (a superset:
// the kernels apply the reference's own tests).
constexpr int kTile = 8;
struct Tiles {
  int tdim[3];
  const int* start;  // tdim0*tdim1*tdim2 + 1
  const int* list;
};

struct AtomsDev {
  int n;
  const double* pos;     // 3n
  const double* rad;
  const double* chg;
  const double* partID;  // n
};

// ------------------------------------------------------------------ charges
// chgm spl2 (vpmg.c:5123-5248) / spl0 (vpmg.c:4986-5089), charge already
// scaled by zmagic/(hx hy hz) per atom order of the reference.
// mode: 0 trilinear (fillcoChargeSpline1), 1 cubic B-spline (fillcoChargeSpline2),
// 2 5th-order B-spline (chgm spl4, fillcoPermanentMultipole with charges only).
void charge_map(hipStream_t s, const Geo& g, const AtomsDev& A, const Tiles& T, int mode,
                double zmagic, int bcfl_lenient, double* charge);

// ------------------------------------------------------------- coefficients
// srfm spl2 (vpmg.c:4631-4854): kappa, epsx/y/z in [0,1] windows, then the
// linear map to epsp..epsw and the ion mask.
// order 2: srfm spl2 (cubic window); 3 / 4: srfm spl3 / spl4 (polynomial
// windows), coef = kSplCoefHost doubles per atom from spline_coefficients.
void coef_spline(hipStream_t s, const Geo& g, const AtomsDev& A, const Tiles& T, double win,
                 double irad, double ionmask, double epsw, double epsp, double* kappa,
                 double* epsx, double* epsy, double* epsz, int order = 2,
                 const double* coef = nullptr);

// Per-atom window coefficients of srfm spl3 / spl4, kSplCoefHost doubles per
// atom: [0, 8) the dielectric window (atom radius +- win; also the gradient
// of the dielectric-boundary force), [8, 16) the ion window of the kappa map,
// [16, 24) the ion window of the ion-boundary force gradient.  spl3 keeps the
// reference's kappa coefficients, which are the dielectric ones divided once
// more by the ion denominator (vpmg.c:7632-7643).
constexpr int kSplCoefHost = 24;
void spline_coefficients(int order, int natoms, const double* radius, double win, double irad,
                         double* out);

// SAS points of every atom (Vacc_atomSurf, vacc.c:866-922): flags[a*nref+q]
// = 1 when the point rad*ref+apos is outside every other inflated atom.
void sas_flags(hipStream_t s, const AtomsDev& A, const Cells& C, const double* ref, int nref,
               double probe, unsigned char* flags);
// compact flagged points to xyz (order of (atom, ref) index); returns count
// on the host (synchronises).
// Writes the accessible SAS points (flag order) to a buffer obtained from
// alloc(doubles), returned in *xyz_out; returns their number.  scratch holds
// 1024 ints.
int sas_compact(hipStream_t s, const AtomsDev& A, const double* ref, int nref, double probe,
                const unsigned char* flags, int* scratch, double** xyz_out,
                double* (*alloc)(size_t));

// markSphere (vpmg.c:6444-6491) for a set of spheres, as flags: every grid
// point of a sphere gets byte 1 in the flag map of each target.  Target t
// (MarkTargets) names the grid (which: 0/1/2 = x/y/z shifted, 3 =
// unshifted), the sphere test (mode 0: atoms, radius r_a + add, strict
// on-mesh test (dielectric); mode 1: atoms, r_a + add, inflated test (ion
// accessibility); mode 2: points of radius add (SAS), no mesh test) and its
// flag map flags + t * (nx ny nz).  The flag maps must start at zero.
struct MarkTarget {
  int which, mode;
  double add;
};
struct MarkTargets {
  int nt;
  MarkTarget t[4];
};
void mark_flags(hipStream_t s, const Geo& g, int npts, const double* pos, const double* rad,
                const MarkTargets& T, unsigned char* flags);
// The molecular-surface maps from the flags (fillcoCoefMolDiel vpmg.c:4335-4344
// with NoSmooth :4346-4498 or Smooth :4500-4628, fillcoCoefMolIon :4237-4333):
// the unsmoothed dielectric value at a shifted point is epsp inside an atom
// sphere (atom flags, 3 maps) and not within a probe radius of an SAS point
// (sas flags, 3 maps; null when the probe radius is 0), else epsw; smooth
// selects the 9-point harmonic mean of those values.  kappa is 0 where the
// kappa flag is set (null: nowhere), else ionmask.  bits: nx*ny*nz bytes of
// scratch (the solute bits of every point).
void mol_maps(hipStream_t s, const Geo& g, const unsigned char* atom_flags,
              const unsigned char* sas_flags, const unsigned char* kappa_flags, double epsp,
              double epsw, double ionmask, int smooth, double* epsx, double* epsy, double* epsz,
              double* kappa, unsigned char* bits);

// ------------------------------------------------------------- boundaries
// bcflnew (vpmg.c:3289-3367): Debye-Hueckel sum over all atoms on the faces.
// The single Debye-Hueckel sphere of bcfl sdh: solute centre and radius, the
// molecule's charge, dipole and traceless quadrupole (already divided by 3),
// pre = e_c^2 / (4 pi eps0 kB T) * 1e10.
struct SdhMoments {
  double pos[3], size, q, u[3], qxx, qxy, qxz, qyy, qyz, qzz;
  double pre, xkappa, eps_p, eps_w;
};
void bc_sdh(hipStream_t s, const Geo& g, const double* xf, const double* yf, const double* zf,
            const SdhMoments& M, double* gx, double* gy, double* gz);
// Doubles of scratch bc_mdh needs for natoms atoms.
size_t bc_mdh_scratch(const Geo& g, int natoms);
// fp32_terms: each atom-point term in FP32 (hardware rsqrt and 2^x), summed
// in FP64; otherwise FP64 terms within a few ulp of the reference's.
void bc_mdh(hipStream_t s, const Geo& g, const double* xf, const double* yf, const double* zf,
            const AtomsDev& A, double pre1, double xkappa, double* part, double* gx, double* gy,
            double* gz, bool fp32_terms = false);
// --------------------------------------------------------------- maps
// An external map on the device or the host (the fields of a Vgrid).
struct MapDev {
  int nx, ny, nz;
  double hx, hy, hz, xmin, ymin, zmin, xmax, ymax, zmax;
  const double* data;
};
// Vgrid_value (vgrid.c:179-260): trilinear value at (x, y, z), false off the
// map.  Coordinates within 1e-4 of a face (Vcompare = 10^-(VGRID_DIGITS-2))
// snap their index to it.  A point whose lower index falls below the map is
// off it here; the reference reads outside its array there.
__host__ __device__ inline bool map_value(const MapDev& m, double x, double y, double z,
                                          double* v) {
  const double kVcompare = 1.0e-4;
  const double ifloat = (x - m.xmin) / m.hx, jfloat = (y - m.ymin) / m.hy,
               kfloat = (z - m.zmin) / m.hz;
  long ihi = (long)ceil(ifloat), jhi = (long)ceil(jfloat), khi = (long)ceil(kfloat);
  long ilo = (long)floor(ifloat), jlo = (long)floor(jfloat), klo = (long)floor(kfloat);
  if (fabs(x - m.xmin) < kVcompare) ilo = 0;
  if (fabs(y - m.ymin) < kVcompare) jlo = 0;
  if (fabs(z - m.zmin) < kVcompare) klo = 0;
  if (fabs(x - m.xmax) < kVcompare) ihi = m.nx - 1;
  if (fabs(y - m.ymax) < kVcompare) jhi = m.ny - 1;
  if (fabs(z - m.zmax) < kVcompare) khi = m.nz - 1;
  if (!(ilo >= 0 && jlo >= 0 && klo >= 0 && ihi >= 0 && jhi >= 0 && khi >= 0 && ilo < m.nx &&
        jlo < m.ny && klo < m.nz && ihi < m.nx && jhi < m.ny && khi < m.nz))
    return false;
  const double dx = ifloat - (double)(ilo), dy = jfloat - (double)(jlo),
               dz = kfloat - (double)(klo);
  const size_t sx = 1, sy = (size_t)m.nx, sz = (size_t)m.nx * m.ny;
  const double* d = m.data;
  *v = dx * dy * dz * (d[ihi * sx + jhi * sy + khi * sz]) +
       dx * (1.0 - dy) * dz * (d[ihi * sx + jlo * sy + khi * sz]) +
       dx * dy * (1.0 - dz) * (d[ihi * sx + jhi * sy + klo * sz]) +
       dx * (1.0 - dy) * (1.0 - dz) * (d[ihi * sx + jlo * sy + klo * sz]) +
       (1.0 - dx) * dy * dz * (d[ilo * sx + jhi * sy + khi * sz]) +
       (1.0 - dx) * (1.0 - dy) * dz * (d[ilo * sx + jlo * sy + khi * sz]) +
       (1.0 - dx) * dy * (1.0 - dz) * (d[ilo * sx + jhi * sy + klo * sz]) +
       (1.0 - dx) * (1.0 - dy) * (1.0 - dz) * (d[ilo * sx + jlo * sy + klo * sz]);
  return true;
}
// Samples of a map (data on the device) at every grid point (xf + ox,
// yf + oy, zf + oz):
//   mode 0: the value (fillcoCoefMap's dielectric maps, vpmg.c:4180-4210)
//   mode 1: the raw value, with one maximum per workgroup into part (the
//           kappa map before its scaling, vpmg.c:4124-4160); bad[1] counts
//           negative values.  part holds map_sample_parts(g) doubles.
//   mode 2: value * scale (fillcoChargeMap, vpmg.c:4938-4985, scale = zmagic)
// bad[0] counts points off the map; out == nullptr only counts.
size_t map_sample_parts(const Geo& g);
void map_sample(hipStream_t s, const Geo& g, const MapDev& m, const double* xf, const double* yf,
                const double* zf, double ox, double oy, double oz, int mode, double scale,
                double* out, double* part, int* bad);
// kappa = (raw < VPMGSMALL ? 0 : raw) / kappamax (vpmg.c:4163-4178)
void kappa_scale(hipStream_t s, size_t n, double kappamax, double* kappa);

// focusFillBound (vpmg.c:1914-2414): trilinear values of the old solution
// (or of the potential map for bcfl map, vpmg.c:1993-2006).
struct OldGrid {
  int nx, ny, nz;
  double hx, hy, hz, xmin, ymin, zmin, xmax, ymax, zmax;
  const double* u;
};
void focus_bound(hipStream_t s, const Geo& g, double xminN, double yminN, double zminN,
                 double xmaxN, double ymaxN, double zmaxN, const OldGrid& o, double* gx,
                 double* gy, double* gz, int* bad);

// --------------------------------------------------------------- partition
// Vpmg_setPart grid part (vpmg.c:720-870) with all bflags 0; invert = 1
// applies extEnergy's flip (vpmg.c:2491-2509).
void set_part(hipStream_t s, const Geo& g, double xcen0, double ycen0, double zcen0,
              const double lo[3], const double hi[3], int invert, double* pvec);

// ---------------------------------------------------------------- energies
// Each writes the raw sum to out (device); callers apply h^3/zmagic.
void energy_qf_volume(hipStream_t s, size_t n, const double* pvec, const double* u,
                      const double* charge, double* part, double* out);
void energy_diel(hipStream_t s, const Geo& g, const double* pvec, const double* u,
                 const double* ex, const double* ey, const double* ez, double* part,
                 double* out);
// nonlinear: sum pvec kappa zks2 c_j (exp(-q_j u) - 1); linear: sum pvec zkappa2 kappa u^2
struct IonBulk {
  int nion;
  double q[10], c[10];
};
void energy_qm(hipStream_t s, size_t n, int nonlin, const double* pvec, const double* kappa,
               const double* u, double zks2, double zkappa2, const IonBulk& ions, double* part,
               double* out, int* nchop);
// spl0 point energy: per atom trilinear u at the atom times charge*partID
void energy_qf_point(hipStream_t s, const Geo& g, const AtomsDev& A, const double* u,
                     double* part, double* out);

// The library's one statically allocated device buffer, for entry points that
// must not allocate (Vpmg_energyNoAlloc).  Layout in doubles: [0,
// kScratchParts) the energy reductions' partial sums (one per block),
// [kScratchParts, kScratchParts + 8) finished sums, the int at double
// kScratchParts + 8 the chop counter, from kScratchAtomOff one partial per
// atom (at most kScratchAtoms atoms).  Every caller on a device shares that
// device's copy, so calls that overlap on one device race.
constexpr size_t kScratchParts = 1024;
constexpr size_t kScratchAtomOff = kScratchParts + 64;
constexpr size_t kScratchAtoms = (size_t)1 << 17;
constexpr size_t kScratchDoubles = kScratchAtomOff + kScratchAtoms;
// The device address of the current device's copy (looked up with
// hipGetSymbolAddress on the first call per device, then cached); nullptr
// when the lookup fails.
double* scratch_address();

// ------------------------------------------------------------------ forces
// Per-atom force components (kT/A), out[3*a..3*a+2]; rc[a] = 1/0 like the
// reference's return codes.
void force_qf(hipStream_t s, const Geo& g, const AtomsDev& A, int mode, const double* u,
              double* out);
// order / coef: the surface method's window (2 = spl2; 3, 4 = spl3, spl4 with
// the per-atom coefficients of spline_coefficients).
void force_ib(hipStream_t s, const Geo& g, const AtomsDev& A, double win, double irad,
              double zkappa2, double izmagic, int nonlin, double ionstr, const IonBulk& ions,
              const double* kappa, const double* u, double* out, int order = 2,
              const double* coef = nullptr);
void force_db(hipStream_t s, const Geo& g, const AtomsDev& A, double win, double srad,
              double epsp, double epsw, double izmagic, const double* epsx, const double* epsy,
              const double* epsz, const double* u, double* out, int* rc, int order = 2,
              const double* coef = nullptr);

}  // namespace pmg
