// vpmg.cpp -- host side of libapbs_pmg.so: the Vpmg C API on one MI355X.
//
// Solve path (src/pmgc/mgdrvd.c, newdrvd.c, mgcsd.c, newtond.c):
//   fine operator VbuildA_fv -> Galerkin coarse operators -> V-cycles with
//   Hackbusch-Reusken damping, or damped inexact Newton with V-cycle inner
//   solves.  All grids live in HBM; the host only reads back stopping norms.
// Deviations from the reference schedule (the fixed point is the same):
//   * coarse 27-point levels use 8-colour Gauss-Seidel;
//   * coarsening continues below the reference's nlev while the grid still
//     halves, and the coarsest level is solved exactly with a dense inverse
//     of the full operator (the reference uses banded LINPACK on oC alone
//     for lpbe and 100 CG iterations for npbe).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <utility>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <limits>
#include <vector>

#include "../../include/apbs_pmg.h"
#include "fill_device.h"
#include "pmg_device.h"
#include "pool.h"
#include "vpmg_internal.h"

#define HIPCHECK(x)                                                                      \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) {                                                              \
      fprintf(stderr, "libapbs_pmg: HIP error %s at %s:%d\n", hipGetErrorString(e_),     \
              __FILE__, __LINE__);                                                       \
      abort();                                                                           \
    }                                                                                    \
  } while (0)

namespace {

// constants of src/generic/vhal.h:377-458, src/generic/vunit.h:68-116 and the
// MALOC values VSMALL / VPI
constexpr double VSMALL = 1.0e-9;
constexpr double VPMGSMALL = 1e-12;
constexpr double VPI = 3.14159265358979323846;
constexpr double Vunit_ec = 1.6021773e-19;
constexpr double Vunit_kb = 1.3806581e-23;
constexpr double Vunit_eps0 = 8.8541878e-12;
constexpr double VREDFRAC = 0.25;
constexpr int VMGNLEV = 4;

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// Device memory comes from the caching pool; every dfree follows a
// synchronisation of the stream that last used the block.
template <class T>
T* dalloc(size_t n) {
  if (n == 0) n = 1;
  return (T*)pmgpool::dev_alloc(n * sizeof(T));
}

template <class T>
void dfree(T*& p) {
  if (p) pmgpool::dev_free((void*)p);
  p = nullptr;
}

int coarsen(int n) { return (n - 1) / 2 + 1; }

}  // namespace

// =================================================================== Valist

extern "C" Valist* Valist_ctor(void) {
  Valist* v = (Valist*)calloc(1, sizeof(Valist));
  return v;
}

extern "C" void Valist_dtor(Valist** thee) {
  if (thee && *thee) {
    free((*thee)->atoms);
    free(*thee);
    *thee = nullptr;
  }
}

extern "C" int Valist_getStatistics(Valist* thee) {
  if (!thee || thee->number <= 0) return 0;
  Vatom* a0 = &thee->atoms[0];
  for (int d = 0; d < 3; d++) thee->mincrd[d] = thee->maxcrd[d] = a0->position[d];
  thee->maxrad = a0->radius;
  thee->charge = 0.0;
  for (int i = 0; i < thee->number; i++) {
    Vatom* a = &thee->atoms[i];
    for (int d = 0; d < 3; d++) {
      if (a->position[d] < thee->mincrd[d]) thee->mincrd[d] = a->position[d];
      if (a->position[d] > thee->maxcrd[d]) thee->maxcrd[d] = a->position[d];
    }
    if (a->radius > thee->maxrad) thee->maxrad = a->radius;
    thee->charge = thee->charge + a->charge;
  }
  for (int d = 0; d < 3; d++) thee->center[d] = 0.5 * (thee->maxcrd[d] + thee->mincrd[d]);
  return 1;
}

extern "C" int Valist_setAtoms(Valist* thee, int n, const double* xyz, const double* radius,
                               const double* charge) {
  if (!thee || n <= 0) return 0;
  free(thee->atoms);
  thee->atoms = (Vatom*)calloc((size_t)n, sizeof(Vatom));
  thee->number = n;
  for (int i = 0; i < n; i++) {
    Vatom* a = &thee->atoms[i];
    for (int d = 0; d < 3; d++) a->position[d] = xyz[3 * i + d];
    a->radius = radius[i];
    a->charge = charge[i];
    a->partID = 1.0;
    a->id = i;
  }
  return Valist_getStatistics(thee);
}

extern "C" int Valist_readPQRFile(Valist* thee, const char* path) {
  FILE* fh = fopen(path, "r");
  if (!fh) return 0;
  std::vector<double> xyz, r, q;
  char line[1024];
  while (fgets(line, sizeof line, fh)) {
    if (strncmp(line, "ATOM", 4) && strncmp(line, "HETATM", 6)) continue;
    std::vector<std::string> tok;
    char* save = nullptr;
    for (char* t = strtok_r(line, " \t\r\n", &save); t; t = strtok_r(nullptr, " \t\r\n", &save))
      tok.emplace_back(t);
    if (tok.size() < 6) continue;
    size_t m = tok.size();
    xyz.push_back(atof(tok[m - 5].c_str()));
    xyz.push_back(atof(tok[m - 4].c_str()));
    xyz.push_back(atof(tok[m - 3].c_str()));
    q.push_back(atof(tok[m - 2].c_str()));
    r.push_back(atof(tok[m - 1].c_str()));
  }
  fclose(fh);
  if (q.empty()) return 0;
  return Valist_setAtoms(thee, (int)q.size(), xyz.data(), r.data(), q.data());
}

extern "C" int Valist_getNumberAtoms(Valist* thee) { return thee ? thee->number : 0; }
extern "C" Vatom* Valist_getAtom(Valist* thee, int i) {
  return (thee && i >= 0 && i < thee->number) ? &thee->atoms[i] : nullptr;
}

// ===================================================================== Vpbe

namespace {
std::vector<double> ref_sphere(int npts);
}

// Vpbe_ctor2 (vpbe.c:264-465)
extern "C" Vpbe* Vpbe_ctor(Valist* alist, int ionNum, double* ionConc, double* ionRadii,
                           double* ionQ, double T, double soluteDiel, double solventDiel,
                           double solventRadius, int /*focusFlag*/, double sdens,
                           double /*z_mem*/, double /*L*/, double /*membraneDiel*/,
                           double /*V*/) {
  if (!alist || alist->number <= 0) return nullptr;
  const double N_A = 6.022045000e+23, e_c = 4.803242384e-10, k_B = 1.380662000e-16;
  const double pi = 4. * atan(1.);
  Vpbe* t = (Vpbe*)calloc(1, sizeof(Vpbe));
  t->alist = alist;
  for (int d = 0; d < 3; d++) t->soluteCenter[d] = alist->center[d];
  double radius = 0;
  Vatom* a0 = &alist->atoms[0];
  double xmin = a0->position[0], xmax = xmin, ymin = a0->position[1], ymax = ymin,
         zmin = a0->position[2], zmax = zmin, charge = 0;
  for (int i = 0; i < alist->number; i++) {
    Vatom* a = &alist->atoms[i];
    double r = a->radius, x = a->position[0], y = a->position[1], z = a->position[2];
    if ((x + r) > xmax) xmax = x + r;
    if ((x - r) < xmin) xmin = x - r;
    if ((y + r) > ymax) ymax = y + r;
    if ((y - r) < ymin) ymin = y - r;
    if ((z + r) > zmax) zmax = z + r;
    if ((z - r) < zmin) zmin = z - r;
    double dx = x - t->soluteCenter[0], dy = y - t->soluteCenter[1], dz = z - t->soluteCenter[2];
    double dist = sqrt(dx * dx + dy * dy + dz * dz) + r;
    if (dist > radius) radius = dist;
    charge += a->charge;
  }
  t->soluteRadius = radius;
  t->soluteXlen = xmax - xmin;
  t->soluteYlen = ymax - ymin;
  t->soluteZlen = zmax - zmin;
  t->soluteCharge = charge;
  if (ionNum >= APBS_PMG_MAXION) {
    free(t);
    return nullptr;
  }
  t->numIon = ionNum;
  double netCharge = 0.0;
  for (int i = 0; i < ionNum; i++) {
    t->ionConc[i] = ionConc[i];
    t->ionRadii[i] = ionRadii[i];
    if (ionRadii[i] > t->maxIonRadius) t->maxIonRadius = ionRadii[i];
    t->ionQ[i] = ionQ[i];
    t->bulkIonicStrength += (0.5 * ionConc[i] * ionQ[i] * ionQ[i]);
    netCharge += (ionConc[i] * ionQ[i]);
  }
  if (fabs(netCharge) > VSMALL) {  // counterion imbalance (vpbe.c:362-373)
    fprintf(stderr, "Vpbe_ctor: counterion charge imbalance (%g M)\n", netCharge);
    free(t);
    return nullptr;
  }
  if (T == 0.0) {
    free(t);
    return nullptr;
  }
  t->T = T;
  t->soluteDiel = soluteDiel;
  t->solventDiel = solventDiel;
  t->solventRadius = solventRadius;
  if (t->bulkIonicStrength == 0.) {
    t->xkappa = t->deblen = t->zkappa2 = 0.;
  } else {
    t->xkappa = sqrt(t->bulkIonicStrength * 1.0e-16 *
                     ((8.0 * pi * N_A * e_c * e_c) / (1000.0 * t->solventDiel * k_B * T)));
    t->deblen = 1. / t->xkappa;
    t->zkappa2 = t->solventDiel * t->xkappa * t->xkappa;
  }
  t->zmagic = ((4.0 * pi * e_c * e_c) / (k_B * t->T)) * 1.0e+8;
  t->ipkey = IPKEY_LPBE;
  t->paramFlag = 1;
  // cell-list radius and hash dimensions (vpbe.c:415-434, MAX_SPLINE_WINDOW
  // 0.5, MAX_HASH_DIM 75); the cell list itself is built on the device
  Vclist* cl = (Vclist*)calloc(1, sizeof(Vclist));
  cl->alist = alist;
  cl->max_radius =
      (t->maxIonRadius > t->solventRadius ? t->maxIonRadius : t->solventRadius) + 0.5;
  const double len[3] = {t->soluteXlen, t->soluteYlen, t->soluteZlen};
  cl->n = 1;
  for (int d = 0; d < 3; d++) {
    int nh = (int)(len[d] / 0.5);
    cl->npts[d] = nh < 3 ? 3 : (nh > 75 ? 75 : nh);
    cl->n *= cl->npts[d];
  }
  t->clist = cl;
  // Vacc_storeParms (vacc.c:146-189): reference sphere of
  // ceil(4 pi (max atom radius + clist radius)^2 sdens) points
  Vacc* acc = (Vacc*)calloc(1, sizeof(Vacc));
  acc->alist = alist;
  acc->clist = cl;
  double maxrad = 0.0;
  for (int i = 0; i < alist->number; i++)
    if (alist->atoms[i].radius > maxrad) maxrad = alist->atoms[i].radius;
  maxrad = maxrad + cl->max_radius;
  const int nsphere = (int)ceil(4.0 * VPI * maxrad * maxrad * sdens);
  std::vector<double> ref = ref_sphere(nsphere);
  VaccSurf* rs = (VaccSurf*)calloc(1, sizeof(VaccSurf));
  rs->npts = (int)(ref.size() / 3);
  rs->xpts = (double*)calloc(rs->npts + 1, sizeof(double));
  rs->ypts = (double*)calloc(rs->npts + 1, sizeof(double));
  rs->zpts = (double*)calloc(rs->npts + 1, sizeof(double));
  rs->bpts = (char*)calloc(rs->npts + 1, 1);
  for (int q = 0; q < rs->npts; q++) {
    rs->xpts[q] = ref[3 * q];
    rs->ypts[q] = ref[3 * q + 1];
    rs->zpts[q] = ref[3 * q + 2];
    rs->bpts[q] = 1;
  }
  rs->area = 4.0 * VPI;
  acc->refSphere = rs;
  t->acc = acc;
  return t;
}

extern "C" void Vpbe_dtor(Vpbe** thee) {
  if (thee && *thee) {
    Vpbe* t = *thee;
    if (t->acc) {
      if (VaccSurf* rs = t->acc->refSphere) {
        free(rs->xpts);
        free(rs->ypts);
        free(rs->zpts);
        free(rs->bpts);
        free(rs);
      }
      free(t->acc);
    }
    free(t->clist);
    free(t);
    *thee = nullptr;
  }
}
extern "C" double Vpbe_getBulkIonicStrength(Vpbe* t) { return t->bulkIonicStrength; }
extern "C" double Vpbe_getMaxIonRadius(Vpbe* t) { return t->maxIonRadius; }
extern "C" double Vpbe_getTemperature(Vpbe* t) { return t->T; }
extern "C" double Vpbe_getSoluteDiel(Vpbe* t) { return t->soluteDiel; }
extern "C" double Vpbe_getSolventDiel(Vpbe* t) { return t->solventDiel; }
extern "C" double Vpbe_getSolventRadius(Vpbe* t) { return t->solventRadius; }
extern "C" double Vpbe_getXkappa(Vpbe* t) { return t->xkappa; }
extern "C" double Vpbe_getZkappa2(Vpbe* t) { return t->zkappa2; }
extern "C" double Vpbe_getZmagic(Vpbe* t) { return t->zmagic; }
extern "C" double Vpbe_getSoluteRadius(Vpbe* t) { return t->soluteRadius; }
extern "C" double Vpbe_getSoluteCharge(Vpbe* t) { return t->soluteCharge; }
extern "C" double* Vpbe_getSoluteCenter(Vpbe* t) { return t->soluteCenter; }
extern "C" Valist* Vpbe_getValist(Vpbe* t) { return t->alist; }
extern "C" int Vpbe_getIons(Vpbe* t, int* nion, double ionConc[APBS_PMG_MAXION],
                            double ionRadii[APBS_PMG_MAXION], double ionQ[APBS_PMG_MAXION]) {
  *nion = t->numIon;
  for (int i = 0; i < t->numIon; i++) {
    ionConc[i] = t->ionConc[i];
    ionRadii[i] = t->ionRadii[i];
    ionQ[i] = t->ionQ[i];
  }
  return *nion;
}

// =================================================================== MGparm

extern "C" MGparm* MGparm_ctor(MGparm_CalcType type) {
  MGparm* m = (MGparm*)calloc(1, sizeof(MGparm));
  m->type = type;
  m->nlev = VMGNLEV;
  m->chgs = VCM_CHARGE;
  m->parsed = 1;
  return m;
}

extern "C" void MGparm_dtor(MGparm** thee) {
  if (thee && *thee) {
    free(*thee);
    *thee = nullptr;
  }
}

// nlev / dime fix-up of MGparm_check (mgparm.c:265-320)
extern "C" int MGparm_check(MGparm* thee) {
  if (!thee || !thee->parsed || !thee->setdime || !thee->setchgm) return 0;
  if (thee->type == MCT_MANUAL || thee->type == MCT_DUMMY) {
    if (!thee->setgrid && !thee->setglen) return 0;
    if (thee->setgrid && thee->setglen) return 0;
    if (!thee->setgcent) return 0;
  }
  if (thee->type == MCT_AUTO || thee->type == MCT_PARALLEL) {
    if (!thee->setcglen || !thee->setfglen || !thee->setcgcent || !thee->setfgcent) return 0;
  }
  if (thee->type == MCT_PARALLEL && (!thee->setpdime || !thee->setofrac)) return 0;
  int tdime[3], tnlev[3];
  if (thee->type != MCT_DUMMY) {
    for (int i = 0; i < 3; i++) {
      int ti = thee->dime[i] - 1;
      if (ti == (int)pow(2, thee->nlev + 1)) {
        tnlev[i] = thee->nlev;
        tdime[i] = thee->dime[i];
      } else {
        tdime[i] = thee->dime[i];
        ti = tdime[i] - 1;
        tnlev[i] = 0;
        while (ti % 2 == 0 && ti > 0) {
          tnlev[i]++;
          ti = (int)ceil(0.5 * ti);
        }
        tnlev[i]--;
        if ((tdime[i] > 65) && (tnlev[i] < VMGNLEV)) {
          ti = (int)(tdime[i] / pow(2., (VMGNLEV + 1)));
          if (ti < 1) ti = 1;
          tdime[i] = ti * (int)(pow(2., (VMGNLEV + 1))) + 1;
          tnlev[i] = 4;
        }
      }
    }
  } else {
    for (int i = 0; i < 3; i++) {
      tnlev[i] = thee->nlev;
      tdime[i] = thee->dime[i];
    }
  }
  int nlev = std::min(std::min(tnlev[0], tnlev[1]), tnlev[2]);
  thee->nlev = nlev;
  if (nlev <= 0) return 0;
  for (int i = 0; i < 3; i++) thee->dime[i] = tdime[i];
  if (thee->type == MCT_MANUAL || thee->type == MCT_DUMMY) {
    // NOsh_setupCalcMGMANUAL (nosh.c:1787-1799)
    if (!thee->setgrid)
      for (int i = 0; i < 3; i++) thee->grid[i] = thee->glen[i] / ((double)(thee->dime[i] - 1));
    if (!thee->setglen)
      for (int i = 0; i < 3; i++) thee->glen[i] = thee->grid[i] * ((double)(thee->dime[i] - 1));
  }
  return 1;
}

// NOsh_setupCalcMGAUTO (nosh.c:1828-2101)
extern "C" int MGparm_setupAuto(MGparm* elec, MGparm** out, int maxcalc) {
  if (!elec || !out) return 0;
  // Every level of an mg-para chain stays MCT_PARALLEL and keeps the
  // processor's disjoint partition (nosh.c:2085-2092 resets it for the
  // others), so setPartMG and the focusing energy use the partition.
  const bool para = elec->type == MCT_PARALLEL;
  double cgrid[3], fgrid[3], redrat[3];
  int tn[3];
  for (int j = 0; j < 3; j++) {
    cgrid[j] = (elec->cglen[j]) / ((double)(elec->dime[j] - 1));
    fgrid[j] = (elec->fglen[j]) / ((double)(elec->dime[j] - 1));
  }
  for (int j = 0; j < 3; j++) {
    if (fgrid[j] / cgrid[j] < VREDFRAC) {
      double redfrac = fgrid[j] / cgrid[j];
      double td = log(redfrac) / log(VREDFRAC);
      tn[j] = (int)ceil(td) + 1;
    } else {
      tn[j] = 2;
    }
  }
  int nfocus = std::max(std::max(tn[0], tn[1]), tn[2]);
  if (nfocus > maxcalc) return 0;
  for (int j = 0; j < 3; j++) redrat[j] = pow((fgrid[j] / cgrid[j]), 1.0 / ((double)nfocus - 1.0));
  for (int f = 0; f < nfocus; f++) {
    MGparm* c = out[f];
    MGparm* p = f > 0 ? out[f - 1] : nullptr;
    *c = *elec;
    c->type = para ? MCT_PARALLEL : MCT_MANUAL;
    for (int j = 0; j < 3; j++) {
      if (f == 0) {
        c->grid[j] = cgrid[j];
        c->glen[j] = elec->cglen[j];
      } else {
        c->grid[j] = redrat[j] * (p->grid[j]);
        c->glen[j] = redrat[j] * (p->glen[j]);
      }
    }
    c->setgrid = c->setglen = 1;
    if (f == 0) {
      c->cmeth = elec->ccmeth;
      for (int j = 0; j < 3; j++) c->center[j] = elec->ccenter[j];
    } else if (f == nfocus - 1) {
      c->cmeth = elec->fcmeth;
      for (int j = 0; j < 3; j++) c->center[j] = elec->fcenter[j];
    } else {
      c->cmeth = MCM_FOCUS;
      for (int j = 0; j < 3; j++) c->center[j] = elec->fcenter[j];
    }
    if (f != 0) {
      for (int j = 0; j < 3; j++) {
        int dofix = 0;
        double minf = c->center[j] - 0.5 * (c->glen[j]);
        double minc = p->center[j] - 0.5 * (p->glen[j]);
        double d = minc - minf;
        if (d >= VSMALL) {
          if (f == nfocus - 1) return 0;
          c->center[j] += d;
          dofix = 1;
        }
        double maxf = c->center[j] + 0.5 * (c->glen[j]);
        double maxc = p->center[j] + 0.5 * (p->glen[j]);
        d = maxf - maxc;
        if (d >= VSMALL) {
          if (f == nfocus - 1 || dofix) return 0;
          c->center[j] -= d;
        }
      }
    }
    if (!para) {
      for (int j = 0; j < 6; j++) c->partDisjOwnSide[j] = 0;
      for (int j = 0; j < 3; j++) {
        c->partDisjCenter[j] = 0;
        c->partDisjLength[j] = c->glen[j];
      }
    }
    c->setgcent = 1;
  }
  return nfocus;
}

// NOsh_setupCalcMGPARA (nosh.c:2103-2394) for the processor `rank` of the
// pdime processor grid (the reference's `async` rank): each axis of the fine
// domain fglen is cut into disjoint slabs of dime/(1 + 2 ofrac) points on a
// global grid of npx*disj points; the processor's fine grid keeps dime points
// and overlaps its neighbours (flush with the domain at the outer
// processors), its disjoint slab becomes partDisjCenter/Length, and
// partDisjOwnSide flags the sides that face another processor.  The result
// is the focusing chain of MGparm_setupAuto over that fine grid.
extern "C" int MGparm_setupPara(MGparm* elec, int rank, MGparm** out, int maxcalc) {
  if (!elec || !out || elec->type != MCT_PARALLEL || !elec->setpdime || !elec->setofrac)
    return 0;
  const int np[3] = {elec->pdime[0], elec->pdime[1], elec->pdime[2]};
  if (np[0] < 1 || np[1] < 1 || np[2] < 1) return 0;
  if (rank < 0 || rank >= np[0] * np[1] * np[2]) return 0;
  const int kp = rank / (np[0] * np[1]);
  const int jp = (rank - kp * np[0] * np[1]) / np[0];
  const int ip = rank - kp * np[0] * np[1] - jp * np[0];
  const int pc[3] = {ip, jp, kp};
  MGparm e = *elec;
  double olapLen[3], olapCen[3], disjLen[3], disjCen[3];
  for (int d = 0; d < 3; d++) {
    const double ofr = np[d] == 1 ? 0.0 : e.ofrac;
    const int disj = (int)floor(e.dime[d] / (1 + 2 * ofr) + 0.5);
    const int effGlob = np[d] * disj;
    const double h = e.fglen[d] / (double)(effGlob - 1);
    const int igmin = pc[d] * disj, igmax = igmin + disj - 1;
    const double mn = std::max(h * (igmin - 0.5), 0.0);
    const double mx = std::min(h * (igmax + 0.5), e.fglen[d]);
    disjLen[d] = mx - mn;
    const double cen = 0.5 * e.fglen[d];
    disjCen[d] = mn + 0.5 * disjLen[d] - cen;
    if (fabs(disjCen[d]) < VSMALL) disjCen[d] = 0.0;
    const int olapReg = np[d] != 1 ? (int)floor(ofr * e.fglen[d] / np[d] / h + 0.5) + 1 : 0;
    int igminO;
    if (pc[d] == 0)
      igminO = 0;
    else if (pc[d] == np[d] - 1)
      igminO = effGlob - e.dime[d];
    else
      igminO = igmin - olapReg;
    const int igmaxO = igminO + e.dime[d] - 1;
    const double mnO = h * igminO, mxO = h * igmaxO;
    olapLen[d] = mxO - mnO;
    olapCen[d] = (mnO + 0.5 * olapLen[d]) - cen;
    if (fabs(olapCen[d]) < VSMALL) olapCen[d] = 0.0;
  }
  // sides facing another processor (VAPBS_RIGHT 0, FRONT 1, UP 2, LEFT 3, BACK 4, DOWN 5)
  e.partDisjOwnSide[3] = ip == 0 ? 0 : 1;
  e.partDisjOwnSide[0] = ip == np[0] - 1 ? 0 : 1;
  e.partDisjOwnSide[4] = jp == 0 ? 0 : 1;
  e.partDisjOwnSide[1] = jp == np[1] - 1 ? 0 : 1;
  e.partDisjOwnSide[5] = kp == 0 ? 0 : 1;
  e.partDisjOwnSide[2] = kp == np[2] - 1 ? 0 : 1;
  for (int d = 0; d < 3; d++) {
    e.fglen[d] = olapLen[d];
    e.partDisjLength[d] = disjLen[d];
    e.partDisjCenter[d] = e.fcenter[d] + disjCen[d];
    e.fcenter[d] += olapCen[d];
  }
  return MGparm_setupAuto(&e, out, maxcalc);
}

// ==================================================================== Vpmgp

extern "C" void Vpmgp_makeCoarse(int numLevel, int nxOld, int nyOld, int nzOld, int* nxNew,
                                 int* nyNew, int* nzNew) {
  *nxNew = nxOld;
  *nyNew = nyOld;
  *nzNew = nzOld;
  for (int l = 0; l < numLevel; l++) {
    *nxNew = coarsen(*nxNew);
    *nyNew = coarsen(*nyNew);
    *nzNew = coarsen(*nzNew);
  }
}

// Vpmgp_size (vpmgp.c:196-291): the reference's workspace accounting, kept
// because callers read these fields.
extern "C" void Vpmgp_size(Vpmgp* t) {
  t->nf = t->nx * t->ny * t->nz;
  t->narr = t->nf;
  int nxf = t->nx, nyf = t->ny, nzf = t->nz;
  t->nxc = t->nx;
  t->nyc = t->ny;
  t->nzc = t->nz;
  for (int l = 2; l <= t->nlev; l++) {
    Vpmgp_makeCoarse(1, nxf, nyf, nzf, &t->nxc, &t->nyc, &t->nzc);
    nxf = t->nxc;
    nyf = t->nyc;
    nzf = t->nzc;
    t->narr += nxf * nyf * nzf;
  }
  t->nc = t->nxc * t->nyc * t->nzc;
  t->narrc = t->narr - t->nf;
  int n_band = 0;
  if (t->mgsolv == 1) {
    int num_band = 1 + (t->nxc - 2) * (t->nyc - 2) + (t->nxc - 2) + 1;
    int nc_band = (t->nxc - 2) * (t->nyc - 2) * (t->nzc - 2);
    n_band = nc_band * num_band;
  }
  t->n_rpc = 100 * (t->nlev + 1);
  t->nrwk = 2 * (size_t)t->narr + (size_t)4 * t->nf + (size_t)(27 + 14) * t->narrc + n_band + t->n_rpc;
  t->n_iz = 50 * (t->nlev + 1);
  t->n_ipc = 100 * (t->nlev + 1);
  t->niwk = t->n_iz + t->n_ipc;
}

// Vpmgp_ctor2 (vpmgp.c:95-170)
extern "C" Vpmgp* Vpmgp_ctor(MGparm* mgparm) {
  if (!mgparm) return nullptr;
  Vpmgp* t = (Vpmgp*)calloc(1, sizeof(Vpmgp));
  t->nx = mgparm->dime[0];
  t->ny = mgparm->dime[1];
  t->nz = mgparm->dime[2];
  t->hx = mgparm->grid[0];
  t->hy = mgparm->grid[1];
  t->hzed = mgparm->grid[2];
  t->xlen = ((double)(mgparm->dime[0] - 1)) * mgparm->grid[0];
  t->ylen = ((double)(mgparm->dime[1] - 1)) * mgparm->grid[1];
  t->zlen = ((double)(mgparm->dime[2] - 1)) * mgparm->grid[2];
  t->nlev = mgparm->nlev;
  t->nonlin = mgparm->nonlintype;
  t->meth = mgparm->method;
  t->ipkey = (t->nonlin == NONLIN_LPBE) ? IPKEY_LPBE
                                        : (t->nonlin == NONLIN_SMPBE ? IPKEY_SMPBE : IPKEY_NPBE);
  t->errtol = mgparm->setetol ? mgparm->etol : 1.0e-6;
  t->itmax = 200;
  t->istop = 1;
  t->iinfo = 1;
  t->bcfl = BCFL_SDH;
  t->mgcoar = 2;
  t->nu1 = 2;
  t->nu2 = 2;
  t->omegal = 19.4e-1;
  t->omegan = 9.0e-1;
  t->ipcon = 3;
  t->irite = 8;
  t->mgsmoo = 1;
  t->mgsolv = (t->nonlin == NONLIN_NPBE || t->nonlin == NONLIN_SMPBE) ? 0 : 1;
  for (int d = 0; d < 3; d++) (&t->xcent)[d] = mgparm->center[d];
  return t;
}

extern "C" void Vpmgp_dtor(Vpmgp** thee) {
  if (thee && *thee) {
    free(*thee);
    *thee = nullptr;
  }
}

// ============================================================ device state

namespace {

struct DLevel {
  int nx = 0, ny = 0, nz = 0, st = 7;
  size_t n = 0;
  double* a[14] = {};
  double *cc = nullptr, *fc = nullptr, *D = nullptr, *x = nullptr, *w0 = nullptr;
  double *cprime = nullptr, *xtmp = nullptr;
  // FP32 copies for the mixed-precision correction cycles
  float* af[14] = {};
  float *Df = nullptr, *xfl = nullptr, *w0f = nullptr;
  // coarse-level ping-pong partners of the 27-point smoother (zero boundary)
  double* xs = nullptr;
  float* xsf = nullptr;

  pmg::Ops<double> ops() const {
    pmg::Ops<double> o;
    for (int s = 0; s < 14; s++) o.a[s] = a[s];
    return o;
  }
  pmg::Ops<float> opsf() const {
    pmg::Ops<float> o;
    for (int s = 0; s < 14; s++) o.a[s] = af[s];
    return o;
  }
};

struct DevCells {
  int* start = nullptr;
  int* list = nullptr;
  pmg::Cells c{};
  void release() {
    dfree(start);
    dfree(list);
  }
};

// Tile lists (pmg::Tiles) on the device, with the host arrays they were
// uploaded from (kept until release, after the kernels that read them).
struct DevTiles {
  int* start = nullptr;
  int* list = nullptr;
  pmg::Tiles t{};
  std::vector<int> hstart, hlist;
  void release() {
    dfree(start);
    dfree(list);
  }
};

struct PmgDev {
  hipStream_t s = nullptr;
  int nx = 0, ny = 0, nz = 0;
  size_t n = 0;
  // fine maps
  double *epsx = nullptr, *epsy = nullptr, *epsz = nullptr, *kappa = nullptr,
         *charge = nullptr, *u = nullptr, *pvec = nullptr, *ccf = nullptr;
  double *gx = nullptr, *gy = nullptr, *gz = nullptr, *xf = nullptr, *yf = nullptr,
         *zf = nullptr;
  // atoms
  int natoms = 0;
  double *apos = nullptr, *arad = nullptr, *achg = nullptr, *apart = nullptr;
  std::vector<double> hpart;  // host partID copy
  // hierarchy
  int nlev_ref = 0, ntot = 0;
  bool h_newton = false, h_mixed = false;
  DLevel L[24];
  bool built = false;
  double *w1 = nullptr, *F = nullptr, *wn1 = nullptr, *wn2 = nullptr, *wn3 = nullptr,
         *nlw = nullptr;
  float* w1f = nullptr;
  double* xs64 = nullptr;  // fine-level ping-pong partners of the smoother (zero boundary)
  float* xs32 = nullptr;
  double *part = nullptr, *scal = nullptr;
  double* hscal = nullptr;  // pinned
  double *dense = nullptr, *dwork = nullptr;
  int dn = 0;
  // forces cache
  double *fqf = nullptr, *fib = nullptr, *fdb = nullptr;
  int fchgm = -1, fsrfm = -1;      // the methods the cached forces were computed with
  double* splcoef = nullptr;       // srfm spl3 / spl4 window coefficients per atom
  int splorder = 0;
  double splwin = 0.0;
  int* frc = nullptr;
  double* atmp = nullptr;  // per-atom scratch
  bool forces_valid = false;
  std::vector<double> hqf, hib, hdb;
  std::vector<int> hrc;
  // options / stats
  int precision = 0;
  double mixed_tol = -1.0;
  // precision mode 2: the MDH boundary from FP32 terms (set before fillco)
  bool fast_boundary = false;
  // fine-level smoother: 0 fused out-of-place sweep, 1 one launch per colour
  // (environment APBS_PMG_SMOOTHER=colour selects 1)
  int smoother = 0;
  // coarse-level smoother: 0 one launch per plane parity or per colour pair
  // on the levels where each is faster (gsrb27_use_planes, gsrb27_use_pairs),
  // 1 one launch per colour everywhere (APBS_PMG_SMOOTHER27=colour selects 1)
  int smoother27 = 0;
  // fillco's MDH boundary, and the coarsest level's dense inverse, on second
  // streams (pmgint::set_side_stream)
  int side = 1;
  // the dense inverse's stream: set_diagonals forks it after the diagonals
  // are formed, the V-cycle joins it before the coarsest solve
  hipStream_t cs = nullptr;
  hipEvent_t cs_fork = nullptr, cs_join = nullptr;
  bool cs_pending = false;
  int mirror = 1;
  bool failed = false;  // the last solve stopped on a failure (solve_failed)
  bool poison = false;  // test hook: NaN-fill the interiors of a new hierarchy
  double stats[8] = {};
  pmg::Geo geo{};
  // kernel timing: HIP event pairs recorded on the object's stream around
  // fine-level smoother calls (kind 0) and whole V-cycles (kind 1)
  int timing = 0;
  std::vector<hipEvent_t> ev;
  struct Mark {
    int e, kind, launches;
  };
  std::vector<Mark> marks;
  int ev_next = 0;    // next free event of the pool
  // V-cycles captured as HIP graphs, keyed by precision, solution and
  // right-hand side, when the environment sets APBS_PMG_GRAPHS=1.  Eager
  // launches are the default: on MI355X the replays were no faster (161^3
  // npbe mixed: 61.2 single-request and 81.4 served solves/s replayed,
  // 61.8 and 84.4 eager), and each new hierarchy pays the capture.
  int graphs = 0;
  struct CycleGraph {
    int word;
    const void *f, *x;
    hipGraph_t g;
    hipGraphExec_t e;
  };
  std::vector<CycleGraph> cycle_graphs;
  // ms0, launches0, ms1, count1, word bytes, fine points, words/pt/launch, vector form
  double kt[8] = {};
};

PmgDev* DEV(Vpmg* p) { return (PmgDev*)p->dev; }

double read_scalar(PmgDev* d, const double* dptr) {
  HIPCHECK(hipMemcpyAsync(d->hscal, dptr, sizeof(double), hipMemcpyDeviceToHost, d->s));
  HIPCHECK(hipStreamSynchronize(d->s));
  return d->hscal[0];
}

// The norm a reduction launched with hout = d->hscal wrote to host-mapped
// memory (visible once the stream has drained).
double read_norm(PmgDev* d) {
  HIPCHECK(hipStreamSynchronize(d->s));
  return d->hscal[0];
}

// hscal slot the coarsest-level inverse flags a pivot in (0 = none): the
// 1-based index of a pivot that was not positive and finite.
constexpr int kPivotFlag = 8;

// Whether the solve has failed, checked after every norm read: a norm that is
// not finite (the iteration diverged or met inf/NaN; the reference's loop test
// (rsnrm/rsden) > errtol is false for NaN and would stop as if converged), or
// a pivot the dense inverse of the coarsest operator flagged (the operator is
// not positive definite: the reference's Vdpbfa test, buildBd.c:106-114).
// The inverse is joined into the object's stream before the coarsest solve,
// so a flag it raised is visible at the next norm read.  Prints once.
bool solve_failed(PmgDev* d, double nrm) {
  const double piv = *(volatile double*)(d->hscal + kPivotFlag);
  if (piv == 0.0 && std::isfinite(nrm)) return false;
  if (!d->failed) {
    if (piv != 0.0)
      fprintf(stderr,
              "Vpmg_solve: coarsest-level operator (%d unknowns) not positive definite: "
              "pivot %d is not positive\n", d->dn, (int)piv);
    if (!std::isfinite(nrm)) fprintf(stderr, "Vpmg_solve: residual norm is not finite\n");
  }
  d->failed = true;
  return true;
}

int tbegin(PmgDev* d) {
  if (!d->timing) return -1;
  const int e = d->ev_next;
  d->ev_next += 2;
  while ((int)d->ev.size() < e + 2) {
    hipEvent_t h;
    HIPCHECK(hipEventCreate(&h));
    d->ev.push_back(h);
  }
  HIPCHECK(hipEventRecord(d->ev[e], d->s));
  return e;
}

void tend(PmgDev* d, int e, int kind, int launches) {
  if (e < 0) return;
  HIPCHECK(hipEventRecord(d->ev[e + 1], d->s));
  d->marks.push_back({e, kind, launches});
}

// Accumulate the recorded intervals (the stream must be idle).
void tcollect(PmgDev* d) {
  for (const auto& m : d->marks) {
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, d->ev[m.e], d->ev[m.e + 1]));
    d->kt[2 * m.kind] += ms;
    d->kt[2 * m.kind + 1] += m.launches;
  }
  d->marks.clear();
  d->ev_next = 0;
}

// Cell list over points with ascending indices per cell (counting sort), for
// gathers that must visit points in index order.
DevCells build_cells(hipStream_t s, const double* hpos, int n, double reach) {
  DevCells dc;
  double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
  for (int a = 0; a < n; a++)
    for (int d = 0; d < 3; d++) {
      mn[d] = std::min(mn[d], hpos[3 * a + d]);
      mx[d] = std::max(mx[d], hpos[3 * a + d]);
    }
  double cs = reach * 1.001 + 1e-6;
  if (cs < 0.25) cs = 0.25;
  // keep the table bounded for very sparse/large sets
  for (;;) {
    long cells = 1;
    for (int d = 0; d < 3; d++) cells *= (long)((mx[d] - (mn[d] - cs)) / cs) + 3;
    if (cells <= 8l * 1024 * 1024) break;
    cs *= 1.5;
  }
  dc.c.cs = cs;
  for (int d = 0; d < 3; d++) {
    dc.c.lo[d] = mn[d] - cs;
    dc.c.dim[d] = (int)((mx[d] - dc.c.lo[d]) / cs) + 2;
  }
  const long nc = (long)dc.c.dim[0] * dc.c.dim[1] * dc.c.dim[2];
  std::vector<int> start(nc + 1, 0), list(std::max(n, 1)), cid(std::max(n, 1));
  for (int a = 0; a < n; a++) {
    int ic[3];
    for (int d = 0; d < 3; d++) ic[d] = (int)floor((hpos[3 * a + d] - dc.c.lo[d]) / cs);
    cid[a] = ic[0] + dc.c.dim[0] * (ic[1] + dc.c.dim[1] * ic[2]);
    start[cid[a] + 1]++;
  }
  for (long i = 0; i < nc; i++) start[i + 1] += start[i];
  std::vector<int> fill(nc, 0);
  for (int a = 0; a < n; a++) list[start[cid[a]] + fill[cid[a]]++] = a;
  dc.start = dalloc<int>(nc + 1);
  dc.list = dalloc<int>(std::max(n, 1));
  HIPCHECK(hipMemcpyAsync(dc.start, start.data(), (nc + 1) * sizeof(int), hipMemcpyHostToDevice, s));
  HIPCHECK(hipMemcpyAsync(dc.list, list.data(), std::max(n, 1) * sizeof(int), hipMemcpyHostToDevice, s));
  HIPCHECK(hipStreamSynchronize(s));
  dc.c.start = dc.start;
  dc.c.list = dc.list;
  return dc;
}

// Tile lists for the order-sensitive gathers (charge and spline maps): atom a
// is listed, in ascending index, in every kTile^3 tile that meets the node box
// [floor(c - R) - 1, ceil(c + R) + 1] per axis, c its position in nodes and R
// its reach (A, reach[a] or reach[0] for all) in nodes.  The box contains the
// points the kernels' own tests can accept, so the lists are supersets.
DevTiles build_tiles(hipStream_t s, const pmg::Geo& g, const double* hpos, int n,
                     const std::vector<double>& reach) {
  DevTiles dt;
  const int nn[3] = {g.nx, g.ny, g.nz};
  const double h[3] = {g.hx, g.hy, g.hz}, lo[3] = {g.xmin, g.ymin, g.zmin};
  for (int d = 0; d < 3; d++) dt.t.tdim[d] = (nn[d] + pmg::kTile - 1) / pmg::kTile;
  const long nt = (long)dt.t.tdim[0] * dt.t.tdim[1] * dt.t.tdim[2];
  dt.hstart.assign(nt + 1, 0);
  std::vector<int> box(6 * (size_t)std::max(n, 1));
  for (int a = 0; a < n; a++) {
    const double r = reach.size() > 1 ? reach[a] : reach[0];
    int* b = &box[6 * (size_t)a];
    for (int d = 0; d < 3; d++) {
      const double c = (hpos[3 * a + d] - lo[d]) / h[d], R = r / h[d];
      const double blo = std::floor(c - R) - 1, bhi = std::ceil(c + R) + 1;
      const int l = (int)std::max(0.0, std::min(blo, (double)nn[d])),
                hgh = (int)std::max(-1.0, std::min(bhi, (double)(nn[d] - 1)));
      b[2 * d] = l / pmg::kTile;
      b[2 * d + 1] = l > hgh ? -1 : hgh / pmg::kTile;
    }
    for (int tz = b[4]; tz <= b[5]; tz++)
      for (int ty = b[2]; ty <= b[3]; ty++)
        for (int tx = b[0]; tx <= b[1]; tx++)
          dt.hstart[tx + (long)dt.t.tdim[0] * (ty + (long)dt.t.tdim[1] * tz) + 1]++;
  }
  for (long t = 0; t < nt; t++) dt.hstart[t + 1] += dt.hstart[t];
  dt.hlist.assign(std::max(dt.hstart[nt], 1), 0);
  std::vector<int> fill(dt.hstart.begin(), dt.hstart.end() - 1);
  for (int a = 0; a < n; a++) {
    const int* b = &box[6 * (size_t)a];
    for (int tz = b[4]; tz <= b[5]; tz++)
      for (int ty = b[2]; ty <= b[3]; ty++)
        for (int tx = b[0]; tx <= b[1]; tx++)
          dt.hlist[fill[tx + (long)dt.t.tdim[0] * (ty + (long)dt.t.tdim[1] * tz)]++] = a;
  }
  dt.start = dalloc<int>(nt + 1);
  dt.list = dalloc<int>(dt.hlist.size());
  HIPCHECK(hipMemcpyAsync(dt.start, dt.hstart.data(), (nt + 1) * sizeof(int),
                          hipMemcpyHostToDevice, s));
  HIPCHECK(hipMemcpyAsync(dt.list, dt.hlist.data(), dt.hlist.size() * sizeof(int),
                          hipMemcpyHostToDevice, s));
  dt.t.start = dt.start;
  dt.t.list = dt.list;
  return dt;
}

// VaccSurf_refSphere (vacc.c:924-978); computed with the host libm, as the
// reference does.
std::vector<double> ref_sphere(int npts) {
  std::vector<double> pts;
  double frac = ((double)npts) / 4.0;
  int ntheta = (int)floor(sqrt(VPI * frac) + 0.5);
  double dtheta = VPI / ((double)ntheta);
  int nphimax = 2 * ntheta;
  for (int it = 0; it < ntheta; it++) {
    double theta = dtheta * ((double)it);
    double sintheta = sin(theta), costheta = cos(theta);
    int nphi = (int)floor(sintheta * nphimax + 0.5);
    if (nphi != 0) {
      double dphi = 2 * VPI / ((double)nphi);
      for (int ip = 0; ip < nphi; ip++) {
        double phi = dphi * ((double)ip);
        pts.push_back(cos(phi) * sintheta);
        pts.push_back(sin(phi) * sintheta);
        pts.push_back(costheta);
      }
    }
  }
  return pts;
}

pmg::AtomsDev atoms_dev(PmgDev* d) {
  pmg::AtomsDev A;
  A.n = d->natoms;
  A.pos = d->apos;
  A.rad = d->arad;
  A.chg = d->achg;
  A.partID = d->apart;
  return A;
}

void upload_part(PmgDev* d) {
  HIPCHECK(hipMemcpyAsync(d->apart, d->hpart.data(), d->natoms * sizeof(double),
                          hipMemcpyHostToDevice, d->s));
}

// --------------------------------------------------------------- hierarchy

void free_graphs(PmgDev* d) {
  for (auto& cg : d->cycle_graphs) {
    HIPCHECK(hipGraphExecDestroy(cg.e));
    HIPCHECK(hipGraphDestroy(cg.g));
  }
  d->cycle_graphs.clear();
}

void free_hierarchy(PmgDev* d) {
  if (d->cs) HIPCHECK(hipStreamSynchronize(d->cs));
  d->cs_pending = false;
  free_graphs(d);
  for (int l = 0; l < d->ntot; l++) {
    DLevel& L = d->L[l];
    for (int s = 0; s < 14; s++) {
      dfree(L.a[s]);
      dfree(L.af[s]);
    }
    dfree(L.cc);
    dfree(L.fc);
    dfree(L.D);
    if (l > 0) dfree(L.x);
    L.x = nullptr;
    dfree(L.w0);
    dfree(L.cprime);
    dfree(L.xtmp);
    dfree(L.Df);
    dfree(L.xfl);
    dfree(L.w0f);
    dfree(L.xs);
    dfree(L.xsf);
    L.xs = nullptr;
    L.xsf = nullptr;
  }
  dfree(d->w1);
  dfree(d->w1f);
  dfree(d->xs64);
  dfree(d->xs32);
  dfree(d->F);
  dfree(d->wn1);
  dfree(d->wn2);
  dfree(d->wn3);
  dfree(d->nlw);
  dfree(d->dense);
  dfree(d->dwork);
  d->ntot = 0;
  d->built = false;
}

// Level sizes: the reference's nlev (Vbuildstr, mgsubd.c:261-342) extended
// while every axis still halves exactly and the coarsest interior exceeds 64.
int plan_levels(int nx, int ny, int nz, int nlev, int dims[][3]) {
  int cx = nx, cy = ny, cz = nz, nt = 0;
  dims[nt][0] = cx;
  dims[nt][1] = cy;
  dims[nt][2] = cz;
  nt++;
  for (int l = 1; l < nlev; l++) {
    cx = coarsen(cx);
    cy = coarsen(cy);
    cz = coarsen(cz);
    dims[nt][0] = cx;
    dims[nt][1] = cy;
    dims[nt][2] = cz;
    nt++;
  }
  while (((cx - 1) % 2 == 0) && ((cy - 1) % 2 == 0) && ((cz - 1) % 2 == 0) && coarsen(cx) >= 3 &&
         coarsen(cy) >= 3 && coarsen(cz) >= 3 && (cx - 2) * (cy - 2) * (cz - 2) > 64 && nt < 23) {
    cx = coarsen(cx);
    cy = coarsen(cy);
    cz = coarsen(cz);
    dims[nt][0] = cx;
    dims[nt][1] = cy;
    dims[nt][2] = cz;
    nt++;
  }
  return nt;
}

}  // namespace

size_t pmgint::footprint_bytes(int nx, int ny, int nz, int nlev, bool newton, bool mixed,
                               int natoms, int nref) {
  auto blk = [](double bytes) {  // pool rounding and reuse slack
    return bytes < (1 << 20) ? bytes + 512.0 : bytes * 1.25 + (1 << 20);
  };
  const double n = (double)nx * ny * nz;
  double b = 0;
  // Vpmg_ctor: 8 fine maps, faces, axes, reduction scratch, atoms, forces
  b += 8 * blk(8 * n);
  b += blk(16.0 * ny * nz) + blk(16.0 * nx * nz) + blk(16.0 * nx * ny);
  b += 3 * blk(8.0 * std::max(nx, std::max(ny, nz))) + blk(16.0 * pmg::kParts) + blk(256);
  b += 5 * blk(8.0 * std::max(natoms, 1) * 3) + 4 * blk(24.0 * std::max(natoms, 1));
  // Vpmg_fillco temporaries (released before the solve): the surface flag
  // maps (8 bytes per point), SAS flags and points; or the spline tiles
  const double sas = (double)natoms * nref;
  const double fill = std::max(blk(8 * n) + blk(sas) + blk(24 * sas), 3 * blk(8 * n)) + (64 << 20);
  // hierarchy
  int dims[24][3];
  const int nt = plan_levels(nx, ny, nz, nlev, dims);
  double h = 0;
  for (int l = 0; l < nt; l++) {
    const double m = (double)dims[l][0] * dims[l][1] * dims[l][2];
    const int na = l == 0 ? 4 : 14;
    h += (na + 4 + (l > 0 ? 2 : 0)) * blk(8 * m);     // ops, cc, fc, D, w0 (+x, xs)
    if (newton) h += 2 * blk(8 * m);                 // cprime, xtmp
    if (mixed) h += (na - 1 + 3 + (l > 0 ? 1 : 0)) * blk(4 * m);  // af, Df, xfl, w0f (+xsf)
  }
  h += 3 * blk(8 * n);                               // w1, F, xs64
  if (newton) h += 4 * blk(8 * n);                   // wn1..3, nlw
  if (mixed) h += 2 * blk(4 * n);                    // w1f, xs32
  const double dn = (double)(dims[nt - 1][0] - 2) * (dims[nt - 1][1] - 2) * (dims[nt - 1][2] - 2);
  h += blk(8 * dn * dn) + blk(8 * dn);
  return (size_t)(b + std::max(fill, h));
}

namespace {

void alloc_hierarchy(PmgDev* d, int nlev, bool newton, bool mixed) {
  // Every array's boundary layer starts at zero, like the reference's zeroed
  // workspace: the operator builders and every solver kernel write interior
  // entries only, and the stencils read coefficients and values of the
  // boundary layer.  The interiors are written before they are read, which is
  // also what lets a hierarchy of the same shape be reused by later solves,
  // so a new one clears the outer layer of each array only (a recycled pool
  // block's interior may hold anything).  With the poison test hook set, the
  // interiors are first filled with NaN.
  if (d->ntot > 0 && d->nlev_ref == nlev && d->h_newton == newton && d->h_mixed == mixed) return;
  free_hierarchy(d);
  struct Grid {
    void* p;
    int nx, ny, nz;
    size_t elem;
  };
  std::vector<Grid> grids;
  auto grid64 = [&](double* p, const DLevel& L) {
    grids.push_back({p, L.nx, L.ny, L.nz, sizeof(double)});
  };
  auto grid32 = [&](float* p, const DLevel& L) {
    grids.push_back({p, L.nx, L.ny, L.nz, sizeof(float)});
  };
  d->h_newton = newton;
  d->h_mixed = mixed;
  int dims[24][3];
  d->nlev_ref = nlev;
  d->ntot = plan_levels(d->nx, d->ny, d->nz, nlev, dims);
  for (int l = 0; l < d->ntot; l++) {
    DLevel& L = d->L[l];
    L.nx = dims[l][0];
    L.ny = dims[l][1];
    L.nz = dims[l][2];
    L.n = (size_t)L.nx * L.ny * L.nz;
    L.st = l == 0 ? 7 : 27;
    const int na = l == 0 ? 4 : 14;
    for (int s = 0; s < na; s++) {
      L.a[s] = dalloc<double>(L.n);
      grid64(L.a[s], L);
    }
    L.cc = dalloc<double>(L.n);
    L.fc = dalloc<double>(L.n);
    L.D = dalloc<double>(L.n);
    grid64(L.cc, L);
    grid64(L.fc, L);
    grid64(L.D, L);
    L.x = l == 0 ? d->u : dalloc<double>(L.n);
    if (l > 0) {
      L.xs = dalloc<double>(L.n);
      grid64(L.xs, L);
      grid64(L.x, L);
    }
    L.w0 = dalloc<double>(L.n);
    grid64(L.w0, L);
    if (newton) {
      L.cprime = dalloc<double>(L.n);
      grid64(L.cprime, L);
      L.xtmp = dalloc<double>(L.n);
      grid64(L.xtmp, L);
    }
    if (mixed) {
      for (int s = 1; s < na; s++) {  // oC is folded into Df
        L.af[s] = dalloc<float>(L.n);
        grid32(L.af[s], L);
      }
      L.Df = dalloc<float>(L.n);
      grid32(L.Df, L);
      L.xfl = dalloc<float>(L.n);
      L.w0f = dalloc<float>(L.n);
      grid32(L.xfl, L);
      grid32(L.w0f, L);
      if (l > 0) {
        L.xsf = dalloc<float>(L.n);
        grid32(L.xsf, L);
      }
    }
  }
  const DLevel& F = d->L[0];
  d->w1 = dalloc<double>(d->n);
  grid64(d->w1, F);
  d->xs64 = dalloc<double>(d->n);
  grid64(d->xs64, F);
  if (mixed) {
    d->w1f = dalloc<float>(d->n);
    grid32(d->w1f, F);
    d->xs32 = dalloc<float>(d->n);
    grid32(d->xs32, F);
  }
  d->F = dalloc<double>(d->n);
  grid64(d->F, F);
  if (newton) {
    d->wn1 = dalloc<double>(d->n);
    d->wn2 = dalloc<double>(d->n);
    d->wn3 = dalloc<double>(d->n);
    d->nlw = dalloc<double>(d->n);
    grid64(d->nlw, F);
    grid64(d->wn1, F);
    grid64(d->wn2, F);
    grid64(d->wn3, F);
  }
  if (d->poison)
    for (const Grid& g : grids) {
      const size_t n = (size_t)g.nx * g.ny * g.nz;
      if (g.elem == sizeof(double))
        pmg::fill<double>(d->s, (double*)g.p, n, std::numeric_limits<double>::quiet_NaN());
      else
        pmg::fill<float>(d->s, (float*)g.p, n, std::numeric_limits<float>::quiet_NaN());
    }
  {
    pmg::ZeroList z(d->s);
    for (const Grid& g : grids) z.add_shell(g.p, g.nx, g.ny, g.nz, g.elem);
  }
  const DLevel& C = d->L[d->ntot - 1];
  d->dn = (C.nx - 2) * (C.ny - 2) * (C.nz - 2);
  d->dense = dalloc<double>((size_t)d->dn * d->dn);
  d->dwork = dalloc<double>(d->dn);
}

// Vbuildops (mgsubd.c:61-258): fine FV operator, Galerkin coarse operators,
// restricted Helmholtz term and source.
void build_operators(PmgDev* d) {
  DLevel& F = d->L[0];
  pmg::buildA_fv(d->s, d->nx, d->ny, d->nz, d->xf, d->yf, d->zf, d->gx, d->gy, d->gz, d->epsx,
                 d->epsy, d->epsz, d->ccf, d->charge, F.a[0], F.cc, F.fc, F.a[1], F.a[2], F.a[3]);
  for (int l = 1; l < d->ntot; l++) {
    DLevel& P = d->L[l - 1];
    DLevel& C = d->L[l];
    pmg::galerkin(d->s, P.st, P.nx, P.ny, P.nz, P.ops(), C.nx, C.ny, C.nz, C.a);
    pmg::restrict3<double>(d->s, P.nx, P.ny, P.nz, C.nx, C.ny, C.nz, P.cc, C.cc);
    pmg::restrict3<double>(d->s, P.nx, P.ny, P.nz, C.nx, C.ny, C.nz, P.fc, C.fc);
  }
}

// D_l = oC_l + c_l for the given per-level Helmholtz arrays (and, for a
// mixed-precision solve, its FP32 image Df_l in the same pass), levels first
// and up.
void diag_levels(PmgDev* d, bool use_cprime, int first) {
  const bool f32 = d->precision == 1 && d->h_mixed;
  pmg::DiagLevels dl{};
  dl.nl = d->ntot - first;
  dl.off[0] = 0;
  for (int q = 0; q < dl.nl; q++) {
    const int l = first + q;
    DLevel& L = d->L[l];
    dl.oC[q] = L.a[0];
    dl.c[q] = use_cprime ? L.cprime : L.cc;
    dl.D[q] = L.D;
    dl.Df[q] = f32 ? L.Df : nullptr;
    dl.off[q + 1] = dl.off[q] + L.n;
  }
  if (dl.nl > 0) pmg::add_diag_levels(d->s, dl);
}

pmg::InvertArgs invert_args(PmgDev* d) {
  const DLevel& C = d->L[d->ntot - 1];
  return pmg::InvertArgs{C.ops(), C.D, d->dense, d->dwork, d->hscal + kPivotFlag};
}

// The diagonals (levels first and up; the Newton Jacobian forms the fine one
// with N'(u)), then the dense inverse of the coarsest level.
void set_diagonals(PmgDev* d, bool use_cprime, int first = 0) {
  diag_levels(d, use_cprime, first);
  const DLevel& C = d->L[d->ntot - 1];
  // The inverse is first read by the coarsest solve, after the downward
  // half of the next V-cycle; a one-workgroup kernel, it runs on its own
  // stream meanwhile.  Captured cycles keep it on the object's stream.
  hipStream_t s = d->s;
  if (d->side && !d->graphs) {
    if (!d->cs) {
      d->cs = pmgpool::stream_acquire();
      HIPCHECK(hipEventCreateWithFlags(&d->cs_fork, hipEventDisableTiming));
      HIPCHECK(hipEventCreateWithFlags(&d->cs_join, hipEventDisableTiming));
    }
    HIPCHECK(hipEventRecord(d->cs_fork, d->s));
    HIPCHECK(hipStreamWaitEvent(d->cs, d->cs_fork, 0));
    s = d->cs;
  }
  const pmg::InvertArgs ia = invert_args(d);
  pmg::dense_assemble_invert_batch(s, C.nx, C.ny, C.nz, d->dn, &ia, 1);
  if (s != d->s) {
    HIPCHECK(hipEventRecord(d->cs_join, s));
    d->cs_pending = true;
  }
}

// set_diagonals for n calculations of one shape on the first one's stream:
// the dense inverses of all n in one launch.
void set_diagonals_batch(PmgDev* const* ds, int n, bool use_cprime, int first = 0) {
  if (n == 1) {
    set_diagonals(ds[0], use_cprime, first);
    return;
  }
  pmg::InvertArgs ia[pmg::kMaxBatch];
  for (int b = 0; b < n; b++) {
    diag_levels(ds[b], use_cprime, first);
    ia[b] = invert_args(ds[b]);
  }
  const DLevel& C = ds[0]->L[ds[0]->ntot - 1];
  pmg::dense_assemble_invert_batch(ds[0]->s, C.nx, C.ny, C.nz, ds[0]->dn, ia, n);
}

// Join the dense inverse's stream before the first read of the inverse.
void join_dense(PmgDev* d) {
  if (!d->cs_pending) return;
  HIPCHECK(hipStreamWaitEvent(d->s, d->cs_join, 0));
  d->cs_pending = false;
}

// FP32 images of the level operators (the stencil arrays, which do not
// change within a solve); set_diagonals writes the FP32 diagonals, which
// change with every Newton Jacobian.
void sync_fp32_ops(PmgDev* d) {
  for (int l = 0; l < d->ntot; l++) {
    DLevel& L = d->L[l];
    pmg::ConvArgs c{};
    c.m = l == 0 ? 3 : 13;
    c.n = L.n;
    for (int s = 1; s <= c.m; s++) {
      c.in[s - 1] = L.a[s];
      c.out[s - 1] = L.af[s];
    }
    pmg::convert_many(d->s, c);
  }
}

// ------------------------------------------------------------------ V-cycle
// One V-cycle of Vmvcs (mgcsd.c:315-523) on per-level arrays.  X[l] are the
// per-level solutions (X[0] fine), f the fine right-hand side, w1 a fine-size
// scratch buffer (the residual of every level on the way down).
template <class T>
struct CycleArgs {
  T* X[24];
  const T* Dl[24];
  pmg::Ops<T> A[24];
  T* w0[24];
  T* xsl[24];  // coarse-level smoother partner buffers
  const T* f;
  T* w1;
  T* xs;  // fine-size smoother partner buffer
  // the fine solution starts at zero (its contents are not read and need not
  // be cleared)
  bool zero_x = false;
};

// The V-cycle of n calculations of one shape (n <= pmg::kMaxBatch), all on
// the stream of the first: every stage is one launch over the batch, so the
// coarse levels' small launches serve all n calculations.  Each calculation's
// arithmetic is its own single V-cycle's.
template <class T>
void vcycle_batch(PmgDev* const* ds, CycleArgs<T>* cs, int n) {
  PmgDev* d0 = ds[0];
  hipStream_t s = d0->s;
  const int nt = d0->ntot;
  DLevel* L = d0->L;  // level shapes (the same for every calculation)
  if (n == 1) {
    d0->kt[4] = sizeof(T);
    d0->kt[5] = (double)L[0].n;
    // algorithmic words per fine point per smoother launch: 7 for a full sweep
    d0->kt[6] = d0->smoother == 0 ? 7.0 : 3.5;
    d0->kt[7] = d0->smoother == 0 && pmg::fused_vector(L[0].nx, L[0].ny, L[0].nz, (int)sizeof(T));
  }
  pmg::LevelArgs<T> la[pmg::kMaxBatch];
  pmg::TransferArgs<T> ta[pmg::kMaxBatch];
  // the level-l arguments of every calculation (with_r: the residual goes to w1)
  auto level = [&](int l, bool with_r) {
    for (int b = 0; b < n; b++)
      la[b] = pmg::LevelArgs<T>{cs[b].A[l], cs[b].Dl[l], l == 0 ? cs[b].f : cs[b].w0[l],
                                cs[b].X[l], l == 0 ? cs[b].xs : cs[b].xsl[l],
                                with_r ? cs[b].w1 : nullptr};
  };
  // (the smoother timing covers full sweeps only: a zero-start pair's first
  // sweep reads no solution)
  auto fine_smooth = [&](int adj, bool zero_x) {
    const int e0 = n == 1 && !zero_x ? tbegin(d0) : -1;
    if (d0->smoother == 0) {
      level(0, false);
      pmg::gsrb7_fused_batch<T>(s, L[0].nx, L[0].ny, L[0].nz, la, n, 2, adj, 0, zero_x);
      if (n == 1) tend(d0, e0, 0, pmg::gsrb7_fused_launches(2));
    } else {
      for (int b = 0; b < n; b++) {
        if (zero_x) pmg::fill<T>(s, cs[b].X[0], L[0].n, T(0));
        pmg::gsrb7<T>(s, L[0].nx, L[0].ny, L[0].nz, cs[b].A[0], cs[b].Dl[0], cs[b].f, cs[b].X[0],
                      2, adj);
      }
      if (n == 1) tend(d0, e0, 0, pmg::gsrb7_launches(2));
    }
  };
  auto coarse_smooth = [&](int l, int adj) {
    level(l, false);
    if (d0->smoother27 == 0 && pmg::gsrb27_use_planes(L[l].nx, L[l].ny, L[l].nz))
      pmg::gsrb27_planes_batch<T>(s, L[l].nx, L[l].ny, L[l].nz, la, n, 2, adj);
    else if (d0->smoother27 == 0 && pmg::gsrb27_use_pairs(L[l].nx))
      pmg::gsrb27_pairs_batch<T>(s, L[l].nx, L[l].ny, L[l].nz, la, n, 2, adj);
    else
      pmg::gsrb27_batch<T>(s, L[l].nx, L[l].ny, L[l].nz, la, n, 2, adj);
  };
  const int ecyc = n == 1 ? tbegin(d0) : -1;
  // calculations of a batch start alike (every mixed-precision cycle from zero)
  fine_smooth(0, cs[0].zero_x);
  level(0, true);
  pmg::resid_batch<T>(s, 7, L[0].nx, L[0].ny, L[0].nz, la, n);
  for (int l = 1; l < nt; l++) {
    // the coarse solution starts at zero (interior cleared with the
    // restriction; the boundary layer is never written)
    for (int b = 0; b < n; b++)
      ta[b] = pmg::TransferArgs<T>{cs[b].w1, cs[b].w0[l], l != nt - 1 ? cs[b].X[l] : nullptr,
                                   nullptr};
    pmg::restrict3_batch<T>(s, L[l - 1].nx, L[l - 1].ny, L[l - 1].nz, L[l].nx, L[l].ny, L[l].nz,
                            ta, n);
    if (l != nt - 1) {
      coarse_smooth(l, 0);
      level(l, true);
      pmg::resid_batch<T>(s, 27, L[l].nx, L[l].ny, L[l].nz, la, n);
    }
  }
  const DLevel& C = L[nt - 1];
  pmg::DenseArgs<T> da[pmg::kMaxBatch];
  for (int b = 0; b < n; b++) {
    join_dense(ds[b]);
    da[b] = pmg::DenseArgs<T>{ds[b]->dense, cs[b].w0[nt - 1], cs[b].X[nt - 1]};
  }
  pmg::dense_solve_batch<T>(s, C.nx, C.ny, C.nz, d0->dn, da, n);
  pmg::DampArgs<T> pa[pmg::kMaxBatch];
  for (int l = nt - 2; l >= 0; l--) {
    const DLevel& Cl = L[l + 1];
    for (int b = 0; b < n; b++) {
      pa[b] = pmg::DampArgs<T>{cs[b].A[l + 1], cs[b].Dl[l + 1], cs[b].X[l + 1], cs[b].w0[l + 1],
                               ds[b]->part, ds[b]->scal + 4};
      ta[b] = pmg::TransferArgs<T>{cs[b].X[l + 1], cs[b].X[l], nullptr, ds[b]->scal + 4};
    }
    pmg::damp_factor_batch<T>(s, Cl.nx, Cl.ny, Cl.nz, pa, n);
    pmg::prolong_add_batch<T>(s, Cl.nx, Cl.ny, Cl.nz, L[l].nx, L[l].ny, L[l].nz, ta, n);
    if (l == 0)
      fine_smooth(1, false);
    else
      coarse_smooth(l, 1);
  }
  if (n == 1) tend(d0, ecyc, 1, 1);
}

template <class T>
void vcycle(PmgDev* d, CycleArgs<T>& c) {
  vcycle_batch<T>(&d, &c, 1);
}

// One V-cycle: replayed from a captured graph, or launched eagerly while
// kernel timing is on (its events are host-side bookkeeping).  The cycle's
// kernels depend only on device data, so the graph of a given (precision,
// solution, right-hand side) is valid for every later cycle on this
// hierarchy.  Batches of several calculations are launched eagerly.
template <class T>
void run_vcycle(PmgDev* const* ds, CycleArgs<T>* cs, int n) {
  PmgDev* d = ds[0];
  if (n > 1 || !d->graphs || d->timing) {
    vcycle_batch<T>(ds, cs, n);
    return;
  }
  CycleArgs<T>& c = cs[0];
  for (auto& cg : d->cycle_graphs)
    if (cg.word == (int)sizeof(T) && cg.f == (const void*)c.f && cg.x == (const void*)c.X[0]) {
      HIPCHECK(hipGraphLaunch(cg.e, d->s));
      return;
    }
  PmgDev::CycleGraph cg;
  cg.word = (int)sizeof(T);
  cg.f = c.f;
  cg.x = c.X[0];
  HIPCHECK(hipStreamBeginCapture(d->s, hipStreamCaptureModeThreadLocal));
  vcycle<T>(d, c);
  HIPCHECK(hipStreamEndCapture(d->s, &cg.g));
  HIPCHECK(hipGraphInstantiate(&cg.e, cg.g, nullptr, nullptr, 0));
  d->cycle_graphs.push_back(cg);
  HIPCHECK(hipGraphLaunch(cg.e, d->s));
}

CycleArgs<double> args64(PmgDev* d, double* const* X, const double* f) {
  CycleArgs<double> c;
  for (int l = 0; l < d->ntot; l++) {
    c.X[l] = X[l];
    c.Dl[l] = d->L[l].D;
    c.A[l] = d->L[l].ops();
    c.w0[l] = d->L[l].w0;
    c.xsl[l] = d->L[l].xs;
  }
  c.f = f;
  c.w1 = d->w1;
  c.xs = d->xs64;
  return c;
}

CycleArgs<float> args32(PmgDev* d) {
  CycleArgs<float> c;
  for (int l = 0; l < d->ntot; l++) {
    c.X[l] = d->L[l].xfl;
    c.Dl[l] = d->L[l].Df;
    c.A[l] = d->L[l].opsf();
    c.w0[l] = d->L[l].w0f;
    c.xsl[l] = d->L[l].xsf;
  }
  c.w1 = d->w1f;
  c.xs = d->xs32;
  // r (FP64) -> FP32 rhs stored in L[0].w0f
  c.f = d->L[0].w0f;
  return c;
}

// Wait for the shared stream and check each active calculation's norm (read
// from its host-mapped scalar); a failure (solve_failed) deactivates it.
void read_norms(PmgDev* const* ds, int n, const bool* active, double* out, bool* failed) {
  HIPCHECK(hipStreamSynchronize(ds[0]->s));
  for (int b = 0; b < n; b++) {
    if (!active[b]) continue;
    out[b] = ds[b]->hscal[0];
    if (solve_failed(ds[b], out[b])) failed[b] = true;
  }
}

// A Newton step's first trial point, formed with the inner solve's residual
// (jres_trial_nrm1): y = u + z for the correction z the cycle produced, rn the
// nonlinear residual at y.  After the inner solve, ok[b] says whether
// calculation b ran a cycle (so y and rn belong to its final correction) and
// nrm[b] holds ||rn||1.
// rn32[b], when set, receives the FP32 image of rn.  A mixed-precision solve
// that starts from zero reads its first correction straight from the FP32
// cycle's solution; x_pending[b] then says that z itself was not written
// (it equals that FP32 solution converted).
struct Trials {
  const double* u[pmg::kMaxBatch];
  double* y[pmg::kMaxBatch];
  double* rn[pmg::kMaxBatch];
  float* rn32[pmg::kMaxBatch];
  const pmg::Ions* I[pmg::kMaxBatch];
  double nrm[pmg::kMaxBatch];
  bool ok[pmg::kMaxBatch];
  bool x_pending[pmg::kMaxBatch];
};

// The inner residual f + J z of calculation b's fine level (its L1 norm to
// hscal[0]; its FP32 image to r32 when given) and, with trials, its trial
// point and nonlinear residual (norm to hscal[1]) in the same pass.  With zf
// given, z is read as (double) zf.
void inner_resid(PmgDev* d, const double* f, const double* z, float* r32, Trials* tr, int b,
                 const float* zf = nullptr) {
  const DLevel& F = d->L[0];
  if (tr) {
    pmg::jres_trial_nrm1(d->s, F.nx, F.ny, F.nz, F.ops(), F.D, F.cc, F.fc, f, tr->u[b], 1.0, z,
                         zf, tr->y[b], tr->rn[b], r32, tr->rn32[b], *tr->I[b], d->part, d->scal,
                         d->hscal, d->scal + 8, d->hscal + 1);
    tr->ok[b] = true;
  } else {
    pmg::resid7_nrm1<double>(d->s, F.nx, F.ny, F.nz, F.ops(), F.D, f, z, nullptr, d->part,
                             d->scal, d->hscal, r32);
  }
}

// The trial norms of the calculations that computed a trial this round
// (read after the stream synchronization of read_norms).
void read_trials(PmgDev* const* ds, int n, const bool* ran, Trials* tr) {
  if (!tr) return;
  for (int b = 0; b < n; b++)
    if (ran[b]) tr->nrm[b] = ds[b]->hscal[1];
}

// Vmvcs with istop 0 (absolute) or 1 (relative to ||f||1) for n calculations
// in lockstep: each runs V-cycles until its own stopping test holds.  its[b]
// gets the cycles, or -1 when that solve failed (solve_failed).  With tr,
// each cycle's residual also forms the Newton trial point (Trials).
void mvcs64_batch(PmgDev* const* ds, int n, double* const* const* X, const double* const* f,
                  int istop, int itmax, const double* errtol, double* relres, int* its,
                  Trials* tr = nullptr) {
  bool active[pmg::kMaxBatch], failed[pmg::kMaxBatch] = {};
  double rsden[pmg::kMaxBatch], rsnrm[pmg::kMaxBatch];
  for (int b = 0; b < n; b++) {
    active[b] = true;
    rsden[b] = 1.0;
    its[b] = 0;
  }
  if (istop == 1) {
    for (int b = 0; b < n; b++) {
      const DLevel& F = ds[b]->L[0];
      pmg::nrm1<double>(ds[b]->s, F.nx, F.ny, F.nz, f[b], ds[b]->part, ds[b]->scal,
                        ds[b]->hscal);
    }
    read_norms(ds, n, active, rsden, failed);
  }
  for (int b = 0; b < n; b++) {
    if (failed[b]) active[b] = false;
    if (rsden[b] == 0.0) rsden[b] = 1.0;
    rsnrm[b] = rsden[b];
  }
  CycleArgs<double> cs[pmg::kMaxBatch];
  PmgDev* act[pmg::kMaxBatch];
  for (;;) {
    int m = 0;
    for (int b = 0; b < n; b++)
      if (active[b]) {
        cs[m] = args64(ds[b], X[b], f[b]);
        act[m++] = ds[b];
      }
    if (m == 0) break;
    run_vcycle<double>(act, cs, m);
    for (int b = 0; b < n; b++) {
      if (!active[b]) continue;
      its[b]++;
      inner_resid(ds[b], f[b], X[b][0], nullptr, tr, b);
    }
    read_norms(ds, n, active, rsnrm, failed);
    read_trials(ds, n, active, tr);
    for (int b = 0; b < n; b++) {
      if (!active[b]) continue;
      if (failed[b] || its[b] >= itmax || !((rsnrm[b] / rsden[b]) > errtol[b])) active[b] = false;
    }
  }
  for (int b = 0; b < n; b++) {
    if (failed[b]) {
      its[b] = -1;
      continue;
    }
    if (relres) relres[b] = rsnrm[b] / rsden[b];
  }
}

// Mixed precision for n calculations in lockstep: FP32 V-cycles on the
// correction equation inside an FP64 defect-correction loop (fine residual,
// update and stopping test in FP64).  Each residual is written straight to
// the FP32 right-hand side of the next cycle (L[0].w0f), the values a
// conversion of the FP64 residual would give.  With x_zero the solution
// starts at zero and is not read: the first correction is assigned to it.
// outer[b] gets the outer iterations, or -1 when that solve failed
// (solve_failed).
void mixed_batch(PmgDev* const* ds, int n, double* const* const* X64, const double* const* f,
                 int istop, int itmax, const double* errtol, double* relres, int* outer,
                 bool x_zero, Trials* tr = nullptr, const double* known_fnrm = nullptr) {
  bool active[pmg::kMaxBatch], failed[pmg::kMaxBatch] = {};
  double rsden[pmg::kMaxBatch], fnrm[pmg::kMaxBatch], rsnrm[pmg::kMaxBatch];
  int it[pmg::kMaxBatch];
  for (int b = 0; b < n; b++) {
    active[b] = true;
    rsden[b] = 1.0;
    fnrm[b] = -1.0;
    it[b] = 0;
  }
  // the first residual: f itself when the solution starts at zero (f + A 0
  // rounds to f, and its FP32 image comes with its norm), else computed; a
  // known norm (known_fnrm[b] >= 0) comes with the image already in w0f
  if (istop == 1 || x_zero) {
    bool launched[pmg::kMaxBatch] = {}, any = false;
    for (int b = 0; b < n; b++) {
      PmgDev* d = ds[b];
      const DLevel& F = d->L[0];
      if (x_zero && known_fnrm && known_fnrm[b] >= 0.0) {
        fnrm[b] = known_fnrm[b];
        continue;
      }
      launched[b] = any = true;
      if (x_zero)
        pmg::nrm1_to32(d->s, F.nx, F.ny, F.nz, f[b], d->L[0].w0f, d->part, d->scal, d->hscal);
      else
        pmg::nrm1<double>(d->s, F.nx, F.ny, F.nz, f[b], d->part, d->scal, d->hscal);
    }
    if (any) read_norms(ds, n, launched, fnrm, failed);
    for (int b = 0; b < n; b++)
      if (istop == 1) rsden[b] = fnrm[b];
  }
  for (int b = 0; b < n; b++) {
    if (failed[b]) active[b] = false;
    if (rsden[b] == 0.0) rsden[b] = 1.0;
  }
  for (int b = 0; b < n; b++) {
    if (x_zero) {
      rsnrm[b] = fnrm[b];
    } else if (active[b]) {
      PmgDev* d = ds[b];
      const DLevel& F = d->L[0];
      pmg::resid7_nrm1<double>(d->s, F.nx, F.ny, F.nz, F.ops(), F.D, f[b], X64[b][0], nullptr,
                               d->part, d->scal, d->hscal, d->L[0].w0f);
    }
  }
  if (!x_zero) read_norms(ds, n, active, rsnrm, failed);
  CycleArgs<float> cs[pmg::kMaxBatch];
  PmgDev* act[pmg::kMaxBatch];
  bool pending[pmg::kMaxBatch] = {};  // X64 not written: it is (double) xfl
  for (;;) {
    int m = 0;
    for (int b = 0; b < n; b++) {
      if (!active[b]) continue;
      if (failed[b] || !(it[b] < itmax && (rsnrm[b] / rsden[b]) > errtol[b])) {
        active[b] = false;
        continue;
      }
      PmgDev* d = ds[b];
      const DLevel& F = d->L[0];
      if (pending[b]) {
        pmg::axpy_interior_f32(d->s, F.nx, F.ny, F.nz, 1.0, F.xfl, X64[b][0], true);
        pending[b] = false;
      }
      // every cycle solves for the correction from zero: the cycle's first
      // sweep does not read the fine solution, which is not cleared
      cs[m] = args32(d);
      cs[m].zero_x = true;
      act[m++] = d;
    }
    if (m == 0) break;
    run_vcycle<float>(act, cs, m);
    for (int b = 0; b < n; b++) {
      if (!active[b]) continue;
      PmgDev* d = ds[b];
      const DLevel& F = d->L[0];
      // X64 += (double) e in one pass (X64 = 0 + (double) e the first time
      // when it starts at zero); with trials, the first correction of a zero
      // start is read from e itself
      const bool direct = tr && x_zero && it[b] == 0;
      if (!direct)
        pmg::axpy_interior_f32(d->s, F.nx, F.ny, F.nz, 1.0, F.xfl, X64[b][0],
                               x_zero && it[b] == 0);
      it[b]++;
      inner_resid(d, f[b], X64[b][0], F.w0f, tr, b, direct ? F.xfl : nullptr);
      pending[b] = direct;
    }
    read_norms(ds, n, active, rsnrm, failed);
    read_trials(ds, n, active, tr);
  }
  if (tr)
    for (int b = 0; b < n; b++) tr->x_pending[b] = pending[b];
  for (int b = 0; b < n; b++) {
    // a solution that started at zero and took no correction is zero
    if (x_zero && it[b] == 0)
      pmg::fill<double>(ds[b]->s, X64[b][0], ds[b]->L[0].n, 0.0);
    if (failed[b]) {
      outer[b] = -1;
      continue;
    }
    if (relres) relres[b] = rsnrm[b] / rsden[b];
    outer[b] = it[b];
  }
}

pmg::Ions ions_scaled(Vpbe* pbe) {
  pmg::Ions I{};
  I.nion = pbe->numIon;
  double zks2 = (pbe->bulkIonicStrength > 0.0) ? 0.5 / pbe->bulkIonicStrength : 0.0;
  for (int i = 0; i < I.nion; i++) {
    I.q[i] = pbe->ionQ[i];
    I.sconc[i] = zks2 * pbe->ionConc[i];  // vpmg.c:340-377
  }
  return I;
}

// Per-calculation state of Vnewton (newtond.c:162-548) in a lockstep batch.
struct NewtonState {
  pmg::Ions I;
  double rsden = 1, xnorm_old = 0, xnorm_new = 0, xnorm_med = 0, rsnrm = 0, damp = 1;
  double errtol = 0;
  int itmax = 0;
  int idamp = 1, iters = 0, inner = 0, iter_d = 0, mode = 0;
  bool active = true, searching = false, failed = false;
  // the FP32 image of F is in the fine partner buffer xs32 (written with F by
  // the step's trial kernel); its norm is xnorm_new
  bool f32 = false;
};

// Vnewton with Vgetjac (:550-598) for n calculations of one shape in lockstep:
// each follows its own damped inexact Newton path (its own step count, line
// search and stopping test); the Jacobian inverses and the V-cycles of the
// calculations still iterating are batched.  ok[b] is false when that solve
// failed (solve_failed).
void newton_batch(Vpmg* const* th, int n, double* relres, int* iters, int* inner, bool* ok) {
  PmgDev* ds[pmg::kMaxBatch];
  NewtonState st[pmg::kMaxBatch];
  bool act[pmg::kMaxBatch];
  double nv[pmg::kMaxBatch];
  bool bad[pmg::kMaxBatch] = {};
  for (int b = 0; b < n; b++) {
    ds[b] = DEV(th[b]);
    st[b].I = ions_scaled(th[b]->pbe);
    st[b].errtol = th[b]->pmgp->errtol;
    st[b].itmax = th[b]->pmgp->itmax;
  }
  auto nresid = [&](int b, const double* x, double* r) {
    PmgDev* d = ds[b];
    const DLevel& F = d->L[0];
    pmg::nresid7_nrm1(d->s, F.nx, F.ny, F.nz, F.ops(), F.cc, F.fc, x, r, st[b].I, d->part, d->scal,
                      d->hscal);
  };
  // the trial point wn1 = u + a xtmp and its nonlinear residual in F
  auto nresid_at = [&](int b, double a, const double* xtmp) {
    PmgDev* d = ds[b];
    const DLevel& F = d->L[0];
    pmg::nresid7_xpay_nrm1(d->s, F.nx, F.ny, F.nz, F.ops(), F.cc, F.fc, d->u, a, xtmp, d->wn1,
                           d->F, st[b].I, d->part, d->scal, d->hscal);
  };
  auto read = [&](bool NewtonState::*flag) {
    for (int b = 0; b < n; b++) act[b] = st[b].*flag && !st[b].failed;
    read_norms(ds, n, act, nv, bad);
    for (int b = 0; b < n; b++)
      if (bad[b]) st[b].failed = true;
  };
  // rsden = ||F(0)||1, then ||F(u)||1
  for (int b = 0; b < n; b++) {
    pmg::fill<double>(ds[b]->s, ds[b]->wn1, ds[b]->n, 0.0);
    nresid(b, ds[b]->wn1, ds[b]->wn2);
  }
  read(&NewtonState::active);
  for (int b = 0; b < n; b++) {
    st[b].rsden = nv[b] == 0.0 ? 1.0 : nv[b];
    nresid(b, ds[b]->u, ds[b]->F);
  }
  read(&NewtonState::active);
  for (int b = 0; b < n; b++) {
    st[b].xnorm_old = st[b].xnorm_new = nv[b];
    st[b].rsnrm = st[b].rsden;
    if (st[b].failed) st[b].active = false;
  }
  double* X[pmg::kMaxBatch][24];
  for (;;) {
    int m = 0;
    PmgDev* ad[pmg::kMaxBatch];
    int idx[pmg::kMaxBatch];
    for (int b = 0; b < n; b++)
      if (st[b].active) {
        ad[m] = ds[b];
        idx[m++] = b;
      }
    if (m == 0) break;
    // Jacobian diagonal on all levels
    for (int q = 0; q < m; q++) {
      PmgDev* d = ad[q];
      DLevel* L = d->L;
      st[idx[q]].iters++;
      // cprime and the fine diagonal oC + cprime (and its FP32 image) in one pass
      pmg::dcvec(d->s, d->n, L[0].cc, d->u, L[0].cprime, st[idx[q]].I, L[0].a[0], L[0].D,
                 d->precision == 1 && d->h_mixed ? L[0].Df : nullptr);
      for (int l = 1; l < d->ntot; l++)
        pmg::restrict3<double>(d->s, L[l - 1].nx, L[l - 1].ny, L[l - 1].nz, L[l].nx, L[l].ny,
                               L[l].nz, L[l - 1].cprime, L[l].cprime);
    }
    set_diagonals_batch(ad, m, true, 1);
    // inner solve: the V-cycle clears the coarse corrections itself; the
    // fine one starts at zero
    double errs[pmg::kMaxBatch], rr[pmg::kMaxBatch];
    int its[pmg::kMaxBatch];
    double* const* Xp[pmg::kMaxBatch];
    const double* fp[pmg::kMaxBatch];
    for (int q = 0; q < m; q++) {
      PmgDev* d = ad[q];
      NewtonState& S = st[idx[q]];
      errs[q] = std::min(0.9 * S.xnorm_old, 1.0 * pow(S.xnorm_old, 2.0));
      for (int l = 0; l < d->ntot; l++) X[q][l] = d->L[l].xtmp;
      // the FP64 cycles start from a zero fine correction; the mixed solve
      // assigns its first correction (x_zero) instead of adding it to zero
      if (d->precision != 1) pmg::fill<double>(d->s, X[q][0], d->L[0].n, 0.0);
      Xp[q] = X[q];
      fp[q] = d->F;
      if (d->precision == 1 && S.iters == 1) sync_fp32_ops(d);
    }
    // the inner solve's last residual also forms each step's first trial
    // point u + xtmp (wn1) and its nonlinear residual (wn3)
    // (mixed precision, full steps: with the FP32 image of the residual,
    // which the next step's inner solve takes as its right-hand side)
    Trials tr;
    double known[pmg::kMaxBatch];
    for (int q = 0; q < m; q++) {
      PmgDev* d = ad[q];
      NewtonState& S = st[idx[q]];
      const bool mixed = d->precision == 1 && d->h_mixed;
      known[q] = -1.0;
      if (mixed && S.f32) {
        // F's FP32 image becomes the right-hand side of the FP32 cycles
        std::swap(d->L[0].w0f, d->xs32);
        known[q] = S.xnorm_new;
      }
      S.f32 = false;
      tr.u[q] = d->u;
      tr.y[q] = d->wn1;
      tr.rn[q] = d->wn3;
      tr.rn32[q] = mixed && S.idamp != 1 ? d->xs32 : nullptr;
      tr.I[q] = &st[idx[q]].I;
      tr.ok[q] = false;
      tr.x_pending[q] = false;
    }
    // calculations of one batch share the precision (batches are formed by
    // shape and precision); each has its own inner tolerance
    if (ad[0]->precision == 1)
      mixed_batch(ad, m, Xp, fp, 0, 1000, errs, rr, its, true, &tr, known);
    else
      mvcs64_batch(ad, m, Xp, fp, 0, 1000, errs, rr, its, &tr);
    // the damping search reads the correction itself
    for (int q = 0; q < m; q++) {
      PmgDev* d = ad[q];
      const DLevel& F = d->L[0];
      if (tr.x_pending[q] && st[idx[q]].idamp == 1)
        pmg::axpy_interior_f32(d->s, F.nx, F.ny, F.nz, 1.0, F.xfl, F.xtmp, true);
    }
    for (int q = 0; q < m; q++) {
      NewtonState& S = st[idx[q]];
      if (its[q] < 0) {
        S.failed = true;
        S.active = false;
        continue;
      }
      S.inner += its[q];
    }
    // first trial point (damped) or the full step: u + xtmp and its residual
    // as the inner solve left them in wn1 and wn3 (the buffers are exchanged:
    // all have a zero boundary layer and are written on the interior only),
    // or formed here when the inner solve ran no cycle
    bool formed[pmg::kMaxBatch] = {}, need_read = false;
    for (int q = 0; q < m; q++) {
      const int b = idx[q];
      NewtonState& S = st[b];
      if (!S.active) continue;
      PmgDev* d = ds[b];
      const DLevel& F = d->L[0];
      double* xtmp = F.xtmp;
      if (tr.ok[q]) {
        formed[q] = true;
        std::swap(d->F, d->wn3);
        if (S.idamp != 1) {
          std::swap(d->u, d->wn1);
          d->L[0].x = d->u;  // the fine level's solution is the potential
          S.f32 = tr.rn32[q] != nullptr;
        }
        nv[b] = tr.nrm[q];
        if (solve_failed(d, nv[b])) S.failed = true;
        continue;
      }
      need_read = true;
      if (S.idamp == 1) {
        nresid_at(b, 1.0, xtmp);
      } else {
        pmg::axpy_interior<double>(d->s, F.nx, F.ny, F.nz, 1.0, xtmp, d->u);
        nresid(b, d->u, d->F);
      }
    }
    if (need_read) {
      double keep[pmg::kMaxBatch];
      for (int q = 0; q < m; q++) keep[q] = nv[idx[q]];
      read(&NewtonState::active);
      for (int q = 0; q < m; q++)
        if (formed[q]) nv[idx[q]] = keep[q];
    }
    for (int q = 0; q < m; q++) {
      NewtonState& S = st[idx[q]];
      if (!S.active) continue;
      if (S.failed) {
        S.active = false;
        continue;
      }
      S.xnorm_new = nv[idx[q]];
      if (S.idamp == 1) {
        S.damp = 1.0;
        S.iter_d = 0;
        S.mode = 0;
        S.searching = true;
      } else {
        S.xnorm_old = S.xnorm_new;
      }
    }
    // the damping line search, one trial per round for every calculation
    // still searching
    for (;;) {
      bool any = false;
      for (int q = 0; q < m; q++) {
        const int b = idx[q];
        NewtonState& S = st[b];
        if (!S.searching) continue;
        bool stop = S.iter_d >= 10;
        if (!stop) {
          if (S.mode == 0) {
            if (S.xnorm_new < S.xnorm_old) S.mode = 1;
          } else if (S.xnorm_new > S.xnorm_med) {
            stop = true;
          }
        }
        PmgDev* d = ds[b];
        const DLevel& F = d->L[0];
        if (stop) {
          pmg::copy_interior<double>(d->s, F.nx, F.ny, F.nz, d->wn2, d->u);
          std::swap(d->F, d->wn3);
          S.xnorm_new = S.xnorm_med;
          S.xnorm_old = S.xnorm_new;
          if ((S.iter_d - 1) == 0) S.idamp = 0;
          S.searching = false;
          continue;
        }
        // keep the candidate and its residual (wn2, wn3) by exchanging the
        // buffers: every one of them has a zero boundary layer and is written
        // on the interior only
        std::swap(d->wn1, d->wn2);
        std::swap(d->F, d->wn3);
        S.xnorm_med = S.xnorm_new;
        S.damp = S.damp / 2.0;
        nresid_at(b, S.damp, F.xtmp);
        any = true;
      }
      if (!any) break;
      read(&NewtonState::searching);
      for (int q = 0; q < m; q++) {
        NewtonState& S = st[idx[q]];
        if (!S.searching) continue;
        if (S.failed) {
          S.searching = false;
          S.active = false;
          continue;
        }
        S.xnorm_new = nv[idx[q]];
        S.iter_d++;
      }
    }
    for (int q = 0; q < m; q++) {
      NewtonState& S = st[idx[q]];
      if (!S.active) continue;
      S.rsnrm = S.xnorm_new;
      if ((S.rsnrm / S.rsden) <= S.errtol || S.iters >= S.itmax) S.active = false;
    }
  }
  for (int b = 0; b < n; b++) {
    ok[b] = !st[b].failed;
    relres[b] = st[b].rsnrm / st[b].rsden;
    iters[b] = st[b].iters;
    inner[b] = st[b].inner;
  }
}

void ensure_forces(Vpmg* thee);

}  // namespace

// ===================================================================== Vpmg

extern "C" int Vpmg_ctor2(Vpmg* thee, Vpmgp* pmgp, Vpbe* pbe, int focusFlag, Vpmg* pmgOLD,
                          MGparm* mgparm, PBEparm_calcEnergy energyFlag);

extern "C" Vpmg* Vpmg_ctor(Vpmgp* pmgp, Vpbe* pbe, int focusFlag, Vpmg* pmgOLD, MGparm* mgparm,
                           PBEparm_calcEnergy energyFlag) {
  Vpmg* t = (Vpmg*)calloc(1, sizeof(Vpmg));
  if (!Vpmg_ctor2(t, pmgp, pbe, focusFlag, pmgOLD, mgparm, energyFlag)) {
    free(t);
    return nullptr;
  }
  return t;
}

static void ext_energy(Vpmg* thee, Vpmg* old, PBEparm_calcEnergy, double* partMin,
                       double* partMax, int* bflags);
static int focus_fill_bound(Vpmg* thee, Vpmg* old);

extern "C" int Vpmg_ctor2(Vpmg* thee, Vpmgp* pmgp, Vpbe* pbe, int focusFlag, Vpmg* pmgOLD,
                          MGparm* mgparm, PBEparm_calcEnergy energyFlag) {
  if (!thee || !pmgp || !pbe || !mgparm) return 0;
  thee->pmgp = pmgp;
  thee->pbe = pbe;
  Vpmgp_size(pmgp);
  if ((pmgp->nonlin == NONLIN_NPBE || pmgp->nonlin == NONLIN_SMPBE) && pmgp->meth == VSOL_Newton)
    pmgp->nrwk += 2 * (size_t)pmgp->nf;
  PmgDev* d = new PmgDev();
  thee->dev = d;
  d->s = pmgpool::stream_acquire();
  if (const char* sm = getenv("APBS_PMG_SMOOTHER")) d->smoother = strcmp(sm, "colour") == 0 ? 1 : 0;
  if (const char* sm = getenv("APBS_PMG_SMOOTHER27"))
    d->smoother27 = strcmp(sm, "colour") == 0 ? 1 : 0;
  if (const char* gr = getenv("APBS_PMG_GRAPHS")) d->graphs = strcmp(gr, "1") == 0 ? 1 : 0;
  d->nx = pmgp->nx;
  d->ny = pmgp->ny;
  d->nz = pmgp->nz;
  d->n = (size_t)d->nx * d->ny * d->nz;
  const size_t n = d->n;
  d->epsx = dalloc<double>(n);
  d->epsy = dalloc<double>(n);
  d->epsz = dalloc<double>(n);
  d->kappa = dalloc<double>(n);
  d->charge = dalloc<double>(n);
  d->u = dalloc<double>(n);
  d->pvec = dalloc<double>(n);
  d->ccf = dalloc<double>(n);
  d->gx = dalloc<double>(2 * (size_t)d->ny * d->nz);
  d->gy = dalloc<double>(2 * (size_t)d->nx * d->nz);
  d->gz = dalloc<double>(2 * (size_t)d->nx * d->ny);
  d->xf = dalloc<double>(d->nx);
  d->yf = dalloc<double>(d->ny);
  d->zf = dalloc<double>(d->nz);
  d->part = dalloc<double>(2 * pmg::kParts);
  d->scal = dalloc<double>(32);
  // device memory is recycled between objects: start every array at zero,
  // as the reference's zeroed allocations do (the boundary values of
  // bcfl zero, for one, are never written)
  {
    pmg::ZeroList z(d->s);
    for (double* a : {d->epsx, d->epsy, d->epsz, d->kappa, d->charge, d->u, d->pvec, d->ccf})
      z.add(a, n * sizeof(double));
    z.add(d->gx, 2 * (size_t)d->ny * d->nz * sizeof(double));
    z.add(d->gy, 2 * (size_t)d->nx * d->nz * sizeof(double));
    z.add(d->gz, 2 * (size_t)d->nx * d->ny * sizeof(double));
    z.add(d->part, 2 * pmg::kParts * sizeof(double));
    z.add(d->scal, 32 * sizeof(double));
  }
  d->hscal = pmgpool::pinned_scalars();
  Valist* al = pbe->alist;
  d->natoms = al->number;
  std::vector<double> pos(3 * (size_t)d->natoms), rad(d->natoms), chg(d->natoms);
  for (int a = 0; a < d->natoms; a++) {
    for (int k = 0; k < 3; k++) pos[3 * a + k] = al->atoms[a].position[k];
    rad[a] = al->atoms[a].radius;
    chg[a] = al->atoms[a].charge;
  }
  d->apos = dalloc<double>(3 * (size_t)d->natoms);
  d->arad = dalloc<double>(d->natoms);
  d->achg = dalloc<double>(d->natoms);
  d->apart = dalloc<double>(d->natoms);
  d->atmp = dalloc<double>(d->natoms);
  {
    pmg::ZeroList z(d->s);
    z.add(d->apart, (size_t)d->natoms * sizeof(double));
    z.add(d->atmp, (size_t)d->natoms * sizeof(double));
  }
  HIPCHECK(hipMemcpyAsync(d->apos, pos.data(), pos.size() * sizeof(double), hipMemcpyHostToDevice, d->s));
  HIPCHECK(hipMemcpyAsync(d->arad, rad.data(), rad.size() * sizeof(double), hipMemcpyHostToDevice, d->s));
  HIPCHECK(hipMemcpyAsync(d->achg, chg.data(), chg.size() * sizeof(double), hipMemcpyHostToDevice, d->s));
  d->hpart.assign(d->natoms, 1.0);
  // host arrays callers dereference (vpmg.h:116-193)
  thee->u = (double*)calloc(n, sizeof(double));
  thee->pvec = (double*)calloc(n, sizeof(double));
  thee->rwork = (double*)calloc(n, sizeof(double));
  if (focusFlag && pmgOLD) {
    if (pmgp->bcfl != BCFL_FOCUS) pmgp->bcfl = BCFL_FOCUS;
    if (!focus_fill_bound(thee, pmgOLD)) {
      fprintf(stderr, "Vpmg_ctor2: new mesh not contained in the old mesh\n");
      return 0;
    }
    if (energyFlag != PCE_NO) {
      double partMin[3], partMax[3];
      for (int j = 0; j < 3; j++) {
        if (mgparm->type == MCT_PARALLEL) {
          partMin[j] = mgparm->partDisjCenter[j] - 0.5 * mgparm->partDisjLength[j];
          partMax[j] = mgparm->partDisjCenter[j] + 0.5 * mgparm->partDisjLength[j];
        } else {
          partMin[j] = mgparm->center[j] - 0.5 * mgparm->glen[j];
          partMax[j] = mgparm->center[j] + 0.5 * mgparm->glen[j];
        }
      }
      ext_energy(thee, pmgOLD, energyFlag, partMin, partMax, mgparm->partDisjOwnSide);
    }
  } else {
    thee->extQmEnergy = thee->extDiEnergy = thee->extQfEnergy = 0.0;
  }
  thee->chargeSrc = mgparm->chgs;
  Vpmg_unsetPart(thee);
  thee->filled = 0;
  Vpmg_dtor(&pmgOLD);
  return 1;
}

extern "C" void Vpmg_dtor2(Vpmg* thee) {
  if (!thee) return;
  PmgDev* d = DEV(thee);
  if (d) {
    HIPCHECK(hipStreamSynchronize(d->s));
    free_hierarchy(d);
    if (d->cs) {
      HIPCHECK(hipEventDestroy(d->cs_fork));
      HIPCHECK(hipEventDestroy(d->cs_join));
      pmgpool::stream_release(d->cs);
    }
    dfree(d->epsx);
    dfree(d->epsy);
    dfree(d->epsz);
    dfree(d->kappa);
    dfree(d->charge);
    dfree(d->u);
    dfree(d->pvec);
    dfree(d->ccf);
    dfree(d->gx);
    dfree(d->gy);
    dfree(d->gz);
    dfree(d->xf);
    dfree(d->yf);
    dfree(d->zf);
    dfree(d->apos);
    dfree(d->arad);
    dfree(d->achg);
    dfree(d->apart);
    dfree(d->atmp);
    dfree(d->part);
    dfree(d->scal);
    dfree(d->fqf);
    dfree(d->splcoef);
    dfree(d->fib);
    dfree(d->fdb);
    dfree(d->frc);
    pmgpool::pinned_release(d->hscal);
    for (hipEvent_t h : d->ev) HIPCHECK(hipEventDestroy(h));
    pmgpool::stream_release(d->s);
    delete d;
    thee->dev = nullptr;
  }
  free(thee->u);
  free(thee->pvec);
  free(thee->rwork);
  free(thee->epsx);
  free(thee->epsy);
  free(thee->epsz);
  free(thee->kappa);
  free(thee->charge);
  thee->u = thee->pvec = thee->rwork = nullptr;
  thee->epsx = thee->epsy = thee->epsz = thee->kappa = thee->charge = nullptr;
}

extern "C" void Vpmg_dtor(Vpmg** thee) {
  if (thee && *thee) {
    Vpmg_dtor2(*thee);
    free(*thee);
    *thee = nullptr;
  }
}

static pmg::Geo geo_of(Vpmg* thee) {
  pmg::Geo g;
  Vpmgp* p = thee->pmgp;
  g.nx = p->nx;
  g.ny = p->ny;
  g.nz = p->nz;
  g.hx = p->hx;
  g.hy = p->hy;
  g.hz = p->hzed;
  g.xmin = p->xmin;
  g.ymin = p->ymin;
  g.zmin = p->zmin;
  g.xmax = p->xmax;
  g.ymax = p->ymax;
  g.zmax = p->zmax;
  return g;
}

static int srfm_order(Vsurf_Meth s) {
  return s == VSM_SPLINE3 ? 3 : (s == VSM_SPLINE4 ? 4 : 2);
}

// Per-atom window coefficients of srfm spl3 / spl4 on the device (computed
// on the host, so the reference's divisions and pow round the same way).
static void ensure_spline_coef(Vpmg* thee, int order) {
  PmgDev* d = DEV(thee);
  if (order == 2 || d->natoms == 0) return;
  if (d->splcoef && d->splorder == order && d->splwin == thee->splineWin) return;
  const Valist* al = thee->pbe->alist;
  std::vector<double> hr(d->natoms), hc((size_t)pmg::kSplCoefHost * d->natoms);
  for (int a = 0; a < d->natoms; a++) hr[a] = al->atoms[a].radius;
  pmg::spline_coefficients(order, d->natoms, hr.data(), thee->splineWin,
                           thee->pbe->maxIonRadius, hc.data());
  if (!d->splcoef) d->splcoef = dalloc<double>(hc.size());
  HIPCHECK(hipMemcpyAsync(d->splcoef, hc.data(), hc.size() * sizeof(double),
                          hipMemcpyHostToDevice, d->s));
  HIPCHECK(hipStreamSynchronize(d->s));
  d->splorder = order;
  d->splwin = thee->splineWin;
}

// ================================================================ Vgrid

// Vgrid_ctor (vgrid.c:86-147): the map borrows data; xmax = xmin + (nx-1) hx.
extern "C" Vgrid* Vgrid_ctor(int nx, int ny, int nz, double hx, double hy, double hzed,
                             double xmin, double ymin, double zmin, double* data) {
  if (nx < 1 || ny < 1 || nz < 1 || !(hx > 0) || !(hy > 0) || !(hzed > 0)) return nullptr;
  Vgrid* g = (Vgrid*)calloc(1, sizeof(Vgrid));
  if (!g) return nullptr;
  g->nx = nx;
  g->ny = ny;
  g->nz = nz;
  g->hx = hx;
  g->hy = hy;
  g->hzed = hzed;
  g->xmin = xmin;
  g->xmax = xmin + (nx - 1) * hx;
  g->ymin = ymin;
  g->ymax = ymin + (ny - 1) * hy;
  g->zmin = zmin;
  g->zmax = zmin + (nz - 1) * hzed;
  g->data = data;
  g->ctordata = data != nullptr;
  g->readdata = 0;
  return g;
}

// Vgrid_dtor (vgrid.c:152-175): the data stays the caller's.
extern "C" void Vgrid_dtor(Vgrid** thee) {
  if (!thee || !*thee) return;
  free(*thee);
  *thee = nullptr;
}

static pmg::MapDev map_fields(const Vgrid* v, const double* data) {
  pmg::MapDev m;
  m.nx = v->nx;
  m.ny = v->ny;
  m.nz = v->nz;
  m.hx = v->hx;
  m.hy = v->hy;
  m.hz = v->hzed;
  m.xmin = v->xmin;
  m.ymin = v->ymin;
  m.zmin = v->zmin;
  m.xmax = v->xmax;
  m.ymax = v->ymax;
  m.zmax = v->zmax;
  m.data = data;
  return m;
}

// Vgrid_value (vgrid.c:179-260), the arithmetic the device maps use.
extern "C" int Vgrid_value(Vgrid* thee, double pt[3], double* value) {
  if (!thee || !thee->data || !(thee->ctordata || thee->readdata)) return 0;
  double v = 0.0;
  if (!pmg::map_value(map_fields(thee, thee->data), pt[0], pt[1], pt[2], &v)) return 0;
  *value = v;
  return 1;
}

// A map's data copied to the device for one Vpmg_fillco.
struct DevMap {
  pmg::MapDev m{};
  double* data = nullptr;
  DevMap(PmgDev* d, const Vgrid* v) {
    const size_t n = (size_t)v->nx * v->ny * v->nz;
    data = dalloc<double>(n);
    HIPCHECK(hipMemcpyAsync(data, v->data, n * sizeof(double), hipMemcpyHostToDevice, d->s));
    m = map_fields(v, data);
  }
  ~DevMap() { dfree(data); }
  DevMap(const DevMap&) = delete;
  DevMap& operator=(const DevMap&) = delete;
};

static bool map_ok(const Vgrid* v) {
  return v && v->data && (v->ctordata || v->readdata) && v->nx >= 1 && v->ny >= 1 &&
         v->nz >= 1 && v->hx > 0 && v->hy > 0 && v->hzed > 0;
}

// Samples one map at the grid points shifted by (ox, oy, oz); returns the
// number of points off the map (and of negative values in *neg, mode 1).
static int sample_map(PmgDev* d, const pmg::Geo& g, const Vgrid* v, double ox, double oy,
                      double oz, int mode, double scale, double* out, double* vmax, int* neg) {
  DevMap dm(d, v);
  const size_t np = pmg::map_sample_parts(g);
  double* part = dalloc<double>(np);
  int* bad = dalloc<int>(2);
  HIPCHECK(hipMemsetAsync(bad, 0, 2 * sizeof(int), d->s));
  pmg::map_sample(d->s, g, dm.m, d->xf, d->yf, d->zf, ox, oy, oz, mode, scale, out, part, bad);
  int hbad[2] = {0, 0};
  std::vector<double> hp(mode == 1 ? np : 0);
  HIPCHECK(hipMemcpyAsync(hbad, bad, 2 * sizeof(int), hipMemcpyDeviceToHost, d->s));
  if (mode == 1)
    HIPCHECK(hipMemcpyAsync(hp.data(), part, np * sizeof(double), hipMemcpyDeviceToHost, d->s));
  HIPCHECK(hipStreamSynchronize(d->s));
  dfree(part);
  dfree(bad);
  if (mode == 1) {
    double mx = -1.0;
    for (double x : hp) mx = std::max(mx, x);
    *vmax = mx;
    *neg = hbad[1];
  }
  return hbad[0];
}

// fillcoCoefMap (vpmg.c:4098-4220): every dielectric map, and the kappa map
// when there is salt; kappa scaled into [0, 1] by the map's maximum.
static int fill_coef_map(Vpmg* thee, const pmg::Geo& g, double ionstr) {
  PmgDev* d = DEV(thee);
  if (!thee->useDielXMap || !thee->useDielYMap || !thee->useDielZMap ||
      (!thee->useKappaMap && ionstr > VPMGSMALL)) {
    fprintf(stderr, "fillcoCoefMap: You need to use all coefficient maps!\n");
    return 0;
  }
  if (ionstr > VPMGSMALL) {
    double kmax = -1.0;
    int neg = 0;
    if (sample_map(d, g, thee->kappaMap, 0, 0, 0, 1, 1.0, d->kappa, &kmax, &neg)) {
      fprintf(stderr, "Vpmg_fillco: Off kappaMap\n");
      return 0;
    }
    if (neg) {
      fprintf(stderr, "Vpmg_fillcoCoefMap: Kappa map less than 0\n");
      return 0;
    }
    if (!(kmax > 1.0)) kmax = 1.0;
    pmg::kappa_scale(d->s, d->n, kmax, d->kappa);
  } else {
    pmg::fill<double>(d->s, d->kappa, d->n, 0.0);
  }
  const Vgrid* maps[3] = {thee->dielXMap, thee->dielYMap, thee->dielZMap};
  double* eps[3] = {d->epsx, d->epsy, d->epsz};
  const char* names[3] = {"dielXMap", "dielYMap", "dielZMap"};
  for (int w = 0; w < 3; w++) {
    const double ox = w == 0 ? 0.5 * g.hx : 0.0, oy = w == 1 ? 0.5 * g.hy : 0.0,
                 oz = w == 2 ? 0.5 * g.hz : 0.0;
    if (sample_map(d, g, maps[w], ox, oy, oz, 0, 1.0, eps[w], nullptr, nullptr)) {
      fprintf(stderr, "Vpmg_fillco: Off %s\n", names[w]);
      return 0;
    }
  }
  return 1;
}

static int focus_fill_bound_grid(Vpmg* thee, const pmg::OldGrid& og);

// Vpmg_fillco (vpmg.c:5250-5415)
extern "C" int Vpmg_fillco(Vpmg* thee, Vsurf_Meth surfMeth, double splineWin,
                           Vchrg_Meth chargeMeth, int useDielXMap, Vgrid* dielXMap,
                           int useDielYMap, Vgrid* dielYMap, int useDielZMap, Vgrid* dielZMap,
                           int useKappaMap, Vgrid* kappaMap, int usePotMap, Vgrid* potMap,
                           int useChargeMap, Vgrid* chargeMap) {
  if (!thee) return 0;
  // vpmg.c:5299-5310
  const struct { int use; Vgrid* map; int* flag; Vgrid** slot; const char* name; } maps[6] = {
      {useDielXMap, dielXMap, &thee->useDielXMap, &thee->dielXMap, "dielXMap"},
      {useDielYMap, dielYMap, &thee->useDielYMap, &thee->dielYMap, "dielYMap"},
      {useDielZMap, dielZMap, &thee->useDielZMap, &thee->dielZMap, "dielZMap"},
      {useKappaMap, kappaMap, &thee->useKappaMap, &thee->kappaMap, "kappaMap"},
      {usePotMap, potMap, &thee->usePotMap, &thee->potMap, "potMap"},
      {useChargeMap, chargeMap, &thee->useChargeMap, &thee->chargeMap, "chargeMap"}};
  for (const auto& m : maps) {
    if (m.use && !map_ok(m.map)) {
      fprintf(stderr, "Vpmg_fillco: %s is missing or has no data\n", m.name);
      return 0;
    }
  }
  for (const auto& m : maps) {
    *m.flag = m.use;
    *m.slot = m.use ? m.map : nullptr;
  }
  const double t0 = now();
  PmgDev* d = DEV(thee);
  Vpmgp* p = thee->pmgp;
  Vpbe* pbe = thee->pbe;
  thee->surfMeth = surfMeth;
  thee->splineWin = splineWin;
  thee->chargeMeth = chargeMeth;
  const double ionstr = pbe->bulkIonicStrength, epsw = pbe->solventDiel, epsp = pbe->soluteDiel;
  p->xmin = p->xcent - (p->xlen / 2.0);
  p->ymin = p->ycent - (p->ylen / 2.0);
  p->zmin = p->zcent - (p->zlen / 2.0);
  p->xmax = p->xcent + (p->xlen / 2.0);
  p->ymax = p->ycent + (p->ylen / 2.0);
  p->zmax = p->zcent + (p->zlen / 2.0);
  // the operator is not a plain Laplacian when maps are read (vpmg.c:5351-5362)
  const bool coef_maps =
      thee->useDielXMap || thee->useDielYMap || thee->useDielZMap || thee->useKappaMap;
  const int islap = !(coef_maps || thee->usePotMap) && (ionstr < VPMGSMALL) &&
                    (fabs(epsp - epsw) < VPMGSMALL);
  std::vector<double> xf(d->nx), yf(d->ny), zf(d->nz);
  for (int i = 0; i < d->nx; i++) xf[i] = p->xmin + i * p->hx;
  for (int i = 0; i < d->ny; i++) yf[i] = p->ymin + i * p->hy;
  for (int i = 0; i < d->nz; i++) zf[i] = p->zmin + i * p->hzed;
  HIPCHECK(hipMemcpyAsync(d->xf, xf.data(), d->nx * sizeof(double), hipMemcpyHostToDevice, d->s));
  HIPCHECK(hipMemcpyAsync(d->yf, yf.data(), d->ny * sizeof(double), hipMemcpyHostToDevice, d->s));
  HIPCHECK(hipMemcpyAsync(d->zf, zf.data(), d->nz * sizeof(double), hipMemcpyHostToDevice, d->s));
  const pmg::Geo g = geo_of(thee);
  d->geo = g;
  const pmg::AtomsDev A = atoms_dev(d);
  // The MDH boundary (FP64-bound, independent of the charge and coefficient
  // maps) runs on a second stream while those are built; the object's
  // stream waits for it at the boundary section.  The guard joins and frees
  // on every return path.
  struct SideBoundary {
    hipStream_t s2 = nullptr;
    hipEvent_t done = nullptr;
    double* part = nullptr;
    ~SideBoundary() {
      if (s2) {
        HIPCHECK(hipStreamSynchronize(s2));
        pmgpool::stream_release(s2);
      }
      if (done) HIPCHECK(hipEventDestroy(done));
      dfree(part);
    }
  } side;
  const double pre1 =
      ((Vunit_ec) / (4 * VPI * Vunit_eps0 * epsw * Vunit_kb * pbe->T)) * (1.0e10);
  if (p->bcfl == BCFL_MDH && d->side) {
    hipEvent_t ready;
    HIPCHECK(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
    HIPCHECK(hipEventRecord(ready, d->s));  // xf, yf, zf uploaded
    side.s2 = pmgpool::stream_acquire();
    HIPCHECK(hipStreamWaitEvent(side.s2, ready, 0));
    HIPCHECK(hipEventDestroy(ready));
    pmg::fill<double>(side.s2, d->gx, 2 * (size_t)d->ny * d->nz, 0.0);
    pmg::fill<double>(side.s2, d->gy, 2 * (size_t)d->nx * d->nz, 0.0);
    pmg::fill<double>(side.s2, d->gz, 2 * (size_t)d->nx * d->ny, 0.0);
    // bcflnew (vpmg.c:3289-3367)
    side.part = dalloc<double>(pmg::bc_mdh_scratch(g, d->natoms));
    pmg::bc_mdh(side.s2, g, d->xf, d->yf, d->zf, A, pre1, pbe->xkappa, side.part, d->gx, d->gy,
                d->gz, d->fast_boundary);
    HIPCHECK(hipEventCreateWithFlags(&side.done, hipEventDisableTiming));
    HIPCHECK(hipEventRecord(side.done, side.s2));
  }
  struct TilesGuard {
    hipStream_t s;
    DevTiles t;
    ~TilesGuard() {
      if (t.start || t.list) {
        HIPCHECK(hipStreamSynchronize(s));
        t.release();
      }
    }
  } charge_tiles{d->s, {}};
  Valist* al = pbe->alist;
  std::vector<double> hpos(3 * (size_t)d->natoms), hrad(d->natoms);
  double rmax = 0;
  for (int a = 0; a < d->natoms; a++) {
    for (int k = 0; k < 3; k++) hpos[3 * a + k] = al->atoms[a].position[k];
    hrad[a] = al->atoms[a].radius;
    rmax = std::max(rmax, hrad[a]);
  }
  const double hmax = std::max(std::max(p->hx, p->hy), p->hzed);
  // charges: an atom reaches at most 2 grid spacings per axis (3 for spl4)
  if (thee->useChargeMap) {
    // fillcoChargeMap (vpmg.c:4938-4985): the map in e/A^3 times zmagic
    if (sample_map(d, g, thee->chargeMap, 0, 0, 0, 2, pbe->zmagic, d->charge, nullptr, nullptr)) {
      fprintf(stderr, "fillcoChargeMap: Error -- fell off of charge map\n");
      return 0;
    }
  } else {
    // the lists are released when fillco returns, after its stream has
    // drained (no synchronization here: the surface maps queue behind)
    const double reach = (chargeMeth == VCM_BSPL4 ? 3.0 : 2.0) * hmax;
    charge_tiles.t = build_tiles(d->s, g, hpos.data(), d->natoms, {reach});
    pmg::charge_map(d->s, g, A, charge_tiles.t.t, (int)chargeMeth, pbe->zmagic, 0, d->charge);
  }
  if (!islap && coef_maps) {
    if (!fill_coef_map(thee, g, ionstr)) return 0;
  } else if (!islap) {
    const double irad = pbe->maxIonRadius;
    const double ionmask = (ionstr > VPMGSMALL) ? 1.0 : 0.0;
    if (surfMeth == VSM_SPLINE || surfMeth == VSM_SPLINE3 || surfMeth == VSM_SPLINE4) {
      // fillcoCoefSpline (vpmg.c:4631), fillcoCoefSpline3 (:7525), fillcoCoefSpline4 (:7052)
      const int order = srfm_order(surfMeth);
      ensure_spline_coef(thee, order);
      // the kernel's box: max(irad + r, r) + win + h/2 around the atom
      std::vector<double> reach(std::max(d->natoms, 1));
      for (int a = 0; a < d->natoms; a++)
        reach[a] = std::max(irad + hrad[a] + splineWin, hrad[a] + splineWin) + 0.5 * hmax;
      DevTiles tiles = build_tiles(d->s, g, hpos.data(), d->natoms, reach);
      pmg::coef_spline(d->s, g, A, tiles.t, splineWin, irad, ionmask, epsw, epsp, d->kappa,
                       d->epsx, d->epsy, d->epsz, order, order == 2 ? nullptr : d->splcoef);
      HIPCHECK(hipStreamSynchronize(d->s));
      tiles.release();
    } else {
      // fillcoCoefMolDiel (vpmg.c:4335-4344) then fillcoCoefMolIon, as flag
      // maps (byte per point): the atom spheres on the three shifted grids
      // and, with salt, the ion-accessibility spheres on the unshifted one;
      // the probe spheres at the SAS points, which re-mark the solvent
      // value; then one pass forms the maps (and the harmonic mean for smol).
      // Layout: [0, 3n) atom flags x, y, z; [3n, 4n) kappa; [4n, 7n) SAS;
      // [7n, 8n) the solute bits mol_maps forms (written whole, not cleared).
      const double srad = pbe->solventRadius;
      const bool ion_marks = !(ionstr < VPMGSMALL);
      const size_t n = d->n;
      const size_t fbytes = (7 * n + 3) & ~(size_t)3;
      unsigned char* mflags = dalloc<unsigned char>(fbytes + n);
      {
        pmg::ZeroList z(d->s);
        z.add(mflags, fbytes);
      }
      pmg::MarkTargets T{};
      for (int w = 0; w < 3; w++) T.t[T.nt++] = pmg::MarkTarget{w, 0, srad};
      if (ion_marks) T.t[T.nt++] = pmg::MarkTarget{3, 1, irad};
      pmg::mark_flags(d->s, g, d->natoms, d->apos, d->arad, T, mflags);
      if (srad > VSMALL) {
        // the accessibility object's reference sphere (vacc.c:146-189), read
        // from the Vpbe so a Vpbe built by the reference is honoured
        const VaccSurf* rs = pbe->acc->refSphere;
        const int nref = rs->npts;
        std::vector<double> ref(3 * (size_t)nref);
        for (int q = 0; q < nref; q++) {
          ref[3 * q] = rs->xpts[q];
          ref[3 * q + 1] = rs->ypts[q];
          ref[3 * q + 2] = rs->zpts[q];
        }
        double* dref = dalloc<double>(ref.size());
        HIPCHECK(hipMemcpyAsync(dref, ref.data(), ref.size() * sizeof(double),
                                hipMemcpyHostToDevice, d->s));
        DevCells cells = build_cells(d->s, hpos.data(), d->natoms, rmax + srad);
        const size_t tot = (size_t)d->natoms * nref;
        unsigned char* flags = dalloc<unsigned char>(tot);
        pmg::sas_flags(d->s, A, cells.c, dref, nref, srad, flags);
        int* counter = dalloc<int>(1024);
        double* sas = nullptr;
        int nsas = pmg::sas_compact(d->s, A, dref, nref, srad, flags, counter, &sas,
                                    [](size_t n) { return dalloc<double>(n); });
        pmg::MarkTargets S{};
        for (int w = 0; w < 3; w++) S.t[S.nt++] = pmg::MarkTarget{w, 2, srad};
        pmg::mark_flags(d->s, g, nsas, sas, nullptr, S, mflags + 4 * n);
        HIPCHECK(hipStreamSynchronize(d->s));
        dfree(flags);
        dfree(sas);
        dfree(counter);
        dfree(dref);
        cells.release();
      }
      pmg::mol_maps(d->s, g, mflags, srad > VSMALL ? mflags + 4 * n : nullptr,
                    ion_marks ? mflags + 3 * n : nullptr, epsp, epsw, ionmask,
                    surfMeth == VSM_MOLSMOOTH ? 1 : 0, d->epsx, d->epsy, d->epsz, d->kappa,
                    mflags + fbytes);
      HIPCHECK(hipStreamSynchronize(d->s));
      dfree(mflags);
    }
  } else {
    pmg::fill<double>(d->s, d->kappa, d->n, 0.0);
    pmg::fill<double>(d->s, d->epsx, d->n, epsp);
    pmg::fill<double>(d->s, d->epsy, d->n, epsp);
    pmg::fill<double>(d->s, d->epsz, d->n, epsp);
  }
  if (side.done) {
    HIPCHECK(hipStreamWaitEvent(d->s, side.done, 0));  // MDH on the second stream
  } else if (p->bcfl != BCFL_FOCUS) {
    pmg::fill<double>(d->s, d->gx, 2 * (size_t)d->ny * d->nz, 0.0);
    pmg::fill<double>(d->s, d->gy, 2 * (size_t)d->nx * d->nz, 0.0);
    pmg::fill<double>(d->s, d->gz, 2 * (size_t)d->nx * d->ny, 0.0);
    if (p->bcfl == BCFL_MDH) {
      // bcflnew (vpmg.c:3289-3367)
      double* part = dalloc<double>(pmg::bc_mdh_scratch(g, d->natoms));
      pmg::bc_mdh(d->s, g, d->xf, d->yf, d->zf, A, pre1, pbe->xkappa, part, d->gx, d->gy, d->gz,
                  d->fast_boundary);
      HIPCHECK(hipStreamSynchronize(d->s));
      dfree(part);
    } else if (p->bcfl == BCFL_SDH) {
      // bcfl_sdh (vpmg.c:3479-3645): moments summed on the host in atom order
      pmg::SdhMoments M{};
      const double* c = pbe->soluteCenter;
      double quad[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, tr[9];
      for (int d = 0; d < 3; d++) M.pos[d] = c[d];
      M.size = pbe->soluteRadius;
      for (int a = 0; a < d->natoms; a++) {
        const Vatom* at = &al->atoms[a];
        const double xr = at->position[0] - c[0], yr = at->position[1] - c[1],
                     zr = at->position[2] - c[2];
        const double ch = at->charge;
        M.q += ch;
        M.u[0] += xr * ch;
        M.u[1] += yr * ch;
        M.u[2] += zr * ch;
        tr[0] = xr * xr * ch; tr[1] = xr * yr * ch; tr[2] = xr * zr * ch;
        tr[3] = yr * xr * ch; tr[4] = yr * yr * ch; tr[5] = yr * zr * ch;
        tr[6] = zr * xr * ch; tr[7] = zr * yr * ch; tr[8] = zr * zr * ch;
        const double qave = (tr[0] + tr[4] + tr[8]) / 3.0;
        for (int e = 0; e < 9; e++) quad[e] += 1.5 * ((e % 4 == 0) ? (tr[e] - qave) : tr[e]);
      }
      M.qxx = quad[0] / 3.0;
      M.qxy = quad[1] / 3.0;
      M.qxz = quad[2] / 3.0;
      M.qyy = quad[4] / 3.0;
      M.qyz = quad[5] / 3.0;
      M.qzz = quad[8] / 3.0;
      M.pre = ((Vunit_ec * Vunit_ec) / (4 * VPI * Vunit_eps0 * Vunit_kb * pbe->T)) * (1.0e10);
      M.xkappa = pbe->xkappa;
      M.eps_p = pbe->soluteDiel;
      M.eps_w = pbe->solventDiel;
      pmg::bc_sdh(d->s, g, d->xf, d->yf, d->zf, M, d->gx, d->gy, d->gz);
    } else if (p->bcfl == BCFL_MAP) {
      // bcfl_map (vpmg.c:3953-3990): every grid point must lie on the
      // potential map; then focusFillBound with the map as the old solution
      // (vpmg.c:1993-2006, 4084-4087)
      if (!thee->usePotMap) {
        fprintf(stderr, "Vpmg_fillco: bcfl map needs a potential map (usemap pot)\n");
        return 0;
      }
      if (sample_map(d, g, thee->potMap, 0, 0, 0, 0, 1.0, nullptr, nullptr, nullptr)) {
        fprintf(stderr, "bcfl_map: Error -- fell off of potential map\n");
        return 0;
      }
      const Vgrid* pm = thee->potMap;
      DevMap dm(d, pm);
      pmg::OldGrid og;
      og.nx = pm->nx;
      og.ny = pm->ny;
      og.nz = pm->nz;
      og.hx = pm->hx;
      og.hy = pm->hy;
      og.hz = pm->hzed;
      og.xmin = pm->xmin;
      og.xmax = pm->xmax;
      og.ymin = pm->ymin;
      og.ymax = pm->ymax;
      og.zmin = pm->zmin;
      og.zmax = pm->zmax;
      og.u = dm.data;
      if (!focus_fill_bound_grid(thee, og)) {
        fprintf(stderr, "Vpmg::focusFillBound -- new mesh not contained in the potential map\n");
        return 0;
      }
    } else if (p->bcfl != BCFL_ZERO) {
      fprintf(stderr, "Vpmg_fillco: bcfl %d is not supported (zero, sdh, mdh, map, focus are)\n",
              (int)p->bcfl);
      return 0;
    }
  }
  HIPCHECK(hipStreamSynchronize(d->s));
  thee->filled = 1;
  d->forces_valid = false;
  d->stats[5] = now() - t0;
  return 1;
}

// Vpmg_solve (vpmg.c:401-570)
static_assert(APBS_PMG_MAXBATCH == pmg::kMaxBatch, "the public batch size is the kernels'");

// Whether two filled calculations can share batched launches: the same grid,
// hierarchy depth, method, precision and smoother forms.
static bool same_batch(Vpmg* a, Vpmg* b) {
  PmgDev* x = DEV(a);
  PmgDev* y = DEV(b);
  const Vpmgp* p = a->pmgp;
  const Vpmgp* q = b->pmgp;
  return x->nx == y->nx && x->ny == y->ny && x->nz == y->nz && p->nlev == q->nlev &&
         p->meth == q->meth && x->precision == y->precision && x->smoother == y->smoother &&
         x->smoother27 == y->smoother27;
}

// Vpmg_solve (vpmg.c:401-570) for n <= pmg::kMaxBatch calculations that
// same_batch accepts, all on the first one's stream (the others' streams are
// drained first and restored after): each calculation's preparation, then
// the lockstep solve, then each one's Dirichlet faces and statistics.
// Returns how many succeeded; ok[b] per calculation.
static int solve_group(Vpmg* const* th, int n, int* ok) {
  PmgDev* ds[pmg::kMaxBatch];
  hipStream_t own[pmg::kMaxBatch];
  for (int b = 0; b < n; b++) {
    ds[b] = DEV(th[b]);
    own[b] = ds[b]->s;
    if (b > 0) {
      HIPCHECK(hipStreamSynchronize(own[b]));
      ds[b]->s = ds[0]->s;
    }
  }
  hipStream_t s = ds[0]->s;
  const double t0 = now();
  Vpmgp* p0 = th[0]->pmgp;
  const bool newton = (p0->meth == VSOL_Newton);
  for (int b = 0; b < n; b++) {
    PmgDev* d = ds[b];
    const double zkappa2 = th[b]->pbe->zkappa2;
    // ccf = zkappa2 * kappa, or 0 (vpmg.c:437-447)
    if (zkappa2 > VPMGSMALL)
      pmg::scale(s, d->n, zkappa2, d->kappa, d->ccf);
    else
      pmg::fill<double>(s, d->ccf, d->n, 0.0);
    alloc_hierarchy(d, th[b]->pmgp->nlev, newton, d->precision == 1);
    // no inverse of an earlier solve is in flight (every solve ends with both
    // streams drained), so the pivot flag can be cleared from the host
    d->hscal[kPivotFlag] = 0.0;
    d->failed = false;
  }
  const double tb = now();
  for (int b = 0; b < n; b++) {
    build_operators(ds[b]);
    pmg::fill<double>(s, ds[b]->u, ds[b]->n, 0.0);
  }
  double relres[pmg::kMaxBatch] = {};
  int iters[pmg::kMaxBatch] = {}, inner[pmg::kMaxBatch] = {}, outer[pmg::kMaxBatch] = {};
  bool good[pmg::kMaxBatch];
  if (newton) {
    newton_batch(th, n, relres, iters, inner, good);
  } else {
    set_diagonals_batch(ds, n, false);
    double* X[pmg::kMaxBatch][24];
    double* const* Xp[pmg::kMaxBatch];
    const double* fp[pmg::kMaxBatch];
    double tol[pmg::kMaxBatch];
    for (int b = 0; b < n; b++) {
      for (int l = 0; l < ds[b]->ntot; l++) X[b][l] = ds[b]->L[l].x;
      Xp[b] = X[b];
      fp[b] = ds[b]->L[0].fc;
    }
    HIPCHECK(hipStreamSynchronize(s));
    for (int b = 0; b < n; b++) ds[b]->stats[6] = now() - tb;
    if (ds[0]->precision == 1) {
      for (int b = 0; b < n; b++) {
        sync_fp32_ops(ds[b]);
        tol[b] = ds[b]->mixed_tol > 0 ? ds[b]->mixed_tol : th[b]->pmgp->errtol;
      }
      mixed_batch(ds, n, Xp, fp, 1, p0->itmax * 4, tol, relres, outer, true);
      for (int b = 0; b < n; b++) iters[b] = outer[b];
    } else {
      for (int b = 0; b < n; b++) tol[b] = th[b]->pmgp->errtol;
      mvcs64_batch(ds, n, Xp, fp, 1, p0->itmax, tol, relres, iters);
    }
    for (int b = 0; b < n; b++) good[b] = iters[b] >= 0;
  }
  int nok = 0;
  for (int b = 0; b < n; b++) {
    PmgDev* d = ds[b];
    ok[b] = good[b] ? 1 : 0;
    if (!good[b]) {
      // every kernel has been launched normally; drain both streams so the
      // object stays usable (a later fillco / solve / dtor)
      if (d->cs) HIPCHECK(hipStreamSynchronize(d->cs));
      d->cs_pending = false;
      d->forces_valid = false;
      fprintf(stderr, "Vpmg_solve: solve failed; the potential is not valid\n");
      continue;
    }
    nok++;
    pmg::set_faces(s, d->nx, d->ny, d->nz, d->u, d->gx, d->gy, d->gz);
    if (d->mirror)
      HIPCHECK(hipMemcpyAsync(th[b]->u, d->u, d->n * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  HIPCHECK(hipStreamSynchronize(s));
  for (int b = 0; b < n; b++) {
    PmgDev* d = ds[b];
    tcollect(d);
    d->stats[4] = now() - t0;
    d->forces_valid = false;
    if (good[b]) {
      d->stats[0] = iters[b];
      d->stats[1] = inner[b];
      d->stats[2] = d->ntot;
      d->stats[3] = relres[b];
      d->stats[7] = outer[b];
    }
    d->s = own[b];
  }
  return nok;
}

static bool solve_ready(Vpmg* thee) {
  if (!thee || !thee->filled) return false;
  const int meth = thee->pmgp->meth;
  if (meth != VSOL_Newton && meth != VSOL_MG) {
    fprintf(stderr, "Vpmg_solve: solver method %d is not supported\n", meth);
    return false;
  }
  return true;
}

extern "C" int Vpmg_solve(Vpmg* thee) {
  if (!solve_ready(thee)) return 0;
  int ok = 0;
  solve_group(&thee, 1, &ok);
  return ok;
}

// Several calculations solved together: consecutive runs of up to
// pmg::kMaxBatch calculations that same_batch accepts share every V-cycle
// launch.  No reference counterpart.
extern "C" int Vpmg_solveBatch(Vpmg** pmgs, int n, int* ok) {
  if (!pmgs || n < 0) return 0;
  int nok = 0;
  int b = 0;
  while (b < n) {
    if (!solve_ready(pmgs[b])) {
      if (ok) ok[b] = 0;
      b++;
      continue;
    }
    int e = b + 1;
    while (e < n && e - b < pmg::kMaxBatch && solve_ready(pmgs[e]) && same_batch(pmgs[b], pmgs[e]))
      e++;
    int res[pmg::kMaxBatch];
    nok += solve_group(pmgs + b, e - b, res);
    if (ok)
      for (int q = b; q < e; q++) ok[q] = res[q - b];
    b = e;
  }
  return nok;
}

// ================================================================ focusing

// focusFillBound (vpmg.c:1914-2414): Dirichlet faces from an old solution on
// the device (a coarser Vpmg's, or the potential map's); 0 when the new mesh
// is not contained in the old one.
static int focus_fill_bound_grid(Vpmg* thee, const pmg::OldGrid& og) {
  PmgDev* d = DEV(thee);
  Vpmgp* p = thee->pmgp;
  double xminN = p->xcent - ((double)(p->nx - 1) * p->hx) / 2.0;
  double xmaxN = p->xcent + ((double)(p->nx - 1) * p->hx) / 2.0;
  double yminN = p->ycent - ((double)(p->ny - 1) * p->hy) / 2.0;
  double ymaxN = p->ycent + ((double)(p->ny - 1) * p->hy) / 2.0;
  double zminN = p->zcent - ((double)(p->nz - 1) * p->hzed) / 2.0;
  double zmaxN = p->zcent + ((double)(p->nz - 1) * p->hzed) / 2.0;
  if (fabs(og.xmin - xminN) < VSMALL) xminN = og.xmin;
  if (fabs(og.xmax - xmaxN) < VSMALL) xmaxN = og.xmax;
  if (fabs(og.ymin - yminN) < VSMALL) yminN = og.ymin;
  if (fabs(og.ymax - ymaxN) < VSMALL) ymaxN = og.ymax;
  if (fabs(og.zmin - zminN) < VSMALL) zminN = og.zmin;
  if (fabs(og.zmax - zmaxN) < VSMALL) zmaxN = og.zmax;
  if ((xmaxN > og.xmax) || (ymaxN > og.ymax) || (zmaxN > og.zmax) || (og.xmin > xminN) ||
      (og.ymin > yminN) || (og.zmin > zminN))
    return 0;
  pmg::Geo g{};
  g.nx = p->nx;
  g.ny = p->ny;
  g.nz = p->nz;
  g.hx = p->hx;
  g.hy = p->hy;
  g.hz = p->hzed;
  int* bad = dalloc<int>(1);
  HIPCHECK(hipMemsetAsync(bad, 0, sizeof(int), d->s));
  pmg::focus_bound(d->s, g, xminN, yminN, zminN, xmaxN, ymaxN, zmaxN, og, d->gx, d->gy, d->gz, bad);
  int hbad = 0;
  HIPCHECK(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, d->s));
  HIPCHECK(hipStreamSynchronize(d->s));
  dfree(bad);
  return hbad ? 0 : 1;
}

static int focus_fill_bound(Vpmg* thee, Vpmg* old) {
  PmgDev* o = DEV(old);
  Vpmgp* q = old->pmgp;
  pmg::OldGrid og;
  og.nx = q->nx;
  og.ny = q->ny;
  og.nz = q->nz;
  og.hx = q->hx;
  og.hy = q->hy;
  og.hz = q->hzed;
  og.xmin = q->xcent - ((double)(q->nx - 1) * q->hx) / 2.0;
  og.xmax = q->xcent + ((double)(q->nx - 1) * q->hx) / 2.0;
  og.ymin = q->ycent - ((double)(q->ny - 1) * q->hy) / 2.0;
  og.ymax = q->ycent + ((double)(q->ny - 1) * q->hy) / 2.0;
  og.zmin = q->zcent - ((double)(q->nz - 1) * q->hzed) / 2.0;
  og.zmax = q->zcent + ((double)(q->nz - 1) * q->hzed) / 2.0;
  og.u = o->u;
  HIPCHECK(hipStreamSynchronize(o->s));
  return focus_fill_bound_grid(thee, og);
}

// Vpmg_setPart (vpmg.c:627-870), atom part on the host, grid part on the
// device when no partition side is shared (sequential focusing), else on the
// host with the reference's arithmetic.
static double part_axis_h(double x, double lo, double hi, double h, int bl, int bh) {
  double ok;
  if ((x < (hi - h / 2)) && (x > (lo + h / 2))) ok = 1.0;
  else if ((fabs(x - lo) < VPMGSMALL) && (bl == 0)) ok = 1.0;
  else if ((fabs(x - lo) < VPMGSMALL) && (bl == 1)) ok = 0.5;
  else if ((fabs(x - hi) < VPMGSMALL) && (bh == 0)) ok = 1.0;
  else if ((fabs(x - hi) < VPMGSMALL) && (bh == 1)) ok = 0.5;
  else if ((x > (hi + h / 2)) || (x < (lo - h / 2))) ok = 0.0;
  else {
    double x0 = std::max(x - h / 2, lo), x1 = std::min(x + h / 2, hi);
    ok = fabs(x1 - x0) / h;
    if (ok < 0.0 && fabs(ok) < VPMGSMALL) ok = 0.0;
    if (ok > 1.0 && fabs(ok - 1.0) < VPMGSMALL) ok = 1.0;
  }
  return ok;
}

static double atom_axis_h(double x, double lo, double hi, int bl, int bh) {
  if ((x < hi) && (x > lo)) return 1;
  if ((fabs(x - lo) < VPMGSMALL) && (bl == 0)) return 1;
  if ((fabs(x - lo) < VPMGSMALL) && (bl == 1)) return 0.5;
  if ((fabs(x - hi) < VPMGSMALL) && (bh == 0)) return 1;
  if ((fabs(x - hi) < VPMGSMALL) && (bh == 1)) return 0.5;
  return 0;
}

// bflags order of src/generic/vhal.h VAPBS_{RIGHT,FRONT,UP,LEFT,BACK,DOWN} = 0..5
// invert = 1 is extEnergy's use (vpmg.c:2416-2545): after the partition of
// [lo, hi] is set, every grid point and atom is flipped to 1 - [weight > 0]
// and then zeroed where it lies beyond the processor partition
// [partMin, partMax] on a side owned by another processor (bflags), so the
// old grid's energy is counted outside the new grid and inside this
// processor's share only.
static inline double beyond_mask(const double x[3], const double* partMin, const double* partMax,
                                 const int bflags[6]) {
  if (!partMin) return 1.0;
  double xv = 1, yv = 1, zv = 1;
  if (x[0] < partMin[0] && bflags[3] == 1) xv = 0;
  else if (x[0] > partMax[0] && bflags[0] == 1) xv = 0;
  if (x[1] < partMin[1] && bflags[4] == 1) yv = 0;
  else if (x[1] > partMax[1] && bflags[1] == 1) yv = 0;
  if (x[2] < partMin[2] && bflags[5] == 1) zv = 0;
  else if (x[2] > partMax[2] && bflags[2] == 1) zv = 0;
  return xv * yv * zv;
}

static void set_part_impl(Vpmg* thee, const double lo[3], const double hi[3], const int bflags[6],
                          int invert, const double* partMin = nullptr,
                          const double* partMax = nullptr) {
  PmgDev* d = DEV(thee);
  Vpmgp* p = thee->pmgp;
  const int bR = bflags[0], bF = bflags[1], bU = bflags[2], bL = bflags[3], bB = bflags[4],
            bD = bflags[5];
  Valist* al = thee->pbe->alist;
  for (int a = 0; a < d->natoms; a++) {
    const double* x = al->atoms[a].position;
    double v = atom_axis_h(x[0], lo[0], hi[0], bL, bR) * atom_axis_h(x[1], lo[1], hi[1], bB, bF) *
               atom_axis_h(x[2], lo[2], hi[2], bD, bU);
    d->hpart[a] = v;
    al->atoms[a].partID = v;
  }
  const double xm = p->xcent - 0.5 * p->hx * (p->nx - 1);
  const double ym = p->ycent - 0.5 * p->hy * (p->ny - 1);
  const double zm = p->zcent - 0.5 * p->hzed * (p->nz - 1);
  const bool shared = bR || bF || bU || bL || bB || bD;
  if (!shared) {
    pmg::Geo g{};
    g.nx = p->nx;
    g.ny = p->ny;
    g.nz = p->nz;
    g.hx = p->hx;
    g.hy = p->hy;
    g.hz = p->hzed;
    pmg::set_part(d->s, g, xm, ym, zm, lo, hi, invert, d->pvec);
  } else {
    std::vector<double> pv(d->n);
    for (int i = 0; i < p->nx; i++) {
      double xok = part_axis_h(i * p->hx + xm, lo[0], hi[0], p->hx, bL, bR);
      for (int j = 0; j < p->ny; j++) {
        double yok = part_axis_h(j * p->hy + ym, lo[1], hi[1], p->hy, bB, bF);
        for (int k = 0; k < p->nz; k++) {
          double zok = part_axis_h(k * p->hzed + zm, lo[2], hi[2], p->hzed, bD, bU);
          double v = xok * yok * zok;
          v = (fabs(v) < VPMGSMALL) ? 0.0 : v;
          if (invert) {
            const double pos[3] = {i * p->hx + xm, j * p->hy + ym, k * p->hzed + zm};
            v = (1 - (v > VSMALL ? 1.0 : v)) * beyond_mask(pos, partMin, partMax, bflags);
          }
          pv[((size_t)k * p->ny + j) * p->nx + i] = v;
        }
      }
    }
    HIPCHECK(hipMemcpyAsync(d->pvec, pv.data(), d->n * sizeof(double), hipMemcpyHostToDevice, d->s));
  }
  if (invert) {
    for (int a = 0; a < d->natoms; a++) {
      double v = d->hpart[a];
      if (v > VSMALL) v = 1.0;
      v = (1 - v) * beyond_mask(al->atoms[a].position, partMin, partMax, bflags);
      d->hpart[a] = v;
      al->atoms[a].partID = v;
    }
  }
  upload_part(d);
  if (d->mirror)
    HIPCHECK(hipMemcpyAsync(thee->pvec, d->pvec, d->n * sizeof(double), hipMemcpyDeviceToHost,
                            d->s));
  HIPCHECK(hipStreamSynchronize(d->s));
  d->forces_valid = false;
}

extern "C" void Vpmg_setPart(Vpmg* thee, double lowerCorner[3], double upperCorner[3],
                             int bflags[6]) {
  set_part_impl(thee, lowerCorner, upperCorner, bflags, 0);
}

extern "C" void Vpmg_unsetPart(Vpmg* thee) {
  PmgDev* d = DEV(thee);
  pmg::fill<double>(d->s, d->pvec, d->n, 1.0);
  Valist* al = thee->pbe->alist;
  for (int a = 0; a < d->natoms; a++) {
    d->hpart[a] = 1.0;
    al->atoms[a].partID = 1.0;
  }
  upload_part(d);
  if (d->mirror)
    for (size_t i = 0; i < d->n; i++) thee->pvec[i] = 1.0;
  HIPCHECK(hipStreamSynchronize(d->s));
  d->forces_valid = false;
}

// extEnergy (vpmg.c:2416-2545): energies of the old solve on the part of its
// domain outside the new one.
static void ext_energy(Vpmg* thee, Vpmg* old, PBEparm_calcEnergy, double* partMin,
                       double* partMax, int* bflags) {
  Vpmgp* p = thee->pmgp;
  double lo[3], hi[3];
  lo[0] = p->xcent - ((double)(p->nx - 1) * p->hx) / 2.0;
  hi[0] = p->xcent + ((double)(p->nx - 1) * p->hx) / 2.0;
  lo[1] = p->ycent - ((double)(p->ny - 1) * p->hy) / 2.0;
  hi[1] = p->ycent + ((double)(p->ny - 1) * p->hy) / 2.0;
  lo[2] = p->zcent - ((double)(p->nz - 1) * p->hzed) / 2.0;
  hi[2] = p->zcent + ((double)(p->nz - 1) * p->hzed) / 2.0;
  set_part_impl(old, lo, hi, bflags, 1, partMin, partMax);
  thee->extQmEnergy = Vpmg_qmEnergy(old, 1);
  thee->extQfEnergy = Vpmg_qfEnergy(old, 1);
  thee->extDiEnergy = Vpmg_dielEnergy(old, 1);
  Vpmg_unsetPart(old);
}

// ================================================================ energies

// Device buffers an energy reduction works in: partial sums (kParts), the
// finished sum, the per-atom partials of the point-charge energy and the chop
// counter.  The public energies use the object's buffers (and a pool block
// for the counter); Vpmg_energyNoAlloc uses the library's static buffer.
struct EnergyBufs {
  double* part;
  double* out;
  double* apart;
  int* nchop;  // nullptr: a pool block per call
};

EnergyBufs object_bufs(PmgDev* d) { return EnergyBufs{d->part, d->scal, d->atmp, nullptr}; }

static double qf_energy(Vpmg* thee, int extFlag, const EnergyBufs& b) {
  PmgDev* d = DEV(thee);
  Vpmgp* p = thee->pmgp;
  double e;
  // Vpmg_qfEnergy (vpmg.c:1687-1702): the volume sum for a charge map or spl2
  if (thee->useChargeMap || thee->chargeMeth == VCM_BSPL2) {  // vpmg.c:1861-1891
    pmg::energy_qf_volume(d->s, d->n, d->pvec, d->u, d->charge, b.part, b.out);
    e = read_scalar(d, b.out);
    e = e * p->hx * p->hy * p->hzed / thee->pbe->zmagic;
    if (extFlag == 1) e += thee->extQfEnergy;
  } else {  // Vpmg_qfEnergyPoint (vpmg.c:1704-1789)
    pmg::energy_qf_point(d->s, d->geo, atoms_dev(d), d->u, b.apart, b.out);
    e = read_scalar(d, b.out);
    if (extFlag) e += thee->extQfEnergy;
  }
  return e;
}

static double diel_energy(Vpmg* thee, int extFlag, const EnergyBufs& b) {
  PmgDev* d = DEV(thee);
  Vpmgp* p = thee->pmgp;
  if (!thee->filled) return 0.0;
  pmg::energy_diel(d->s, d->geo, d->pvec, d->u, d->epsx, d->epsy, d->epsz, b.part, b.out);
  double e = read_scalar(d, b.out);
  e = 0.5 * e * p->hx * p->hy * p->hzed;
  e = e / thee->pbe->zmagic;
  if (extFlag == 1) e += thee->extDiEnergy;
  return e;
}

static double qm_energy(Vpmg* thee, int extFlag, const EnergyBufs& b) {
  PmgDev* d = DEV(thee);
  Vpmgp* p = thee->pmgp;
  Vpbe* pbe = thee->pbe;
  if (pbe->zkappa2 < VSMALL) return 0.0;
  const double zks2 = 0.5 * pbe->zkappa2 / pbe->bulkIonicStrength;
  pmg::IonBulk I{};
  I.nion = pbe->numIon;
  for (int i = 0; i < I.nion; i++) {
    I.q[i] = pbe->ionQ[i];
    I.c[i] = pbe->ionConc[i];
  }
  int* nchop = b.nchop ? b.nchop : dalloc<int>(1);
  pmg::energy_qm(d->s, d->n, p->nonlin ? 1 : 0, d->pvec, d->kappa, d->u, zks2, pbe->zkappa2, I,
                 b.part, b.out, nchop);
  double e = read_scalar(d, b.out);
  int hchop = 0;
  HIPCHECK(hipMemcpyAsync(&hchop, nchop, sizeof(int), hipMemcpyDeviceToHost, d->s));
  HIPCHECK(hipStreamSynchronize(d->s));
  if (!b.nchop) dfree(nchop);
  if (p->nonlin && hchop > 0) {
    // the reference aborts here (vpmg.c:1468-1472)
    fprintf(stderr, "Vpmg_qmEnergy: chopped EXP %d times; large potential values\n", hchop);
    return NAN;
  }
  if (!p->nonlin) e = 0.5 * e;
  e = e * p->hx * p->hy * p->hzed;
  e = e / pbe->zmagic;
  if (extFlag == 1) e += thee->extQmEnergy;
  return e;
}

// Vpmg_energy (vpmg.c:1248-1277)
static double total_energy(Vpmg* thee, int extFlag, const EnergyBufs& b) {
  if ((thee->pmgp->nonlin) && (thee->pbe->bulkIonicStrength > 0.)) {
    double qm = qm_energy(thee, extFlag, b);
    double qf = qf_energy(thee, extFlag, b);
    double di = diel_energy(thee, extFlag, b);
    return qf - di - qm;
  }
  return 0.5 * qf_energy(thee, extFlag, b);
}

extern "C" double Vpmg_qfEnergy(Vpmg* thee, int extFlag) {
  return qf_energy(thee, extFlag, object_bufs(DEV(thee)));
}

extern "C" double Vpmg_dielEnergy(Vpmg* thee, int extFlag) {
  return diel_energy(thee, extFlag, object_bufs(DEV(thee)));
}

extern "C" double Vpmg_qmEnergy(Vpmg* thee, int extFlag) {
  return qm_energy(thee, extFlag, object_bufs(DEV(thee)));
}

extern "C" double Vpmg_energy(Vpmg* thee, int extFlag) {
  return total_energy(thee, extFlag, object_bufs(DEV(thee)));
}

// Vpmg_energy's value computed without any allocation: the reductions work in
// the library's static device buffer (pmg::scratch_address) and the result is
// read through the object's pinned scalars.  The point-charge energy (chgm
// spl0 without a charge map) keeps one partial per atom there, so it is
// refused above pmg::kScratchAtoms atoms.  Calls on one device share the
// buffer: calls that overlap on one device race.
extern "C" int Vpmg_energyNoAlloc(Vpmg* thee, int extFlag, double* energy) {
  if (!thee || !thee->dev || !energy) {
    fprintf(stderr, "Vpmg_energyNoAlloc: null argument\n");
    return 0;
  }
  PmgDev* d = DEV(thee);
  const bool point = !(thee->useChargeMap || thee->chargeMeth == VCM_BSPL2);
  if (point && (size_t)d->natoms > pmg::kScratchAtoms) {
    fprintf(stderr,
            "Vpmg_energyNoAlloc: the point-charge energy of %d atoms needs one partial per "
            "atom; the static buffer holds %zu (use Vpmg_energy)\n",
            d->natoms, pmg::kScratchAtoms);
    return 0;
  }
  double* sc = pmg::scratch_address();
  if (!sc) return 0;
  const EnergyBufs b{sc, sc + pmg::kScratchParts, sc + pmg::kScratchAtomOff,
                     reinterpret_cast<int*>(sc + pmg::kScratchParts + 8)};
  *energy = total_energy(thee, extFlag, b);
  return std::isfinite(*energy) ? 1 : 0;
}

// Vpmg_qfAtomEnergy (vpmg.c:1791-1859): trilinear u at the atom
extern "C" double Vpmg_qfAtomEnergy(Vpmg* thee, Vatom* atom) {
  PmgDev* d = DEV(thee);
  Vpmgp* p = thee->pmgp;
  if (!thee->u) return 0.0;
  if (!d->mirror)
    HIPCHECK(hipMemcpy(thee->u, d->u, d->n * sizeof(double), hipMemcpyDeviceToHost));
  const int nx = p->nx, ny = p->ny, nz = p->nz;
  const double xmin = p->xmin, ymin = p->ymin, zmin = p->zmin;
  const double fi = (atom->position[0] - xmin) / p->hx, fj = (atom->position[1] - ymin) / p->hy,
               fk = (atom->position[2] - zmin) / p->hzed;
  const int ihi = (int)ceil(fi), ilo = (int)floor(fi), jhi = (int)ceil(fj), jlo = (int)floor(fj),
            khi = (int)ceil(fk), klo = (int)floor(fk);
  if (!(atom->partID > 0)) return 0.0;
  if (!((ihi < nx) && (jhi < ny) && (khi < nz) && (ilo >= 0) && (jlo >= 0) && (klo >= 0))) return 0.0;
  const double dx = fi - ilo, dy = fj - jlo, dz = fk - klo;
  auto U = [&](int i, int j, int k) { return thee->u[((size_t)k * ny + j) * nx + i]; };
  const double uval = dx * dy * dz * U(ihi, jhi, khi) + dx * (1.0 - dy) * dz * U(ihi, jlo, khi) +
                      dx * dy * (1.0 - dz) * U(ihi, jhi, klo) +
                      dx * (1.0 - dy) * (1.0 - dz) * U(ihi, jlo, klo) +
                      (1.0 - dx) * dy * dz * U(ilo, jhi, khi) +
                      (1.0 - dx) * (1.0 - dy) * dz * U(ilo, jlo, khi) +
                      (1.0 - dx) * dy * (1.0 - dz) * U(ilo, jhi, klo) +
                      (1.0 - dx) * (1.0 - dy) * (1.0 - dz) * U(ilo, jlo, klo);
  return uval * atom->charge * atom->partID;
}

// ================================================================== forces

namespace {
// The three force kinds for every atom, with the charge method chgm and the
// surface method srfm the caller asks for (Vpmg_qfForce / Vpmg_ibForce /
// Vpmg_dbForce take them as arguments); cached until the next solve or
// partition change.
void ensure_forces(Vpmg* thee, int chgm, int srfm) {
  PmgDev* d = DEV(thee);
  if (d->forces_valid && d->fchgm == chgm && d->fsrfm == srfm) return;
  const int order = srfm_order((Vsurf_Meth)srfm);
  ensure_spline_coef(thee, order);
  const double* coef = order == 2 ? nullptr : d->splcoef;
  const int na = d->natoms;
  if (!d->fqf) {
    d->fqf = dalloc<double>(3 * (size_t)na);
    d->fib = dalloc<double>(3 * (size_t)na);
    d->fdb = dalloc<double>(3 * (size_t)na);
    d->frc = dalloc<int>(na);
  }
  Vpbe* pbe = thee->pbe;
  const pmg::AtomsDev A = atoms_dev(d);
  pmg::force_qf(d->s, d->geo, A, chgm, d->u, d->fqf);
  pmg::IonBulk I{};
  I.nion = pbe->numIon;
  for (int i = 0; i < I.nion; i++) {
    I.q[i] = pbe->ionQ[i];
    I.c[i] = pbe->ionConc[i];
  }
  pmg::force_ib(d->s, d->geo, A, thee->splineWin, pbe->maxIonRadius, pbe->zkappa2,
                1.0 / pbe->zmagic, thee->pmgp->nonlin ? 1 : 0, pbe->bulkIonicStrength, I,
                d->kappa, d->u, d->fib, order, coef);
  pmg::force_db(d->s, d->geo, A, thee->splineWin, pbe->solventRadius, pbe->soluteDiel,
                pbe->solventDiel, 1.0 / pbe->zmagic, d->epsx, d->epsy, d->epsz, d->u, d->fdb,
                d->frc, order, coef);
  d->hqf.resize(3 * (size_t)na);
  d->hib.resize(3 * (size_t)na);
  d->hdb.resize(3 * (size_t)na);
  d->hrc.resize(na);
  HIPCHECK(hipMemcpyAsync(d->hqf.data(), d->fqf, 3 * na * sizeof(double), hipMemcpyDeviceToHost, d->s));
  HIPCHECK(hipMemcpyAsync(d->hib.data(), d->fib, 3 * na * sizeof(double), hipMemcpyDeviceToHost, d->s));
  HIPCHECK(hipMemcpyAsync(d->hdb.data(), d->fdb, 3 * na * sizeof(double), hipMemcpyDeviceToHost, d->s));
  HIPCHECK(hipMemcpyAsync(d->hrc.data(), d->frc, na * sizeof(int), hipMemcpyDeviceToHost, d->s));
  HIPCHECK(hipStreamSynchronize(d->s));
  d->forces_valid = true;
  d->fchgm = chgm;
  d->fsrfm = srfm;
}
}  // namespace

extern "C" int Vpmg_qfForce(Vpmg* thee, double* force, int atomID, Vchrg_Meth chgm) {
  force[0] = force[1] = force[2] = 0.0;
  if (chgm != VCM_TRIL && chgm != VCM_BSPL2 && chgm != VCM_BSPL4) return 0;
  PmgDev* d = DEV(thee);
  if (atomID < 0 || atomID >= d->natoms) return 0;
  ensure_forces(thee, chgm, thee->surfMeth);
  for (int k = 0; k < 3; k++) force[k] = d->hqf[3 * atomID + k];
  return 1;
}

static bool spline_srfm(Vsurf_Meth s) {
  return s == VSM_SPLINE || s == VSM_SPLINE3 || s == VSM_SPLINE4;
}

extern "C" int Vpmg_ibForce(Vpmg* thee, double* force, int atomID, Vsurf_Meth srfm) {
  force[0] = force[1] = force[2] = 0.0;
  if (!spline_srfm(srfm)) return 0;  // vpmg.c:5471-5477
  PmgDev* d = DEV(thee);
  if (atomID < 0 || atomID >= d->natoms) return 0;
  ensure_forces(thee, thee->chargeMeth, srfm);
  for (int k = 0; k < 3; k++) force[k] = d->hib[3 * atomID + k];
  return 1;
}

extern "C" int Vpmg_dbForce(Vpmg* thee, double* dbForce, int atomID, Vsurf_Meth srfm) {
  dbForce[0] = dbForce[1] = dbForce[2] = 0.0;
  if (!thee->filled) return 0;
  if (!spline_srfm(srfm)) return 0;  // vpmg.c:5638-5644
  PmgDev* d = DEV(thee);
  if (atomID < 0 || atomID >= d->natoms) return 0;
  ensure_forces(thee, thee->chargeMeth, srfm);
  for (int k = 0; k < 3; k++) dbForce[k] = d->hdb[3 * atomID + k];
  return d->hrc[atomID];
}

// Vpmg_force (vpmg.c:5417-5438)
extern "C" int Vpmg_force(Vpmg* thee, double* force, int atomID, Vsurf_Meth srfm,
                          Vchrg_Meth chgm) {
  double qf[3] = {0, 0, 0}, db[3] = {0, 0, 0}, ib[3] = {0, 0, 0};
  int rc = 1;
  rc = rc && Vpmg_dbForce(thee, db, atomID, srfm);
  rc = rc && Vpmg_ibForce(thee, ib, atomID, srfm);
  rc = rc && Vpmg_qfForce(thee, qf, atomID, chgm);
  for (int k = 0; k < 3; k++) force[k] = qf[k] + db[k] + ib[k];
  return rc;
}

// ================================================================== arrays

extern "C" int Vpmg_copyArray(Vpmg* thee, int which, double* out, size_t n) {
  PmgDev* d = DEV(thee);
  const double* src = nullptr;
  size_t need = d->n;
  switch (which) {
    case 0: src = d->u; break;
    case 1: src = d->epsx; break;
    case 2: src = d->epsy; break;
    case 3: src = d->epsz; break;
    case 4: src = d->kappa; break;
    case 5: src = d->charge; break;
    case 6: src = d->pvec; break;
    case 7: src = d->gx; need = 2 * (size_t)d->ny * d->nz; break;
    case 8: src = d->gy; need = 2 * (size_t)d->nx * d->nz; break;
    case 9: src = d->gz; need = 2 * (size_t)d->nx * d->ny; break;
    default: return 0;
  }
  if (n < need) return 0;
  HIPCHECK(hipMemcpyAsync(out, src, need * sizeof(double), hipMemcpyDeviceToHost, d->s));
  HIPCHECK(hipStreamSynchronize(d->s));
  return 1;
}

extern "C" int Vpmg_fillArray(Vpmg* thee, double* vec, Vdata_Type type, double, Vhal_PBEType,
                              void*) {
  PmgDev* d = DEV(thee);
  if (!thee->filled) return 0;
  switch (type) {
    case VDT_CHARGE: {
      if (!Vpmg_copyArray(thee, 5, vec, d->n)) return 0;
      const double zm = thee->pbe->zmagic;
      for (size_t i = 0; i < d->n; i++) vec[i] = vec[i] / zm;
      return 1;
    }
    case VDT_DIELX: return Vpmg_copyArray(thee, 1, vec, d->n);
    case VDT_DIELY: return Vpmg_copyArray(thee, 2, vec, d->n);
    case VDT_DIELZ: return Vpmg_copyArray(thee, 3, vec, d->n);
    case VDT_KAPPA: return Vpmg_copyArray(thee, 4, vec, d->n);
    case VDT_POT: return Vpmg_copyArray(thee, 0, vec, d->n);
    default:
      fprintf(stderr, "Vpmg_fillArray: data type %d is not supported\n", (int)type);
      return 0;
  }
}

extern "C" unsigned long int Vpmg_memChk(Vpmg* thee) {
  if (!thee || !thee->dev) return 0;
  PmgDev* d = DEV(thee);
  size_t bytes = 10 * d->n * sizeof(double);
  for (int l = 0; l < d->ntot; l++)
    bytes += (size_t)(d->L[l].st == 7 ? 8 : 18) * d->L[l].n * sizeof(double);
  return (unsigned long)bytes;
}

// ============================================================= extensions

extern "C" int Vpmg_setPrecision(Vpmg* thee, int mode, double tol) {
  if (!thee || mode < 0 || mode > 2) return 0;
  PmgDev* d = DEV(thee);
  d->precision = mode ? 1 : 0;
  d->fast_boundary = mode == 2;
  d->mixed_tol = tol;
  return 1;
}

void pmgint::set_side_stream(Vpmg* pmg, bool on) {
  if (pmg) DEV(pmg)->side = on ? 1 : 0;
}

// Test hook: hierarchies this object allocates from now on have their
// interiors filled with NaN before the boundary layers are cleared, so a
// kernel that read an interior entry before writing it would spread NaN.
extern "C" int pmgtest_poison_hierarchy(void* pmg, int on) {
  Vpmg* thee = (Vpmg*)pmg;
  if (!thee || !thee->dev) return 0;
  DEV(thee)->poison = on != 0;
  return 1;
}

extern "C" int Vpmg_setHostMirror(Vpmg* thee, int on) {
  if (!thee) return 0;
  DEV(thee)->mirror = on ? 1 : 0;
  return 1;
}

extern "C" int Vpmg_getStats(Vpmg* thee, double* out, int n) {
  if (!thee) return 0;
  PmgDev* d = DEV(thee);
  int m = std::min(n, 8);
  for (int i = 0; i < m; i++) out[i] = d->stats[i];
  return m;
}

extern "C" int Vpmg_setKernelTiming(Vpmg* thee, int on) {
  if (!thee) return 0;
  PmgDev* d = DEV(thee);
  d->timing = on ? 1 : 0;
  for (double& v : d->kt) v = 0.0;
  d->marks.clear();
  d->ev_next = 0;
  return 1;
}

extern "C" int Vpmg_getKernelTiming(Vpmg* thee, double* out, int n) {
  if (!thee) return 0;
  PmgDev* d = DEV(thee);
  int m = std::min(n, 8);
  for (int i = 0; i < m; i++) out[i] = d->kt[i];
  return m;
}

extern "C" int apbs_pmg_synchronize(void) { return hipDeviceSynchronize() == hipSuccess; }

extern "C" const char* apbs_pmg_version(void) { return "apbs_pmg 0.1 (gfx950)"; }
