// server.cpp -- continuous-batching PB solve service (apbs_pmg_server.h).
//
// One worker thread per slot.  A worker takes the oldest queued request as
// soon as its footprint fits in the unused HBM budget, runs the mg-manual
// calculation through the public Vpmg entry points on the Vpmg's own stream,
// posts the result and frees the footprint.  The solves of calculations of
// one shape are batched across requests (SolveGate), so their V-cycles share
// kernel launches; everything else of a request overlaps with the others'
// work on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/apbs_pmg_server.h"
#include "pool.h"
#include "vpmg_internal.h"

namespace {

constexpr double kVPI = 3.14159265358979323846;

struct Job {
  long id;
  PmgRequest req;
  std::vector<double> xyz, rad, chg;
  size_t footprint;
  double t_submit;
};

int nlev_of(const PmgRequest& r) {
  MGparm* m = MGparm_ctor(MCT_MANUAL);
  for (int j = 0; j < 3; j++) {
    m->dime[j] = r.dime[j];
    m->grid[j] = r.grid[j];
    m->center[j] = r.center[j];
  }
  m->setdime = m->setgrid = m->setgcent = 1;
  m->nlev = 4;
  m->chgm = (Vchrg_Meth)r.chgm;
  m->setchgm = 1;
  int nlev = MGparm_check(m) ? m->nlev : -1;
  MGparm_dtor(&m);
  return nlev;
}

size_t footprint_of(const PmgRequest& r) {
  const int nlev = nlev_of(r);
  if (nlev < 1) return 0;
  int nref = 0;
  if ((r.srfm == VSM_MOL || r.srfm == VSM_MOLSMOOTH) && r.srad > 1e-9) {
    double rmax = 0, imax = 0;
    for (int a = 0; a < r.natoms; a++) rmax = std::max(rmax, r.radius[a]);
    for (int i = 0; i < r.nion; i++) imax = std::max(imax, r.ionr[i]);
    const double mr = rmax + std::max(r.srad, imax);
    nref = (int)std::ceil(4.0 * kVPI * mr * mr * r.sdens);
  }
  return pmgint::footprint_bytes(r.dime[0], r.dime[1], r.dime[2], nlev, r.npbe != 0,
                                 r.precision != 0, r.natoms, nref);
}

// Cross-request batching of the solves.  A worker whose calculation is
// filled waits here; a waiting worker leads a batch when fewer than
// max_running batches run: it takes the oldest waiting calculations of its own
// shape (grid, level count, method, precision), at most max_batch, and solves
// them with Vpmg_solveBatch, which shares every V-cycle launch among them.
// Calculations of one batch run in lockstep on one stream; batches of
// different shapes and the other workers' setup overlap on the device.
struct SolveGate {
  struct Key {
    int nx, ny, nz, nlev, meth, precision;
    bool operator==(const Key& o) const {
      return nx == o.nx && ny == o.ny && nz == o.nz && nlev == o.nlev && meth == o.meth &&
             precision == o.precision;
    }
  };
  struct Entry {
    Vpmg* pmg;
    Key key;
    int ok = 0;
    bool taken = false, done = false;
  };
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Entry*> waiting;
  int running = 0;
  int max_batch = APBS_PMG_MAXBATCH;
  int max_running = 1;  // set to the slot count by PmgServer_ctor
  long batches = 0, batched = 0;

  int solve(Vpmg* pmg, int precision) {
    Entry e{pmg, Key{pmg->pmgp->nx, pmg->pmgp->ny, pmg->pmgp->nz, pmg->pmgp->nlev,
                     (int)pmg->pmgp->meth, precision}};
    std::unique_lock<std::mutex> lk(mu);
    waiting.push_back(&e);
    cv.notify_all();
    for (;;) {
      if (e.done) return e.ok;
      if (!e.taken && running < max_running) {
        std::vector<Entry*> batch;
        for (auto it = waiting.begin(); it != waiting.end() && (int)batch.size() < max_batch;) {
          if ((*it)->key == e.key) {
            (*it)->taken = true;
            batch.push_back(*it);
            it = waiting.erase(it);
          } else {
            ++it;
          }
        }
        running++;
        batches++;
        batched += (long)batch.size();
        lk.unlock();
        std::vector<Vpmg*> list;
        for (Entry* x : batch) list.push_back(x->pmg);
        std::vector<int> ok(batch.size(), 0);
        Vpmg_solveBatch(list.data(), (int)list.size(), ok.data());
        lk.lock();
        running--;
        for (size_t q = 0; q < batch.size(); q++) {
          batch[q]->ok = ok[q];
          batch[q]->done = true;
        }
        cv.notify_all();
        continue;
      }
      cv.wait(lk);
    }
  }
};

// One mg-manual calculation (routines.c initMG / solveMG / setPartMG /
// energyMG) through the public API; the solve goes through the gate when
// batching is on.
void run_job(Job& j, PmgResult& res, SolveGate* gate) {
  PmgRequest& r = j.req;
  res.status = 0;
  Valist* al = Valist_ctor();
  Vpbe* pbe = nullptr;
  MGparm* m = nullptr;
  Vpmgp* pmgp = nullptr;
  Vpmg* pmg = nullptr;
  do {
    if (!Valist_setAtoms(al, r.natoms, j.xyz.data(), j.rad.data(), j.chg.data())) break;
    double ic[APBS_PMG_MAXION + 1] = {0}, ir[APBS_PMG_MAXION + 1] = {0},
           iq[APBS_PMG_MAXION + 1] = {0};
    for (int i = 0; i < r.nion; i++) {
      ic[i] = r.ionc[i];
      ir[i] = r.ionr[i];
      iq[i] = r.ionq[i];
    }
    const double sparm = r.srfm == VSM_SPLINE ? r.swin : r.srad;
    pbe = Vpbe_ctor(al, r.nion, ic, ir, iq, r.temp, r.pdie, r.sdie, sparm, 0, r.sdens, 0, 0, 0,
                    0);
    if (!pbe) break;
    m = MGparm_ctor(MCT_MANUAL);
    for (int k = 0; k < 3; k++) {
      m->dime[k] = r.dime[k];
      m->grid[k] = r.grid[k];
      m->center[k] = r.center[k];
    }
    m->setdime = m->setgrid = m->setgcent = 1;
    m->nlev = 4;
    m->chgm = (Vchrg_Meth)r.chgm;
    m->setchgm = 1;
    m->nonlintype = r.npbe ? NONLIN_NPBE : NONLIN_LPBE;
    m->method = r.npbe ? VSOL_Newton : VSOL_MG;
    if (r.etol > 0) {
      m->etol = r.etol;
      m->setetol = 1;
    }
    if (!MGparm_check(m)) break;
    pmgp = Vpmgp_ctor(m);
    if (!pmgp) break;
    pmgp->bcfl = (Vbcfl)r.bcfl;
    pmgp->xcent = m->center[0];
    pmgp->ycent = m->center[1];
    pmgp->zcent = m->center[2];
    pmg = Vpmg_ctor(pmgp, pbe, 0, nullptr, m, PCE_TOTAL);
    if (!pmg) break;
    if (r.precision) Vpmg_setPrecision(pmg, r.precision, r.mixed_tol);
    Vpmg_setHostMirror(pmg, 0);
    pmgint::set_side_stream(pmg, false);
    if (!Vpmg_fillco(pmg, (Vsurf_Meth)r.srfm, r.swin, (Vchrg_Meth)r.chgm, 0, nullptr, 0, nullptr,
                     0, nullptr, 0, nullptr, 0, nullptr, 0, nullptr))
      break;
    // modes 1 and 2 solve alike (2 differs in fillco only)
    if (!(gate ? gate->solve(pmg, r.precision ? 1 : 0) : Vpmg_solve(pmg))) break;
    double lo[3], hi[3];
    for (int k = 0; k < 3; k++) {
      lo[k] = m->center[k] - 0.5 * m->glen[k];
      hi[k] = m->center[k] + 0.5 * m->glen[k];
    }
    Vpmg_setPart(pmg, lo, hi, m->partDisjOwnSide);
    res.energy_kT = Vpmg_energy(pmg, 1);
    double st[8];
    Vpmg_getStats(pmg, st, 8);
    res.iters = (int)st[0];
    res.inner = (int)st[1];
    res.relres = st[3];
    if (r.potential) {
      const size_t n = (size_t)r.dime[0] * r.dime[1] * r.dime[2];
      if (!Vpmg_copyArray(pmg, 0, r.potential, n)) break;
    }
    res.status = 1;
  } while (false);
  if (pmg) Vpmg_dtor(&pmg);
  if (pmgp) Vpmgp_dtor(&pmgp);
  if (m) MGparm_dtor(&m);
  if (pbe) Vpbe_dtor(&pbe);
  Valist_dtor(&al);
}

}  // namespace

struct PmgServer {
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Job*> queue;
  std::vector<PmgResult> done;
  std::vector<std::thread> workers;
  size_t budget = 0, admitted = 0, peak_admitted = 0;
  long next_id = 0, finished = 0;
  int running = 0, slots = 0;
  bool stop = false;
  bool batching = true;
  SolveGate gate;

  void work() {
    for (;;) {
      Job* j = nullptr;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_work.wait(lk, [&] {
          return stop || (!queue.empty() && admitted + queue.front()->footprint <= budget);
        });
        if (queue.empty() && stop) return;
        if (!queue.empty() && admitted + queue.front()->footprint <= budget) {
          j = queue.front();
          queue.pop_front();
          admitted += j->footprint;
          peak_admitted = std::max(peak_admitted, admitted);
          running++;
        } else {
          continue;
        }
      }
      PmgResult res{};
      res.id = j->id;
      res.tag = j->req.tag;
      res.footprint = j->footprint;
      res.t_submit = j->t_submit;
      res.t_start = apbs_pmg_now();
      run_job(*j, res, batching ? &gate : nullptr);
      res.t_done = apbs_pmg_now();
      {
        std::lock_guard<std::mutex> lk(mu);
        admitted -= j->footprint;
        running--;
        finished++;
        done.push_back(res);
      }
      delete j;
      cv_work.notify_all();
      cv_done.notify_all();
    }
  }
};

extern "C" double apbs_pmg_now(void) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

extern "C" void apbs_pmg_poolStats(size_t* inuse, size_t* peak, size_t* cached, int reset) {
  pmgpool::stats(inuse, peak, cached);
  if (reset) pmgpool::reset_peak();
}

extern "C" size_t PmgServer_footprint(const PmgRequest* req) {
  return req ? footprint_of(*req) : 0;
}

extern "C" PmgServer* PmgServer_ctor(int slots, size_t hbm_budget) {
  if (slots < 1) return nullptr;
  PmgServer* s = new PmgServer();
  s->budget = hbm_budget;
  s->slots = slots;
  s->gate.max_running = slots;  // no limit: a ready solve never waits for a batch
  for (int i = 0; i < slots; i++) s->workers.emplace_back([s] { s->work(); });
  return s;
}

extern "C" long PmgServer_submit(PmgServer* s, const PmgRequest* req) {
  if (!s || !req || req->natoms < 0 || req->nion < 0 || req->nion > APBS_PMG_MAXION ||
      req->precision < 0 || req->precision > 2)
    return -1;
  for (int k = 0; k < 3; k++)
    if (req->dime[k] < 5) return -1;
  Job* j = new Job();
  j->req = *req;
  j->xyz.assign(req->xyz, req->xyz + 3 * (size_t)req->natoms);
  j->rad.assign(req->radius, req->radius + req->natoms);
  j->chg.assign(req->charge, req->charge + req->natoms);
  j->req.xyz = j->xyz.data();
  j->req.radius = j->rad.data();
  j->req.charge = j->chg.data();
  j->footprint = footprint_of(j->req);
  if (j->footprint == 0 || j->footprint > s->budget) {
    delete j;
    return -1;
  }
  j->t_submit = apbs_pmg_now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    j->id = s->next_id++;
    s->queue.push_back(j);
  }
  s->cv_work.notify_all();
  return j->id;
}

extern "C" int PmgServer_poll(PmgServer* s, PmgResult* out, int maxn) {
  if (!s || !out || maxn <= 0) return 0;
  std::lock_guard<std::mutex> lk(s->mu);
  const int n = std::min<int>(maxn, (int)s->done.size());
  for (int i = 0; i < n; i++) out[i] = s->done[i];
  s->done.erase(s->done.begin(), s->done.begin() + n);
  return n;
}

extern "C" void PmgServer_drain(PmgServer* s) {
  if (!s) return;
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] { return s->queue.empty() && s->running == 0; });
}

extern "C" void PmgServer_stats(PmgServer* s, double out[8]) {
  if (!s || !out) return;
  size_t inuse = 0, peak = 0, cached = 0;
  pmgpool::stats(&inuse, &peak, &cached);
  std::lock_guard<std::mutex> lk(s->mu);
  out[0] = (double)s->next_id;
  out[1] = (double)s->finished;
  out[2] = (double)s->running;
  out[3] = (double)s->queue.size();
  out[4] = (double)s->admitted;
  out[5] = (double)s->peak_admitted;
  out[6] = (double)peak;
  out[7] = (double)s->slots;
}

extern "C" int PmgServer_setBatching(PmgServer* s, int max_batch, int max_running) {
  if (!s || max_batch < 1 || max_batch > APBS_PMG_MAXBATCH || max_running < 1) return 0;
  std::lock_guard<std::mutex> lk(s->gate.mu);
  s->batching = max_batch > 1;
  s->gate.max_batch = max_batch;
  s->gate.max_running = max_running;
  return 1;
}

extern "C" void PmgServer_batchStats(PmgServer* s, long* batches, long* batched) {
  if (!s) return;
  std::lock_guard<std::mutex> lk(s->gate.mu);
  if (batches) *batches = s->gate.batches;
  if (batched) *batched = s->gate.batched;
}

extern "C" void PmgServer_dtor(PmgServer** sp) {
  if (!sp || !*sp) return;
  PmgServer* s = *sp;
  PmgServer_drain(s);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto& t : s->workers) t.join();
  delete s;
  *sp = nullptr;
}
