/* apbs_pmg_server.h -- continuous-batching service for PB solves on one GPU.
 *
 * No reference counterpart: APBS runs one calculation at a time
 * (src/routines.c:1314-1941).  The service runs many of those calculations at
 * once.  Each request is one mg-manual calculation (Vpbe_ctor, Vpmgp_ctor,
 * Vpmg_ctor, Vpmg_fillco, Vpmg_solve, Vpmg_setPart, Vpmg_energy), run through
 * the same entry points as apbs_pmg.h on its own HIP stream.  Requests are
 * admitted first-in first-out, whenever a slot is free and the estimated HBM
 * footprint of their multigrid hierarchy fits in the remaining budget.
 * Requests that arrive while others run join as soon as a slot and memory
 * free up, and their GPU work overlaps on the device.  The solves of requests
 * of one grid, level count, method and precision that are ready at the same
 * time are batched (Vpmg_solveBatch): their V-cycles share every kernel
 * launch, and each result is bitwise the request's result alone.  Device
 * memory is recycled through the library's caching pool, so a finished
 * request frees nothing back to the driver.
 */
#ifndef APBS_PMG_SERVER_H
#define APBS_PMG_SERVER_H

#include <stddef.h>

#include "apbs_pmg.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct PmgRequest {
    int natoms;
    const double *xyz;      /* 3*natoms, copied at submit */
    const double *radius;   /* natoms */
    const double *charge;   /* natoms */
    int dime[3];
    double grid[3];         /* spacing (A) */
    double center[3];
    int npbe;               /* 0 linearized (VSOL_MG), 1 nonlinear (VSOL_Newton) */
    int nion;
    double ionq[APBS_PMG_MAXION], ionc[APBS_PMG_MAXION], ionr[APBS_PMG_MAXION];
    double pdie, sdie, srad, swin, sdens, temp;
    int srfm, chgm, bcfl;   /* Vsurf_Meth, Vchrg_Meth, Vbcfl */
    double etol;            /* <= 0: 1e-6 */
    int precision;          /* 0 FP64, 1 mixed, 2 mixed + FP32 MDH terms (Vpmg_setPrecision) */
    double mixed_tol;       /* <= 0: etol */
    double *potential;      /* optional host buffer of dime[0]*dime[1]*dime[2] doubles,
                               filled before the result is posted */
    long tag;               /* returned unchanged in the result */
} PmgRequest;

typedef struct PmgResult {
    long id, tag;
    int status;             /* 1 solved, 0 failed */
    double energy_kT;       /* Vpmg_energy(pmg, 1) */
    int iters, inner;
    double relres;
    double t_submit, t_start, t_done;   /* apbs_pmg_now() seconds */
    size_t footprint;       /* admitted HBM estimate (bytes) */
} PmgResult;

typedef struct PmgServer PmgServer;

/* slots: requests in flight at once (one host thread and one stream each);
 * hbm_budget: bytes of device memory the admitted requests may hold. */
PmgServer *PmgServer_ctor(int slots, size_t hbm_budget);
/* Queue a request; returns its id (>= 0), or -1 if it can never be admitted
 * (footprint above the budget) or is malformed. */
long PmgServer_submit(PmgServer *thee, const PmgRequest *req);
/* Move up to maxn finished results (in completion order) to out; returns the
 * number moved.  Does not block. */
int PmgServer_poll(PmgServer *thee, PmgResult *out, int maxn);
/* Block until every submitted request has finished. */
void PmgServer_drain(PmgServer *thee);
/* [0] submitted, [1] finished, [2] running, [3] queued, [4] admitted
 * footprint now (bytes), [5] peak admitted footprint, [6] peak device bytes
 * held by the library's pool, [7] slots. */
void PmgServer_stats(PmgServer *thee, double out[8]);
/* Cross-request batching of the solves (default: max_batch 8, max_running
 * the slot count, so no ready solve waits).  A request whose coefficient maps
 * are built waits for its solve; a waiting request's worker starts a batch
 * whenever fewer than max_running batches run, with the oldest waiting
 * requests of its shape (grid, level count, method, precision), at most
 * max_batch (<= APBS_PMG_MAXBATCH).  max_batch 1 solves every request alone
 * on its own stream.  Returns 0 on invalid values. */
int PmgServer_setBatching(PmgServer *thee, int max_batch, int max_running);
/* Batches solved so far and the requests they held. */
void PmgServer_batchStats(PmgServer *thee, long *batches, long *batched);
/* Drains, then stops the workers and frees the service. */
void PmgServer_dtor(PmgServer **thee);
/* Estimated device bytes one request holds while it runs. */
size_t PmgServer_footprint(const PmgRequest *req);
/* Steady-clock seconds used by PmgResult times. */
double apbs_pmg_now(void);
/* Bytes in use, peak bytes in use and bytes cached by the device pool; the
 * peak restarts from the current use when reset != 0. */
void apbs_pmg_poolStats(size_t *inuse, size_t *peak, size_t *cached, int reset);

#ifdef __cplusplus
}
#endif

#endif
