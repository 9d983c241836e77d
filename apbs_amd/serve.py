"""Python side of the continuous-batching PB service (include/apbs_pmg_server.h).

`Server` wraps PmgServer; `request` builds one mg-manual calculation;
`binding_stream` builds the served workload of binding triplets (a complex and
the two partners made by cutting its atoms with a random plane, on one grid),
and `run_stream` replays it with Poisson arrivals and reports latency and
throughput.
"""
import ctypes
import threading
import time

import numpy as np

from . import (BCFL_MDH, PmgRequest, PmgResult, VCM_BSPL2, VCM_TRIL, VSM_MOL, VSM_MOLSMOOTH,
               VSM_SPLINE, lib, synth)

SRFM = {"mol": VSM_MOL, "smol": VSM_MOLSMOOTH, "spl2": VSM_SPLINE}
CHGM = {"spl0": VCM_TRIL, "spl2": VCM_BSPL2}
_dp = ctypes.POINTER(ctypes.c_double)


def _ptr(a):
    return a.ctypes.data_as(_dp)


class Request:
    """One calculation; keeps the arrays the C struct points to alive."""

    def __init__(self, xyz, charge, radius, dime, grid, center, pbe, precision=1,
                 mixed_tol=-1.0, want_potential=False, tag=0):
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        self.q = np.ascontiguousarray(charge, dtype=np.float64)
        self.r = np.ascontiguousarray(radius, dtype=np.float64)
        self.dime = tuple(int(d) for d in dime)
        self.potential = (np.zeros(self.dime[0] * self.dime[1] * self.dime[2])
                          if want_potential else None)
        c = PmgRequest()
        c.natoms = len(self.q)
        c.xyz, c.radius, c.charge = _ptr(self.xyz), _ptr(self.r), _ptr(self.q)
        for j in range(3):
            c.dime[j] = self.dime[j]
            c.grid[j] = float(grid[j])
            c.center[j] = float(center[j])
        c.npbe = 1 if pbe.get("npbe") else 0
        ions = pbe.get("ions", [])
        c.nion = len(ions)
        for i, (qi, ci, ri) in enumerate(ions):
            c.ionq[i], c.ionc[i], c.ionr[i] = qi, ci, ri
        c.pdie, c.sdie = pbe["pdie"], pbe["sdie"]
        c.srad, c.swin = pbe.get("srad", 1.4), pbe.get("swin", 0.3)
        c.sdens, c.temp = pbe.get("sdens", 10.0), pbe.get("temp", 298.15)
        c.srfm, c.chgm = SRFM[pbe["srfm"]], CHGM[pbe["chgm"]]
        c.bcfl = BCFL_MDH
        c.etol = pbe.get("etol", -1.0)
        c.precision = precision
        c.mixed_tol = mixed_tol
        c.potential = _ptr(self.potential) if self.potential is not None else None
        c.tag = tag
        self.c = c

    def footprint(self):
        return lib().PmgServer_footprint(ctypes.byref(self.c))


def _result(r):
    return {"id": r.id, "tag": r.tag, "status": r.status, "energy_kT": r.energy_kT,
            "iters": r.iters, "inner": r.inner, "relres": r.relres, "t_submit": r.t_submit,
            "t_start": r.t_start, "t_done": r.t_done, "footprint": r.footprint,
            "latency": r.t_done - r.t_submit}


class Server:
    """PmgServer: `slots` requests in flight; their solves are batched across
    requests of one shape, at most `max_batch` per batch and `max_running`
    batches at once (default: the slot count; max_batch 1: every request
    solved alone)."""

    def __init__(self, slots=8, hbm_budget=200 * 2 ** 30, max_batch=8, max_running=None):
        self.h = lib().PmgServer_ctor(int(slots), int(hbm_budget))
        if not self.h:
            raise RuntimeError("PmgServer_ctor failed")
        if max_running is None:
            max_running = slots
        if not lib().PmgServer_setBatching(self.h, int(max_batch), int(max_running)):
            raise ValueError("invalid batching parameters")
        self._keep = {}

    def submit(self, req):
        rid = lib().PmgServer_submit(self.h, ctypes.byref(req.c))
        if rid < 0:
            raise ValueError("request rejected (footprint above the budget or malformed)")
        self._keep[rid] = req
        return rid

    def poll(self):
        buf = (PmgResult * 64)()
        out = []
        while True:
            n = lib().PmgServer_poll(self.h, buf, 64)
            out.extend(_result(buf[i]) for i in range(n))
            if n < 64:
                break
        for r in out:
            r["request"] = self._keep.pop(r["id"], None)
        return out

    def drain(self):
        lib().PmgServer_drain(self.h)

    def stats(self):
        o = np.zeros(8)
        lib().PmgServer_stats(self.h, _ptr(o))
        nb, nr = ctypes.c_long(), ctypes.c_long()
        lib().PmgServer_batchStats(self.h, ctypes.byref(nb), ctypes.byref(nr))
        return {"submitted": int(o[0]), "finished": int(o[1]), "running": int(o[2]),
                "queued": int(o[3]), "admitted_bytes": o[4], "peak_admitted_bytes": o[5],
                "peak_pool_bytes": o[6], "slots": int(o[7]), "batches": nb.value,
                "batched_requests": nr.value}

    def close(self):
        if self.h:
            h = ctypes.c_void_p(self.h)
            lib().PmgServer_dtor(ctypes.byref(h))
            self.h = None

    def __del__(self):
        self.close()


# ----------------------------------------------------------------- workload

LPBE = dict(pdie=2.0, sdie=78.54, srfm="spl2", chgm="spl2", swin=0.3, srad=1.4, sdens=10.0,
            temp=298.15, bcfl="mdh", ions=[(1, 0.15, 2.0), (-1, 0.15, 2.0)])
NPBE = dict(npbe=True, pdie=4.0, sdie=78.54, srfm="smol", chgm="spl2", swin=0.3, srad=1.4,
            sdens=10.0, temp=298.15, bcfl="mdh",
            ions=[(1, 0.10, 2.0), (-1, 0.10, 2.0), (1, 0.05, 2.0), (-1, 0.05, 2.0)])


def binding_stream(n_requests, seed=3, dims=(65, 97, 129, 161), natoms=(300, 8000)):
    """Request specs grouped as binding triplets.  Each triplet draws a grid
    size, a type (lpbe / npbe, 50/50) and a globule of natoms atoms (the
    seeded generator, FCC density at 2.5 A separation); the complex and the
    two halves cut by a random plane through its centre share one grid whose
    length is the molecule's extent plus 2 x (largest radius + 5 A)."""
    rng = np.random.default_rng(seed)
    specs = []
    t = 0
    while len(specs) < n_requests:
        n = int(rng.integers(natoms[0], natoms[1] + 1))
        dime = int(rng.choice(dims))
        npbe = bool(rng.integers(0, 2))
        rad = (n / (np.sqrt(2.0) / 2.5 ** 3 * 0.95) * 3 / (4 * np.pi)) ** (1 / 3)
        axes = rad * rng.uniform(0.85, 1.15, 3)
        xyz, q, r, _ = synth.globule(n, axes, int(rng.integers(1 << 30)))
        normal = rng.normal(size=3)
        side = (xyz - xyz.mean(0)) @ normal > 0
        center = 0.5 * (xyz.min(0) + xyz.max(0))
        glen = (xyz.max(0) - xyz.min(0)) + 2 * (r.max() + 5.0)
        grid = glen / (dime - 1)
        pbe = NPBE if npbe else LPBE
        for part, sel in (("complex", np.ones(n, bool)), ("a", side), ("b", ~side)):
            specs.append(dict(xyz=xyz[sel], q=q[sel], r=r[sel], dime=(dime,) * 3, grid=grid,
                              center=center, pbe=pbe, triplet=t, part=part))
        t += 1
    return specs[:n_requests]


def run_stream(server, specs, rate, seed=0, precision=1, want_potential=False):
    """Submit specs with Poisson arrivals at `rate` requests/s (rate <= 0:
    all at once) and wait for every result.  Returns (results in submission
    order, wall time from first arrival to last completion)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, len(specs)) if rate > 0 else np.zeros(len(specs))
    reqs = [Request(s["xyz"], s["q"], s["r"], s["dime"], s["grid"], s["center"], s["pbe"],
                    precision=precision, want_potential=want_potential, tag=i)
            for i, s in enumerate(specs)]
    results = {}
    lock = threading.Lock()
    done = threading.Event()

    def collector():
        # stops once submission has ended (whatever happened to it) and the
        # main thread has taken over the last poll after draining
        while not done.is_set():
            got = server.poll()
            with lock:
                for r in got:
                    results[r["tag"]] = r
            time.sleep(0.002)

    th = threading.Thread(target=collector, daemon=True)
    th.start()
    t0 = time.perf_counter()
    due = t0
    try:
        for req, g in zip(reqs, gaps):
            due += g
            d = due - time.perf_counter()
            if d > 0:
                time.sleep(d)
            server.submit(req)
    finally:
        done.set()
        th.join()
        server.drain()
    for r in server.poll():
        results[r["tag"]] = r
    wall = max(r["t_done"] for r in results.values()) - min(r["t_submit"] for r in results.values())
    return [results[i] for i in range(len(reqs))], wall


def latency_summary(results, wall):
    lat = np.array([r["latency"] for r in results]) * 1e3
    return {"requests": len(results), "solves_per_s": len(results) / wall,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "failed": int(sum(r["status"] != 1 for r in results))}


def main(argv=None):
    """The served stream of binding triplets at a sweep of Poisson arrival
    rates: one line per rate with p50/p99 latency, solves/s and the peak
    device footprint."""
    import argparse
    import json
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--requests", type=int, default=512)
    ap.add_argument("--rates", default="10,20,40,0", help="requests/s; 0 = all at once")
    ap.add_argument("--slots", type=int, default=8)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--precision", type=int, default=1)
    ap.add_argument("--seed", type=int, default=3)
    a = ap.parse_args(argv)
    specs = binding_stream(a.requests, seed=a.seed)
    for rate in [float(x) for x in a.rates.split(",")]:
        srv = Server(slots=a.slots, hbm_budget=int(a.budget_gb * 2 ** 30))
        res, wall = run_stream(srv, specs, rate, seed=a.seed, precision=a.precision)
        st = srv.stats()
        srv.close()
        out = latency_summary(res, wall)
        out.update(rate=rate, slots=a.slots, precision=a.precision,
                   peak_admitted_GB=st["peak_admitted_bytes"] / 1e9,
                   peak_pool_GB=st["peak_pool_bytes"] / 1e9,
                   dims=sorted(set(s["dime"][0] for s in specs)),
                   npbe_fraction=float(np.mean([bool(s["pbe"].get("npbe")) for s in specs])))
        print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
