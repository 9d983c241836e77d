"""Benchmark: PB solves/s and p50/p99 latency, mixed precision vs FP64, at 161^3 npbe.

Workload (the metric's configuration): nonlinear PB, Newton-multigrid, grid
161^3 at 0.4 A, a seeded 5,000-atom globule in a 28x24x20 A ellipsoid with net
charge -10 e, pdie 4, sdie 78.54, srfm smol, chgm spl2, bcfl mdh, 0.10 M and
0.05 M monovalent salt.

One request = one complete PB calculation: Vpmg_ctor + Vpmg_fillco
(coefficient maps from the atom list) + Vpmg_solve (Newton multigrid to etol
1e-6) + Vpmg_energy.  The potential stays in HBM (no host mirror).  The atom
list (120 KB) is uploaded inside fillco; that copy is part of the request.

Precision modes (Vpmg_setPrecision): fp64 (FP64 throughout); mixed (FP32
V-cycles in an FP64 defect-correction loop: the FP64 answer to the solver
tolerance); mixed_fastbc (mixed, and the MDH boundary summed in FP64 from
FP32 atom terms: boundary values within about 1e-6 relative, energy within
1e-10 of the oracle's here).  The headline is the fastest mode held to the
metric's 1e-6 relative energy bar, mixed_fastbc.  Paths, each in every mode:
  * single: one request at a time on one Vpmg (--warmup, then --steps timed
    requests; the Vpmg and its hierarchy are reused);
  * served: the continuous-batching service (--slots requests in flight,
    solves of ready requests batched up to 8 per launch), closed loop: --warmup
    rounds, then exactly --steps timed rounds of --slots requests, all
    submitted at once.  `value` is the served headline mode's requests/s;
    `steps` counts those rounds, `ms_per_step` is the time per round;
  * served_stream: the served workload of binding triplets (512 requests,
    dims 65..161, lpbe/npbe 50/50, 300-8000 atoms) with Poisson arrivals at
    two rates below saturation: p50/p99 latency (completion - arrival),
    solves/s and peak device memory.
Beside them: the energy and potential error of each mixed mode against FP64,
the roofline of the dominant kernel (the fine-level red-black smoother, from
HIP events recorded on the solver's stream during the single-request run of
the headline mode)
and the oracle (the FP64 CPU restatement of the reference, built -O3, one
host thread) timed on one complete calculation of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
WORKLOAD = "npbe161"
PRECISION = {"fp64": 0, "mixed": 1, "mixed_fastbc": 2}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default=WORKLOAD)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--modes", default="fp64,mixed,mixed_fastbc")
    ap.add_argument("--slots", type=int, default=8, help="concurrent requests of the served path")
    ap.add_argument("--served", type=int, default=1, help="run the served closed loop (0: skip)")
    ap.add_argument("--stream", type=int, default=1,
                    help="run the served stream of binding triplets at two rates (0: skip)")
    ap.add_argument("--pmc", action="store_true",
                    help="counter-collection workload: the fine level's fused sweeps in each "
                         "precision and a known-size copy kernel (run under rocprofv3 --pmc)")
    return ap.parse_args()


def run_mode(mol, cfg, precision, steps, warmup):
    from apbs_amd import driver, lib
    L = lib()
    pbe = dict(cfg["pbe"])
    center = mol.center
    m = driver.manual_mgparm([cfg["dime"]] * 3, center, grid=[cfg["h"]] * 3, chgm=pbe["chgm"])
    calc = driver.Calc(mol, m, pbe, precision=precision, mixed_tol=-1.0, mirror=False)

    def step():
        calc.fillco()
        calc.solve()
        return calc.energy(1)

    for _ in range(warmup):
        step()
    lat = []
    L.apbs_pmg_synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        e = step()
        lat.append(time.perf_counter() - ts)
    L.apbs_pmg_synchronize()
    total = time.perf_counter() - t0
    # kernel times from HIP events on the solver's stream, in extra steps
    # (event recording launches the V-cycle eagerly instead of from its graph)
    calc.kernel_timing(True)
    for _ in range(max(2, steps // 4)):
        step()
    L.apbs_pmg_synchronize()
    kt = calc.kernel_times()
    calc.kernel_timing(False)
    stats = calc.stats()
    lat_ms = np.array(lat) * 1e3
    res = {
        "solves_per_s": steps / total,
        "ms_per_solve": total / steps * 1e3,
        "p50_ms": float(np.percentile(lat_ms, 50)),
        "p99_ms": float(np.percentile(lat_ms, 99)),
        "energy_kT": e,
        "newton_steps": stats["iters"],
        "inner_cycles": stats["inner"],
        "levels": stats["nlev"],
        "relres": stats["relres"],
        "kernel": kt,
    }
    return res, calc


COPY_DOUBLES = 32 * 2 ** 20   # calibration copy: 256 MiB read + 256 MiB written per launch
PMC_SUMMARY = os.path.join(ROOT, "profiles", "pmc_smoother_npbe161.md")


def measured_traffic(word):
    """HBM bytes per fused-sweep launch from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE, each scaled by the copy-kernel calibration), or
    None when it has not been measured."""
    key = "traffic_bytes_per_launch_" + ("fp32" if word == 4 else "fp64")
    try:
        for line in open(PMC_SUMMARY):
            if line.startswith(key + ":"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


PMC_SWEEPS = 20


def pmc_workload(cfg):
    """The fine level's fused sweep launches alone (FP64 and FP32, on a
    synthetic operator of the workload's grid: the kernel's memory traffic
    depends on the grid shape only), then the known-size copy kernel."""
    from apbs_amd import lib
    import ctypes
    L = lib()
    L.pmgtest_fused_sweeps.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_double)]
    us = ctypes.c_double()
    for word in (8, 4):
        if L.pmgtest_fused_sweeps(cfg["dime"], word, PMC_SWEEPS, 0, ctypes.byref(us)) != 1:
            raise RuntimeError("pmgtest_fused_sweeps failed")
        print(f"# pmc workload: {PMC_SWEEPS} fused sweeps, {cfg['dime']}^3, word {word}: "
              f"{us.value:.2f} us per launch", flush=True)
    L.pmgtest_stream_copy.argtypes = [ctypes.c_size_t, ctypes.c_int]
    L.pmgtest_stream_copy(COPY_DOUBLES, 3)
    print(f"# pmc workload done: copy kernel {COPY_DOUBLES * 8} bytes each way x 3", flush=True)


def smoother_roofline(kt):
    """Fine-level red-black smoother: algorithmic bytes 7 words per point per
    full sweep (D = oC + cc, oE, oN, uC, f, x in, x out); the library reports
    the words per point one launch moves (7 for the fused sweep kernel) and
    times only such sweeps (a mixed-precision cycle's first sweep starts from
    a zero correction and reads no x: that pair is not timed)."""
    if kt["smoother_launches"] == 0:
        return None
    per_launch_ms = kt["smoother_ms"] / kt["smoother_launches"]
    bytes_per_launch = kt["words_per_point"] * kt["word"] * kt["points"]
    achieved = bytes_per_launch / (per_launch_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(kt["word"]),
            "kernel": "%s<%s>" % ("k_gsrb7_vfused" if kt.get("vector_form") else
                                  "k_gsrb7_fused" if kt["words_per_point"] == 7 else "k_gsrb7",
                                  "float" if kt["word"] == 4 else "double"),
            "bytes_per_launch": bytes_per_launch, "avg_launch_us": per_launch_ms * 1e3}


LARGE_GRID = 193
LARGE_SWEEPS = 20


def large_grid_roofline(n=LARGE_GRID, sweeps=LARGE_SWEEPS):
    """The fused sweep on a grid whose fine arrays do not fit the 256 MB
    Infinity Cache (193^3: 7.2M points, 201 MB per FP32 sweep, 403 MB per
    FP64 sweep): back-to-back launches on a synthetic operator, timed with
    HIP events around the launches (pmgtest_fused_sweeps), 7 words per point."""
    from apbs_amd import lib
    import ctypes
    L = lib()
    L.pmgtest_fused_sweeps.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_double)]
    us = ctypes.c_double()
    out = {}
    for word, name in ((4, "fp32"), (8, "fp64")):
        if L.pmgtest_fused_sweeps(n, word, sweeps, 0, ctypes.byref(us)) != 1:
            raise RuntimeError("pmgtest_fused_sweeps failed")
        b = 7 * word * n ** 3
        gbs = b / (us.value * 1e-6) / 1e9
        out[name] = {"grid": n, "avg_launch_us": round(us.value, 2), "bytes_per_launch": b,
                     "achieved": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    return out


def vcycle_bandwidth(kt, cfg, levels):
    """Achieved GB/s per V-cycle against the algorithmic V-cycle bytes
    w * (44 N0 + 106 * sum of the intermediate levels' points); the coarsest
    level (solved by a dense inverse) is excluded."""
    if kt["vcycles"] == 0:
        return None
    n = cfg["dime"]
    dims = [n]
    for _ in range(levels - 1):
        dims.append((dims[-1] - 1) // 2 + 1)
    inter = sum(d ** 3 for d in dims[1:-1])
    words = 44 * n ** 3 + 106 * inter
    ms = kt["vcycle_ms"] / kt["vcycles"]
    return {"ms_per_vcycle": ms, "algorithmic_GB": words * kt["word"] / 1e9,
            "GBs": words * kt["word"] / (ms * 1e-3) / 1e9, "levels": dims}


def cpu_baseline(cfg):
    """The oracle (FP64 CPU restatement of the reference's pmgc path, built
    -O3, one host thread) on one complete calculation of the workload:
    Vpmg_fillco then the whole Newton solve, timed."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import orc_driver as od
    from apbs_amd import synth
    atoms = synth.make(WORKLOAD)
    c = od.mol_center(atoms[0])
    nlev, dime = od.check_nlev([cfg["dime"]] * 3)
    p = od.make_params(dime, [cfg["h"]] * 3, c, nlev, dict(cfg["pbe"]))
    o = od.OrcCalc(p, atoms)
    t = time.perf_counter()
    o.fillco()
    t_fill = time.perf_counter() - t
    t = time.perf_counter()
    o.solve()
    t_solve = time.perf_counter() - t
    st = o.stats()
    total = t_fill + t_solve
    return {"value": round(1.0 / total, 5), "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"{WORKLOAD}: one complete calculation by the oracle (gcc -O3, one host "
                      f"thread): fillco {t_fill:.1f} s + Newton solve {t_solve:.1f} s "
                      f"({st['iters']} steps, reference schedule)"}


def spec_of(xyz, q, r, cfg):
    from apbs_amd import driver
    center = driver.Molecule(xyz, q, r).center
    return dict(xyz=xyz, q=q, r=r, dime=(cfg["dime"],) * 3, grid=[cfg["h"]] * 3, center=center,
                pbe=dict(cfg["pbe"]))


def run_served(spec, precision, rounds, slots, warmup):
    """The served path, closed loop: `warmup` rounds, then exactly `rounds`
    timed rounds of `slots` identical requests submitted at once.  solves/s
    is the timed requests over the time from the first submission to the
    last completion; latency is completion minus submission per request."""
    from apbs_amd import lib, serve
    import ctypes
    srv = serve.Server(slots=slots)
    serve.run_stream(srv, [spec] * (max(warmup, 1) * slots), rate=0, precision=precision)
    z = ctypes.c_size_t()
    lib().apbs_pmg_poolStats(ctypes.byref(z), ctypes.byref(z), ctypes.byref(z), 1)
    b0 = srv.stats()
    res, wall = serve.run_stream(srv, [spec] * (rounds * slots), rate=0, precision=precision)
    st = srv.stats()
    srv.close()
    summ = serve.latency_summary(res, wall)
    summ["ms_per_round"] = wall / rounds * 1e3
    summ["energy_kT"] = res[0]["energy_kT"]
    summ["peak_admitted_GB"] = st["peak_admitted_bytes"] / 1e9
    summ["peak_pool_GB"] = st["peak_pool_bytes"] / 1e9
    summ["slots"] = slots
    nb = st["batches"] - b0["batches"]
    summ["batches"] = nb
    summ["mean_batch"] = (st["batched_requests"] - b0["batched_requests"]) / max(nb, 1)
    return summ


STREAM_REQUESTS = 512
STREAM_RATES = (25.0, 50.0)


def served_stream(precision, rates=STREAM_RATES, n=STREAM_REQUESTS, slots=16):
    """The served workload of binding triplets (serve.binding_stream: dims
    65..161, lpbe/npbe 50/50, 300-8000 atoms) in one precision mode with
    Poisson arrivals at each rate."""
    from apbs_amd import lib, serve
    import ctypes
    specs = serve.binding_stream(n, seed=3)
    out = []
    for rate in rates:
        z = ctypes.c_size_t()
        lib().apbs_pmg_poolStats(ctypes.byref(z), ctypes.byref(z), ctypes.byref(z), 1)
        srv = serve.Server(slots=slots)
        res, wall = serve.run_stream(srv, specs, rate, seed=3, precision=precision)
        st = srv.stats()
        srv.close()
        summ = serve.latency_summary(res, wall)
        summ.update(rate=rate, slots=slots, peak_pool_GB=round(st["peak_pool_bytes"] / 1e9, 3),
                    peak_admitted_GB=round(st["peak_admitted_bytes"] / 1e9, 3),
                    mean_batch=round(st["batched_requests"] / max(st["batches"], 1), 2))
        out.append(summ)
        print(f"# served stream at {rate}/s: {summ}", file=sys.stderr, flush=True)
    return {"workload": f"{n} requests, binding triplets, dims 65/97/129/161, lpbe/npbe 50/50, "
                        f"300-8000 atoms, precision mode {precision}, Poisson arrivals",
            "rates": out}


def main():
    a = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print("bench.py: this build runs on one GPU only; a multi-GPU figure is not measured "
              "(use --gpus 1)", file=sys.stderr)
        sys.exit(2)
    from apbs_amd import driver, synth
    cfg = synth.CONFIGS[a.workload]
    xyz, q, r = synth.make(a.workload)
    mol = driver.Molecule(xyz, q, r)
    if a.pmc:
        pmc_workload(cfg)
        return
    modes = a.modes.split(",")
    out = {}
    calcs = {}
    for mode in modes:
        res, calc = run_mode(mol, cfg, PRECISION[mode], a.steps, a.warmup)
        out[mode] = res
        calcs[mode] = calc
        print(f"# {mode}: {res['solves_per_s']:.2f} solves/s, p50 {res['p50_ms']:.2f} ms, "
              f"p99 {res['p99_ms']:.2f} ms, newton {res['newton_steps']}, "
              f"inner {res['inner_cycles']}, kernel {res['kernel']}", file=sys.stderr, flush=True)
    head = next((m for m in ("mixed_fastbc", "mixed") if m in out), modes[0])
    err = {}
    if "fp64" in out:
        e64 = out["fp64"]["energy_kT"]
        u64 = calcs["fp64"].array("u")
        for mode in out:
            if mode == "fp64":
                continue
            emx = out[mode]["energy_kT"]
            d = calcs[mode].array("u") - u64
            err[mode] = {"energy_rel": abs(emx - e64) / abs(e64),
                         "potential_rel_L2": float(np.linalg.norm(d) / np.linalg.norm(u64)),
                         "potential_rel_Linf": float(np.abs(d).max() / np.abs(u64).max())}
    kt = out[head]["kernel"]
    for c in calcs.values():
        c.release()
    calcs.clear()
    served = {}
    if a.served:
        spec = spec_of(xyz, q, r, cfg)
        for mode in modes:
            served[mode] = run_served(spec, PRECISION[mode], a.steps, a.slots, a.warmup)
            print(f"# served {mode}: {served[mode]}", file=sys.stderr, flush=True)
    stream = served_stream(PRECISION[head]) if a.stream else None
    top = served.get(head)
    if top is not None:
        value, ms, p50, p99 = (top["solves_per_s"], top["ms_per_round"], top["p50_ms"],
                               top["p99_ms"])
    else:
        value, ms, p50, p99 = (out[head]["solves_per_s"], out[head]["ms_per_solve"],
                               out[head]["p50_ms"], out[head]["p99_ms"])
    from apbs_amd import synth
    line = {
        "metric": "solves/s + p50/p99 latency, mixed vs FP64 PMG at 161^3 npbe; HBM GB/s",
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms, 3),
        "p50_ms": round(p50, 3),
        "p99_ms": round(p99, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head == "fp64" else "f32+f64",
        "data": "synthetic (seeded globule)",
        "config": {"workload": f"{a.workload}: npbe 161^3 at 0.4 A, 5000 atoms, smol, spl2, mdh",
                   "molecule": synth.describe(a.workload),
                   "path": (f"{head}, served ({a.slots} slots, batched solves)"
                            if top is not None else head),
                   "step": (f"one round of {a.slots} requests (fillco + solve + energy each)"
                            if top is not None else "one request: fillco + solve + energy")},
        "served": served,
        "headline_mode": head,
        "speedup_served_over_single_fp64": (
            {m: served[m]["solves_per_s"] / out["fp64"]["solves_per_s"] for m in served}
            if "fp64" in out else None),
        "modes": {k: {kk: v[kk] for kk in ("solves_per_s", "ms_per_solve", "p50_ms", "p99_ms",
                                           "newton_steps", "inner_cycles", "levels", "relres",
                                           "energy_kT")} for k, v in out.items()},
        "mixed_vs_fp64": err or None,
        "speedup_single_over_fp64": (
            {m: out[m]["solves_per_s"] / out["fp64"]["solves_per_s"] for m in out if m != "fp64"}
            if "fp64" in out else None),
        "roofline": smoother_roofline(kt),
        "roofline_fp64": smoother_roofline(out["fp64"]["kernel"]) if "fp64" in out else None,
        "vcycle": vcycle_bandwidth(kt, cfg, out[head]["levels"]),
        "roofline_large_grid": large_grid_roofline(),
        "served_stream": stream,
    }
    line["cpu_baseline"] = None if a.no_cpu_baseline else cpu_baseline(cfg)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
