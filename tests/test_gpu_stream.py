"""The served stream of binding triplets through the continuous-batching
service: 512 requests (complexes and their two plane-cut halves), grid dims
drawn from {65, 97, 129, 161}, lpbe and npbe 50/50, 300-8000 atoms, mixed
precision (FP32 V-cycles, FP64 defect correction to the FP64 relative
residual 1e-6), Poisson arrivals at 60 requests/s, 16 slots, solves batched
across requests.

Checks:
  * every request succeeds, and the stream reports p50/p99 latency, solves/s
    and the peak device footprint;
  * a fixed sample of requests, each rerun alone (one slot, no batching),
    gives bitwise the energy it got in the stream (batching and overlap do
    not change a result);
  * the same sample solved in FP64 agrees with the mixed-precision energy to
    1e-6 relative.
"""
import numpy as np
import pytest

from apbs_amd import serve

pytestmark = pytest.mark.gpu

RATE = 60.0
SAMPLE = (0, 1, 2, 97, 250, 301, 410, 511)


def make_req(s, precision, tag):
    return serve.Request(s["xyz"], s["q"], s["r"], s["dime"], s["grid"], s["center"], s["pbe"],
                         precision=precision, tag=tag)


def test_binding_stream_512():
    specs = serve.binding_stream(512, seed=3)
    dims = sorted(set(s["dime"][0] for s in specs))
    npbe = np.mean([bool(s["pbe"].get("npbe")) for s in specs])
    print(f"stream: {len(specs)} requests, dims {dims}, npbe fraction {npbe:.2f}, atoms "
          f"{min(len(s['q']) for s in specs)}..{max(len(s['q']) for s in specs)}", flush=True)
    assert dims == [65, 97, 129, 161]
    srv = serve.Server(slots=16)
    res, wall = serve.run_stream(srv, specs, RATE, seed=3, precision=1)
    st = srv.stats()
    srv.close()
    summ = serve.latency_summary(res, wall)
    print(f"stream at {RATE}/s: {summ}, peak pool {st['peak_pool_bytes'] / 1e9:.2f} GB, "
          f"batches {st['batches']} holding {st['batched_requests']}", flush=True)
    assert summ["failed"] == 0 and summ["requests"] == 512
    assert st["batched_requests"] == 512
    # the sample alone, and in FP64
    alone = serve.Server(slots=1, max_batch=1)
    for i in SAMPLE:
        s = specs[i]
        for precision in (1, 0):
            alone.submit(make_req(s, precision, i))
            alone.drain()
            (r,) = alone.poll()
            assert r["status"] == 1
            if precision == 1:
                assert r["energy_kT"] == res[i]["energy_kT"], (i, r["energy_kT"],
                                                              res[i]["energy_kT"])
            else:
                rel = abs(res[i]["energy_kT"] - r["energy_kT"]) / abs(r["energy_kT"])
                print(f"request {i} (dime {s['dime'][0]}, npbe {bool(s['pbe'].get('npbe'))}, "
                      f"{len(s['q'])} atoms): mixed {res[i]['energy_kT']!r}, fp64 "
                      f"{r['energy_kT']!r}, rel {rel:.2e}", flush=True)
                assert rel <= 1e-6
    alone.close()
