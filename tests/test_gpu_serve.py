"""The continuous-batching service (include/apbs_pmg_server.h) on the GPU.

Every kernel of a calculation is deterministic (ordered gathers, fixed
two-stage reductions, order-free scatters), and a batched solve runs each
calculation's own arithmetic, so a calculation's result does not depend on
what else runs on the device or shares its launches: the bar is bitwise
equality of energy and potential between a request run alone (one slot, no
batching) and the same request run in a mix (several slots, solves batched
across requests), and against the same calculation run directly through
Vpmg.
"""
import numpy as np
import pytest

from apbs_amd import driver, serve

pytestmark = pytest.mark.gpu


def mix_specs():
    specs = serve.binding_stream(6, seed=11, dims=(33, 49), natoms=(150, 600))
    for s in specs[3:]:     # one npbe and one lpbe triplet
        s["pbe"] = serve.LPBE
    return specs


def make_req(s, precision, tag=0):
    return serve.Request(s["xyz"], s["q"], s["r"], s["dime"], s["grid"], s["center"], s["pbe"],
                         precision=precision, want_potential=True, tag=tag)


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_alone_equals_in_mix(precision):
    specs = mix_specs()
    alone = []
    srv = serve.Server(slots=1, max_batch=1)
    for i, s in enumerate(specs):
        req = make_req(s, precision, i)
        srv.submit(req)
        srv.drain()
        (r,) = srv.poll()
        assert r["status"] == 1
        alone.append((r["energy_kT"], req.potential.copy()))
    srv.close()
    # one batch at a time, so requests that finish their setup meanwhile
    # wait at the gate and are solved together
    srv = serve.Server(slots=4, max_running=1)
    reqs = [make_req(s, precision, i) for i, s in enumerate(specs)]
    for req in reqs:
        srv.submit(req)
    srv.drain()
    res = {r["tag"]: r for r in srv.poll()}
    st = srv.stats()
    srv.close()
    assert len(res) == len(specs) and st["finished"] == len(specs)
    print(f"mix: {st['batches']} batches holding {st['batched_requests']} requests")
    assert st["batched_requests"] == len(specs)
    overlap = 0
    for i, req in enumerate(reqs):
        assert res[i]["status"] == 1
        assert res[i]["energy_kT"] == alone[i][0], (i, res[i]["energy_kT"], alone[i][0])
        assert np.array_equal(req.potential, alone[i][1]), i
    # at least two requests were on the device at the same time
    iv = sorted((r["t_start"], r["t_done"]) for r in res.values())
    for a, b in zip(iv, iv[1:]):
        overlap += b[0] < a[1]
    assert overlap > 0


def test_server_matches_direct_vpmg():
    s = mix_specs()[0]
    req = make_req(s, 0)
    srv = serve.Server(slots=2)
    srv.submit(req)
    srv.drain()
    (r,) = srv.poll()
    srv.close()
    mol = driver.Molecule(s["xyz"], s["q"], s["r"])
    e, calc = driver.run_manual(mol, s["dime"], s["center"], s["pbe"], grid=s["grid"])
    assert r["energy_kT"] == calc.energy(1)
    assert np.array_equal(req.potential, calc.array("u").ravel())


def test_footprint_bounds_pool_use_and_budget_is_respected():
    from apbs_amd import lib
    import ctypes
    specs = mix_specs()
    fps = [make_req(s, 1).footprint() for s in specs]
    inuse, peak, cached = (ctypes.c_size_t() for _ in range(3))
    # one request at a time: the pool's peak growth stays under its estimate
    srv = serve.Server(slots=1)
    for s, fp in zip(specs, fps):
        lib().apbs_pmg_poolStats(ctypes.byref(inuse), ctypes.byref(peak), ctypes.byref(cached), 1)
        base = inuse.value
        srv.submit(make_req(s, 1))
        srv.drain()
        srv.poll()
        lib().apbs_pmg_poolStats(ctypes.byref(inuse), ctypes.byref(peak), ctypes.byref(cached), 0)
        assert peak.value - base <= fp, (peak.value - base, fp)
    srv.close()
    # a budget that admits at most two of the largest at once
    budget = 2 * max(fps)
    srv = serve.Server(slots=6, hbm_budget=budget)
    for s in specs:
        srv.submit(make_req(s, 1))
    srv.drain()
    res = srv.poll()
    st = srv.stats()
    srv.close()
    assert len(res) == len(specs) and all(r["status"] == 1 for r in res)
    assert st["peak_admitted_bytes"] <= budget
    # a request larger than the budget is refused at submit
    srv = serve.Server(slots=1, hbm_budget=min(fps) - 1)
    with pytest.raises(ValueError):
        srv.submit(make_req(specs[0], 1))
    srv.close()
