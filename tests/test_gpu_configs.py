"""The larger workload configurations on the GPU, checked through properties
that do not need the oracle at full size (the FP64 GPU path they are
compared with is pinned to the oracle on the 65^3 workload):

  * lpbe 193^3, 20,000 atoms: mixed precision (FP32 V-cycles, FP64 defect
    correction to a relative FP64 residual of 1e-10) against a tightly
    converged FP64 solve; potential L2 / Linf and energy errors stated.
  * npbe focusing 97^3 (1.2 A) -> 353^3 (0.3 A), 60,000 atoms: the fine
    solve in mixed precision against the FP64 fine solve on the same coarse
    boundary; energy error and achieved GB/s per V-cycle stated.
"""
import time

import numpy as np
import pytest

from apbs_amd import driver, synth

pytestmark = pytest.mark.gpu


def errors(u, ref):
    d = u - ref
    return (float(np.linalg.norm(d) / np.linalg.norm(ref)),
            float(np.abs(d).max() / np.abs(ref).max()))


def test_lpbe193_mixed_against_fp64():
    cfg = synth.CONFIGS["lpbe193"]
    mol = driver.Molecule(*synth.make("lpbe193"))
    c = mol.center
    t = time.perf_counter()
    e64, g64 = driver.run_manual(mol, [193] * 3, c, dict(cfg["pbe"], etol=1e-12),
                                 grid=[cfg["h"]] * 3, mirror=False)
    t64 = time.perf_counter() - t
    t = time.perf_counter()
    emx, gmx = driver.run_manual(mol, [193] * 3, c, cfg["pbe"], grid=[cfg["h"]] * 3,
                                 precision=1, mixed_tol=1e-10, mirror=False)
    tmx = time.perf_counter() - t
    l2, linf = errors(gmx.array("u"), g64.array("u"))
    rel = abs(emx - e64) / abs(e64)
    s = gmx.stats()
    print(f"lpbe193: fp64 {e64:.8f} kJ/mol ({t64:.2f} s, {g64.stats()['iters']} cycles), "
          f"mixed {emx:.8f} ({tmx:.2f} s, {s['outer']} outer, relres {s['relres']:.2e}); "
          f"energy rel {rel:.2e}, potential L2 {l2:.2e} Linf {linf:.2e}")
    assert s["relres"] <= 1e-10
    assert rel < 1e-8 and l2 < 1e-7 and linf < 1e-7


def test_focus353_mixed_against_fp64():
    cfg = synth.CONFIGS["focus353"]
    xyz, q, r = synth.make("focus353")
    mol = driver.Molecule(xyz, q, r)
    c = mol.center
    out = {}
    for prec in (0, 1):
        t = time.perf_counter()
        coarse = driver.Calc(mol, driver.manual_mgparm([cfg["coarse_dime"]] * 3, c,
                                                       grid=[cfg["coarse_h"]] * 3),
                             cfg["pbe"], mirror=False)
        coarse.fillco()
        coarse.solve()
        fine = driver.Calc(mol, driver.manual_mgparm([cfg["dime"]] * 3, c, grid=[cfg["h"]] * 3),
                           cfg["pbe"], old=coarse, precision=prec, mirror=False)
        fine.kernel_timing(True)
        fine.fillco()
        fine.solve()
        fine.set_part_domain()
        e = fine.energy_kJ(1)
        kt = fine.kernel_times()
        st = fine.stats()
        out[prec] = (e, fine.array("u"), time.perf_counter() - t, kt, st)
    e64, u64, t64, kt64, st64 = out[0]
    emx, umx, tmx, ktmx, stmx = out[1]
    l2, linf = errors(umx, u64)
    rel = abs(emx - e64) / abs(e64)
    n0 = cfg["dime"] ** 3
    dims = [cfg["dime"]]
    for _ in range(stmx["nlev"] - 1):
        dims.append((dims[-1] - 1) // 2 + 1)
    words = 44 * n0 + 106 * sum(d ** 3 for d in dims[1:-1])
    for name, kt in (("fp64", kt64), ("mixed", ktmx)):
        if kt["vcycles"]:
            ms = kt["vcycle_ms"] / kt["vcycles"]
            print(f"focus353 {name}: {ms:.3f} ms per V-cycle, "
                  f"{words * kt['word'] / (ms * 1e-3) / 1e9:.0f} GB/s algorithmic; smoother "
                  f"{kt['smoother_ms'] / max(kt['smoother_launches'], 1) * 1e3:.1f} us per sweep")
    print(f"focus353: fp64 {e64:.6f} kJ/mol ({t64:.1f} s, {st64}), mixed {emx:.6f} "
          f"({tmx:.1f} s, {stmx}); energy rel {rel:.2e}, potential L2 {l2:.2e} Linf {linf:.2e}")
    assert rel < 1e-6
    assert l2 < 1e-8 and linf < 1e-8
