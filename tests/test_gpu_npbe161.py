"""The nonlinear 161^3 workload at full size against the oracle.

Workload: npbe, Newton multigrid, 161^3 at 0.4 A, the seeded 5,000-atom
globule (28 x 24 x 20 A ellipsoid, net charge -10 e), pdie 4, sdie 78.54,
srfm smol, chgm spl2, bcfl mdh, 0.10 M and 0.05 M monovalent salt.

The oracle's answers are committed fixtures (tests/golden/make_npbe161.py
wrote them from the oracle's reference schedule at etol 1e-12): the energy,
||u||_2, ||u||_inf, sum(u), the projections of u on 8 seeded random sign
fields and u at 8192 seeded points.  A difference field e moves each
projection by about ||e||_2 wherever it sits on the grid, so bounding the 8
projection differences by 5e-9 ||u||_2 bounds the relative L2 difference of
the whole potential at the 1e-9 level without the 33 MB field itself.

Bars (FP64 GPU at etol 1e-12): energy <= 1e-9 relative; ||u||_2, ||u||_inf,
sum(u) <= 1e-9 relative; sampled values <= 1e-9 of ||u||_inf; projections
<= 5e-9 ||u||_2.  Mixed precision at the default etol (1e-6) is held to the
oracle's energy at 1e-6 relative.

Forces: srfm smol with forces aborts the reference (vpmg.c:5471-5477,
routines.c:1840-1842), so they are checked on the companion solve of the same
molecule and grid with srfm spl2: qf, ib and db forces of 32 seeded atoms
within 1e-8 of the largest force magnitude among them.
"""
import json
import os

import numpy as np
import pytest

from apbs_amd import driver, synth

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


@pytest.fixture(scope="module")
def ref():
    with open(os.path.join(GOLD, "npbe161.json")) as fh:
        meta = json.load(fh)
    arr = np.load(os.path.join(GOLD, "npbe161.npz"))
    return meta, {k: arr[k] for k in arr.files}


@pytest.fixture(scope="module")
def workload():
    cfg = synth.CONFIGS["npbe161"]
    mol = driver.Molecule(*synth.make("npbe161"))
    return cfg, mol


def sign_field(k, n):
    return np.random.default_rng(1000 + k).integers(0, 2, n, dtype=np.int8) * 2.0 - 1.0


def rel(a, b):
    return abs(a - b) / abs(b)


def check_potential(u, pot, idx, samples, label):
    u = u.ravel()
    l2, linf = np.linalg.norm(u), np.abs(u).max()
    errs = {"l2": rel(l2, pot["l2"]), "linf": rel(linf, pot["linf"]),
            "sum": rel(u.sum(), pot["sum"]),
            "sample": (float(np.abs(u[idx] - samples).max() / pot["linf"])
                       if idx is not None else 0.0),
            "proj": max(abs(u @ sign_field(k, u.size) - p) for k, p in enumerate(pot["proj"]))
            / pot["l2"]}
    print(f"{label}: potential against the oracle {errs}")
    assert errs["l2"] <= 1e-9 and errs["linf"] <= 1e-9 and errs["sum"] <= 1e-9, errs
    assert errs["sample"] <= 1e-9 and errs["proj"] <= 5e-9, errs
    return errs


def test_npbe161_fp64_matches_oracle(ref, workload):
    meta, arr = ref
    cfg, mol = workload
    want = meta["smol"]
    e, g = driver.run_manual(mol, [cfg["dime"]] * 3, mol.center, dict(cfg["pbe"], etol=1e-12),
                             grid=[cfg["h"]] * 3, mirror=False)
    e_kT = g.energy(1)
    comp = g.components(1)
    print(f"npbe161 fp64 etol 1e-12: {e_kT!r} kT against {want['energy_kT']!r}, stats "
          f"{g.stats()}, components {comp}")
    assert rel(e_kT, want["energy_kT"]) <= 1e-9
    for k in ("qf", "qm", "diel"):
        assert rel(comp[k], want["components"][k]) <= 1e-9, k
    check_potential(g.array("u"), want["potential"], arr["sample_index"], arr["sample_u"],
                    "fp64")


@pytest.mark.parametrize("precision", [0, 1])
def test_npbe161_default_tolerance_energy(ref, workload, precision):
    """The bench's calculation (default etol 1e-6): FP64 and mixed energies
    against the oracle's converged energy."""
    meta, _ = ref
    cfg, mol = workload
    e, g = driver.run_manual(mol, [cfg["dime"]] * 3, mol.center, cfg["pbe"],
                             grid=[cfg["h"]] * 3, precision=precision, mirror=False)
    r = rel(g.energy(1), meta["smol"]["energy_kT"])
    u = g.array("u").ravel()
    pot = meta["smol"]["potential"]
    l2 = rel(np.linalg.norm(u), pot["l2"])
    print(f"npbe161 precision {precision} etol 1e-6: energy rel {r:.2e}, ||u||_2 rel {l2:.2e}, "
          f"stats {g.stats()}")
    assert r <= 1e-6 and l2 <= 1e-6


def test_npbe161_spl2_forces_match_oracle(ref, workload):
    meta, arr = ref
    cfg, mol = workload
    want = meta["spl2"]
    pbe = dict(cfg["pbe"], srfm="spl2", etol=1e-12)
    e, g = driver.run_manual(mol, [cfg["dime"]] * 3, mol.center, pbe, grid=[cfg["h"]] * 3,
                             mirror=False)
    assert rel(g.energy(1), want["energy_kT"]) <= 1e-9
    check_potential(g.array("u"), want["potential"], None, None, "spl2")
    ref_f = arr["forces"]                      # (atoms, [qf, ib, db], xyz)
    fmax = np.abs(ref_f.sum(axis=1)).max()
    worst = 0.0
    for i, a in enumerate(arr["force_atoms"]):
        ok, qf, ib, db = g.force(int(a))
        assert ok
        got = np.stack([qf, ib, db])
        worst = max(worst, float(np.abs(got - ref_f[i]).max() / fmax))
    print(f"npbe161 spl2 companion: energy {g.energy(1)!r} kT against {want['energy_kT']!r}; "
          f"forces of {len(arr['force_atoms'])} atoms within {worst:.2e} of max |F| {fmax:.3e}")
    assert worst <= 1e-8
