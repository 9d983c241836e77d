"""Precision mode 2 (Vpmg_setPrecision): the mixed-precision solve with the
MDH boundary summed from FP32 terms.

Each atom's term c_a exp(-kappa d) / d is evaluated in FP32 (coordinates
relative to the box centre, hardware reciprocal square root and 2^x) and the
terms are added in FP64 in the FP64 kernel's order.  A term is within about
1e-6 relative of the FP64 term, so the bars are:
  * boundary values against the oracle's FP64 ones: within 1e-5 of the
    largest boundary value on every face (41^3, 80 atoms; and the 161^3
    5,000-atom workload against the FP64 GPU boundary, which is itself within
    32 ulp of the oracle's);
  * the 161^3 npbe workload at the default etol: energy within 1e-6 of the
    oracle's converged energy (the metric's error bar) and of the mode-1
    energy;
  * the solve itself is mode 1's: the same boundary gives bitwise the same
    potential whichever mode solved it.
"""
import json
import os

import numpy as np
import pytest

import orc_driver as od
from apbs_amd import driver, synth

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
SALT = [(1, 0.15, 2.0), (-1, 0.15, 2.0)]


def rel(a, b):
    return abs(a - b) / abs(b)


def face_err(ga, gb):
    return max(float(np.abs(a - b).max() / np.abs(b).max()) for a, b in zip(ga, gb))


@pytest.mark.parametrize("salt", [True, False])
def test_boundary_against_oracle(salt):
    xyz, q, r, _ = synth.globule(80, (6.0,) * 3, 11)
    atoms = (xyz, q, r)
    c = od.mol_center(xyz)
    pbe = dict(pdie=2.0, sdie=78.54, srfm="smol", chgm="spl2", srad=1.4, swin=0.3, sdens=10.0,
               temp=298.15, bcfl="mdh", ions=SALT if salt else [])
    mol = driver.Molecule(*atoms)
    g = driver.Calc(mol, driver.manual_mgparm([41] * 3, c, grid=[0.45] * 3), pbe, precision=2)
    g.fillco()
    o = od.OrcCalc(od.make_params([41] * 3, [0.45] * 3, c, 4, pbe), atoms)
    o.fillco()
    err = face_err(g.boundary(), o.boundary())
    print(f"salt {salt}: FP32-term boundary within {err:.2e} of the largest value")
    assert err <= 1e-5
    # everything else fillco computes is unchanged
    for name in ("epsx", "epsy", "epsz", "kappa", "charge"):
        assert np.array_equal(g.array(name), o.array(name)), name


@pytest.fixture(scope="module")
def workload():
    cfg = synth.CONFIGS["npbe161"]
    mol = driver.Molecule(*synth.make("npbe161"))
    return cfg, mol


def calc(cfg, mol, precision):
    m = driver.manual_mgparm([cfg["dime"]] * 3, mol.center, grid=[cfg["h"]] * 3,
                             chgm=cfg["pbe"]["chgm"])
    return driver.Calc(mol, m, cfg["pbe"], precision=precision, mirror=False)


def test_npbe161_energy(workload):
    cfg, mol = workload
    with open(os.path.join(HERE, "golden", "npbe161.json")) as fh:
        e_oracle = json.load(fh)["smol"]["energy_kT"]
    out = {}
    for prec in (0, 1, 2):
        g = calc(cfg, mol, prec)
        g.fillco()
        bnd = [b.copy() for b in g.boundary()]
        g.solve()
        out[prec] = (g.energy(1), bnd, g.array("u"), g.stats())
        g.release()
    e0, b0, u0, _ = out[0]
    e1, _, u1, _ = out[1]
    e2, b2, u2, s2 = out[2]
    berr = face_err(b2, b0)
    d = u2 - u0
    l2 = float(np.linalg.norm(d) / np.linalg.norm(u0))
    print(f"npbe161: boundary within {berr:.2e}; energy fp64 {e0!r} mixed {e1!r} "
          f"mixed+fp32 boundary {e2!r}; rel to oracle {rel(e2, e_oracle):.2e}, to mixed "
          f"{rel(e2, e1):.2e}; potential rel L2 to fp64 {l2:.2e}; stats {s2}")
    assert berr <= 1e-5
    assert rel(e2, e_oracle) <= 1e-6 and rel(e2, e1) <= 1e-6
    assert l2 <= 1e-5


def test_solve_is_mode_one(workload):
    """Mode 2 changes fillco only: a mode-2 object whose boundary is then
    solved gives the same bits as mode 1 on that boundary (setPrecision(1)
    after fillco keeps the FP32-term boundary)."""
    cfg, mol = workload
    a = calc(cfg, mol, 2)
    a.fillco()
    a.solve()
    b = calc(cfg, mol, 2)
    b.fillco()
    b.set_precision(1)
    b.solve()
    assert a.energy(1) == b.energy(1)
    assert np.array_equal(a.array("u"), b.array("u"))
