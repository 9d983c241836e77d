"""The multigrid hierarchy's interiors are written before they are read.

A new hierarchy clears only the boundary layer of each array (a recycled
pool block's interior may hold anything, as a reused hierarchy's holds the
last solve's values).  With the poison hook every interior starts as NaN; a
kernel that read an interior entry before writing it would spread NaN into
the solution.  The poisoned solve must give the same bits as a clean one, for
linear and nonlinear solves in FP64 and mixed precision.
"""
import ctypes

import numpy as np
import pytest

from apbs_amd import driver, lib, synth

pytestmark = pytest.mark.gpu

SALT = [(1, 0.15, 2.0), (-1, 0.15, 2.0)]


def solve(mol, npbe, precision, poison):
    pbe = dict(pdie=2.0, sdie=78.54, srfm="smol", chgm="spl2", srad=1.4, swin=0.3, sdens=10.0,
               temp=298.15, bcfl="mdh", ions=SALT, npbe=npbe)
    c = driver.Calc(mol, driver.manual_mgparm([65] * 3, mol.center, grid=[0.5] * 3), pbe,
                    precision=precision, mirror=False)
    L = lib()
    L.pmgtest_poison_hierarchy.argtypes = [ctypes.c_void_p, ctypes.c_int]
    assert L.pmgtest_poison_hierarchy(ctypes.cast(c.pmg, ctypes.c_void_p), 1 if poison else 0)
    c.fillco()
    c.solve()
    return c.energy(1), c.array("u"), c.stats()


@pytest.mark.parametrize("npbe,precision", [(False, 0), (False, 1), (True, 0), (True, 1)])
def test_poisoned_hierarchy_gives_the_same_bits(npbe, precision):
    xyz, q, r, _ = synth.globule(300, (11.0,) * 3, 21, net_charge=-3)
    mol = driver.Molecule(xyz, q, r)
    e0, u0, st0 = solve(mol, npbe, precision, False)
    e1, u1, st1 = solve(mol, npbe, precision, True)
    print(f"npbe {npbe} precision {precision}: clean {e0!r} {st0}, poisoned {e1!r} {st1}")
    assert np.isfinite(e1) and np.all(np.isfinite(u1))
    assert e1 == e0
    assert np.array_equal(u1, u0)
    assert st1["iters"] == st0["iters"] and st1["inner"] == st0["inner"]
