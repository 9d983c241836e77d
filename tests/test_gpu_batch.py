"""Batched solves (Vpmg_solveBatch) against each calculation solved alone.

A batch shares every V-cycle launch and the coarsest-level inverses among
calculations of one shape; each calculation keeps its own arithmetic, Newton
path and stopping test.  So every calculation's energy, potential and
iteration counts must be bitwise those of Vpmg_solve on its own, for lpbe
and npbe, FP64 and mixed precision, with molecules that take different
numbers of Newton steps and line-search trials.  Lists longer than a batch
and lists mixing shapes are split into runs; the results stay the same.
"""
import numpy as np
import pytest

from apbs_amd import driver, synth

pytestmark = pytest.mark.gpu

SALT = [(1, 0.15, 2.0), (-1, 0.15, 2.0)]


def molecules(k, seed0):
    out = []
    for i in range(k):
        n = 60 + 70 * (i % 4)
        xyz, q, r, _ = synth.globule(n, (6.0 + 1.5 * (i % 4),) * 3, seed0 + i,
                                     net_charge=(-1) ** i * (1 + 2 * i))
        out.append(driver.Molecule(xyz, q, r))
    return out


def calc(mol, dime, npbe, precision, srfm="smol"):
    pbe = dict(pdie=2.0, sdie=78.54, srfm=srfm, chgm="spl2", srad=1.4, swin=0.3, sdens=10.0,
               temp=298.15, bcfl="mdh", ions=SALT, npbe=npbe)
    c = driver.Calc(mol, driver.manual_mgparm([dime] * 3, mol.center, grid=[0.5] * 3), pbe,
                    precision=precision, mirror=False)
    c.fillco()
    return c


def results(c):
    return c.energy(1), c.array("u"), c.stats()


def check_same(alone, batched, label):
    for i, (a, b) in enumerate(zip(alone, batched)):
        ea, ua, sa = a
        eb, ub, sb = b
        print(f"{label} #{i}: alone {ea!r} ({sa['iters']} steps, {sa['inner']} inner), "
              f"batched {eb!r} ({sb['iters']}, {sb['inner']})")
        assert eb == ea, (label, i)
        assert np.array_equal(ub, ua), (label, i)
        assert (sb["iters"], sb["inner"], sb["outer"]) == (sa["iters"], sa["inner"], sa["outer"])


@pytest.mark.parametrize("npbe,precision", [(False, 0), (False, 1), (True, 0), (True, 1)])
def test_batch_equals_alone(npbe, precision):
    mols = molecules(5, 40)
    alone = []
    for m in mols:
        c = calc(m, 49, npbe, precision)
        c.solve()
        alone.append(results(c))
        c.release()
    calcs = [calc(m, 49, npbe, precision) for m in mols]
    assert driver.solve_batch(calcs) == [1] * len(calcs)
    check_same(alone, [results(c) for c in calcs], f"npbe {npbe} precision {precision}")
    iters = {r[2]["iters"] for r in alone}
    if npbe:
        print(f"Newton step counts in the batch: {sorted(iters)}")


def test_long_and_mixed_lists_split_into_runs():
    """11 calculations: two grid sizes and both precisions interleaved, more
    than one batch of one shape."""
    mols = molecules(11, 70)
    specs = [(33 if i % 3 == 0 else 49, i % 2) for i in range(len(mols))]
    alone = []
    for m, (dime, prec) in zip(mols, specs):
        c = calc(m, dime, True, prec)
        c.solve()
        alone.append(results(c))
        c.release()
    calcs = [calc(m, dime, True, prec) for m, (dime, prec) in zip(mols, specs)]
    assert driver.solve_batch(calcs) == [1] * len(calcs)
    check_same(alone, [results(c) for c in calcs], "mixed list")
    # the same shape and precision throughout: 10 calculations, runs of 8 and 2
    calcs = [calc(m, 49, True, 1) for m in mols[:10]]
    alone = []
    for m in mols[:10]:
        c = calc(m, 49, True, 1)
        c.solve()
        alone.append(results(c))
        c.release()
    assert driver.solve_batch(calcs) == [1] * 10
    check_same(alone, [results(c) for c in calcs], "ten of one shape")
