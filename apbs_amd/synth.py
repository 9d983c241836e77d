"""Seeded synthetic molecules for the benchmark workloads.

Atoms are placed on a seeded, jittered FCC lattice inside a sphere or an
ellipsoid with a minimum pair distance of 2.5 A (the requested densities are
above what random sequential placement can reach at that separation); radii are uniform in
[1.2, 2.0] A; charges are -1 (10 %), +1 (10 %) or a uniform partial charge in
[-0.4, 0.4] (80 %).  A requested net charge is reached deterministically:
first by flipping the sign of +1/-1 atoms, then by a uniform shift of the
partial charges.
"""
import numpy as np


def _fcc_sites(semi_axes, d):
    """FCC lattice points (nearest-neighbour distance d) inside the ellipsoid."""
    a = np.asarray(semi_axes, dtype=np.float64)
    c = d * np.sqrt(2.0)                       # cubic cell edge
    m = np.ceil(a / c).astype(int) + 1
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    ii = [np.arange(-m[k], m[k] + 1) for k in range(3)]
    g = np.stack(np.meshgrid(*ii, indexing="ij"), axis=-1).reshape(-1, 1, 3)
    pts = ((g + base[None]) * c).reshape(-1, 3)
    inside = np.sum((pts / a) ** 2, axis=1) <= 1.0
    return pts[inside]


def globule(n, semi_axes, seed, min_dist=2.5, net_charge=None):
    """n atoms inside the ellipsoid with the given semi-axes (A).

    Sites are a seeded random subset of an FCC lattice whose nearest-neighbour
    distance d is the largest (<= 1.5 * min_dist) that still leaves about 5 %
    spare sites; each site is jittered uniformly in a cube of half-width
    (d - min_dist) / (2 sqrt 3), which keeps every pair >= min_dist apart.  When
    even d = min_dist leaves too few sites, d shrinks until n fit and the
    achieved separation d is returned (the requested one cannot be met).
    Returns xyz (n, 3), charge (n,), radius (n,), achieved minimum distance.
    """
    rng = np.random.default_rng(seed)
    d = 1.5 * min_dist
    while True:
        sites = _fcc_sites(semi_axes, d)
        if len(sites) >= 1.05 * n or (d <= min_dist and len(sites) >= n):
            break
        # shrink by 2 %, trying d = min_dist itself before going below it
        d = max(0.98 * d, min_dist) if d > min_dist else 0.98 * d
    if len(sites) < n:
        raise RuntimeError("ellipsoid too small")
    pick = np.sort(rng.choice(len(sites), size=n, replace=False))
    xyz = sites[pick]
    jit = max(0.0, d - min_dist) / (2.0 * np.sqrt(3.0))
    if jit > 0:
        xyz = xyz + rng.uniform(-jit, jit, xyz.shape)
    achieved = min(d, min_dist) if d < min_dist else min_dist
    radius = rng.uniform(1.2, 2.0, n)
    kind = rng.uniform(0, 1, n)
    charge = np.where(kind < 0.1, -1.0, np.where(kind < 0.2, 1.0, 0.0))
    partial = kind >= 0.2
    charge[partial] = rng.uniform(-0.4, 0.4, partial.sum())
    if net_charge is not None:
        target = float(net_charge)
        # flip unit charges towards the target, in index order
        for i in range(n):
            cur = charge.sum()
            if abs(cur - target) < 1.0:
                break
            if cur > target and charge[i] == 1.0:
                charge[i] = -1.0
            elif cur < target and charge[i] == -1.0:
                charge[i] = 1.0
        if partial.any():
            charge[partial] += (target - charge.sum()) / partial.sum()
    return xyz, charge, radius, achieved


# The workload shapes (grid, spacing, molecule) of the benchmark configurations.
CONFIGS = {
    "lpbe65": dict(dime=65, h=0.5, natoms=500, axes=(12, 12, 12), net=None, seed=1,
                   pbe=dict(pdie=2.0, sdie=78.54, srfm="spl2", chgm="spl2", swin=0.3, srad=1.4,
                            sdens=10.0, temp=298.15, bcfl="mdh",
                            ions=[(1, 0.15, 2.0), (-1, 0.15, 2.0)])),
    "npbe161": dict(dime=161, h=0.4, natoms=5000, axes=(28, 24, 20), net=-10, seed=2,
                    pbe=dict(npbe=True, pdie=4.0, sdie=78.54, srfm="smol", chgm="spl2", swin=0.3,
                             srad=1.4, sdens=10.0, temp=298.15, bcfl="mdh",
                             ions=[(1, 0.10, 2.0), (-1, 0.10, 2.0), (1, 0.05, 2.0),
                                   (-1, 0.05, 2.0)])),
    "lpbe193": dict(dime=193, h=0.35, natoms=20000, axes=(30, 15, 15), net=None, seed=3,
                    pbe=dict(pdie=2.0, sdie=78.54, srfm="spl2", chgm="spl2", swin=0.3, srad=1.4,
                             sdens=10.0, temp=298.15, bcfl="mdh",
                             ions=[(1, 0.15, 2.0), (-1, 0.15, 2.0)])),
    # focusing: coarse 97^3 at 1.2 A gives the boundary of a 353^3 grid at 0.3 A
    "focus353": dict(dime=353, h=0.3, coarse_dime=97, coarse_h=1.2, natoms=60000,
                     axes=(50, 45, 40), net=-40, seed=4,
                     pbe=dict(npbe=True, pdie=4.0, sdie=78.54, srfm="smol", chgm="spl2", swin=0.3,
                              srad=1.4, sdens=10.0, temp=298.15, bcfl="mdh",
                              ions=[(1, 0.15, 2.0), (-1, 0.15, 2.0)])),
}


def make(name, natoms=None):
    """(xyz, charge, radius) of a named workload molecule."""
    c = CONFIGS[name]
    n = natoms or c["natoms"]
    xyz, q, r, _ = globule(n, c["axes"], c["seed"], net_charge=c["net"])
    return xyz, q, r


def min_separation(xyz):
    """The smallest distance between two atoms (A)."""
    from scipy.spatial import cKDTree
    d, _ = cKDTree(xyz).query(xyz, k=2)
    return float(d[:, 1].min())


def describe(name):
    """What the generator achieved for a named workload: atoms, net charge and
    the measured minimum pair distance (the requested 2.5 A cannot be met at
    every workload's density; the generator then packs as loosely as fits)."""
    xyz, q, r = make(name)
    return {"atoms": len(q), "net_charge": round(float(q.sum()), 6),
            "min_separation_A": round(min_separation(xyz), 4), "requested_separation_A": 2.5}
