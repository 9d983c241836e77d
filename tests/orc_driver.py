"""ctypes driver for the oracle (oracle/liborc.so) -- test infrastructure only.

Besides the binding, this module restates the input-level semantics the
regression examples depend on, so golden values can be reproduced from the
reference's own inputs:

* nlev derivation of ``MGparm_check`` (src/generic/mgparm.c:265-320);
* mg-manual grid set-up (src/generic/nosh.c:1760-1826);
* mg-auto focusing chain (src/generic/nosh.c:1828-2101);
* the per-calculation sequence initMG -> solveMG -> setPartMG -> energyMG ->
  forceMG (src/routines.c:1314-1941);
* kT -> kJ/mol conversion (src/routines.c:1758-1759).
"""
import ctypes
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ORC_SO = os.path.join(ROOT, "oracle", "liborc.so")

SRFM = {"mol": 0, "smol": 1, "spl2": 2, "spl3": 3, "spl4": 4}
CHGM = {"spl0": 0, "spl2": 1, "spl4": 2}
BCFL = {"zero": 0, "sdh": 1, "mdh": 2, "focus": 4, "map": 6}
MAPS = {"dielx": 0, "diely": 1, "dielz": 2, "kappa": 3, "pot": 4, "charge": 5}

# src/generic/vunit.h:75-80
VUNIT_KB = 1.3806581e-23
VUNIT_NA = 6.0221367e+23
VREDFRAC = 0.25    # src/generic/vhal.h:392
VMGNLEV = 4        # src/generic/vhal.h:387
VSMALL = 1.0e-9


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("dime", ctypes.c_int * 3),
        ("grid", ctypes.c_double * 3),
        ("center", ctypes.c_double * 3),
        ("nlev", ctypes.c_int),
        ("pbetype", ctypes.c_int),
        ("nion", ctypes.c_int),
        ("ionq", ctypes.c_double * 10),
        ("ionc", ctypes.c_double * 10),
        ("ionr", ctypes.c_double * 10),
        ("pdie", ctypes.c_double),
        ("sdie", ctypes.c_double),
        ("srad", ctypes.c_double),
        ("swin", ctypes.c_double),
        ("sdens", ctypes.c_double),
        ("temp", ctypes.c_double),
        ("srfm", ctypes.c_int),
        ("chgm", ctypes.c_int),
        ("bcfl", ctypes.c_int),
        ("etol", ctypes.c_double),
        ("coarse_smoother", ctypes.c_int),
        ("deep_coarse", ctypes.c_int),
        ("newton_itmax", ctypes.c_int),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORC_SO):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(ORC_SO)
        dp = ctypes.POINTER(ctypes.c_double)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.POINTER(OrcParams), ctypes.c_int, dp, dp, dp,
                                 ctypes.c_void_p, ctypes.c_int]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        for f in ("orc_fillco", "orc_solve"):
            getattr(L, f).argtypes = [ctypes.c_void_p]
            getattr(L, f).restype = ctypes.c_int
        for f in ("orc_energy", "orc_qf_energy", "orc_qm_energy", "orc_diel_energy"):
            getattr(L, f).argtypes = [ctypes.c_void_p, ctypes.c_int]
            getattr(L, f).restype = ctypes.c_double
        L.orc_force.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, dp, dp]
        L.orc_force.restype = ctypes.c_int
        L.orc_set_part_domain.argtypes = [ctypes.c_void_p]
        L.orc_get_array.argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
        L.orc_get_boundary.argtypes = [ctypes.c_void_p, dp, dp, dp]
        L.orc_stats.argtypes = [ctypes.c_void_p, dp]
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_set_map.argtypes = [ctypes.c_void_p, ctypes.c_int, ip, dp, dp, dp]
        L.orc_set_map.restype = ctypes.c_int
        L.orc_map_value.argtypes = [ip, dp, dp, dp, dp, dp]
        L.orc_map_value.restype = ctypes.c_int
        _lib = L
    return _lib


def dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


# ----------------------------------------------------------------- molecules

def read_pqr(path):
    """Atoms of a PQR file: the last five fields of ATOM/HETATM records are
    x y z charge radius (src/generic/valist.c:606-723)."""
    xyz, q, r = [], [], []
    with open(path) as fh:
        for line in fh:
            if not (line.startswith("ATOM") or line.startswith("HETATM")):
                continue
            tok = line.split()
            x, y, z, c, rad = (float(t) for t in tok[-5:])
            xyz.append((x, y, z))
            q.append(c)
            r.append(rad)
    return (np.ascontiguousarray(xyz, dtype=np.float64), np.asarray(q, dtype=np.float64),
            np.asarray(r, dtype=np.float64))


def mol_center(xyz):
    """Valist center: bounding-box midpoint of atom centres (valist.c:908-910)."""
    return 0.5 * (xyz.max(axis=0) + xyz.min(axis=0))


# --------------------------------------------------------------- input rules

def check_nlev(dime, nlev_user=4):
    """MGparm_check (mgparm.c:265-320): keep nlev if dime-1 == 2^(nlev+1),
    else the number of times dime-1 halves, minus one.  Returns (nlev, dime)."""
    tnlev, tdime = [], []
    for d in dime:
        ti = d - 1
        if ti == 2 ** (nlev_user + 1):
            tnlev.append(nlev_user)
            tdime.append(d)
            continue
        n = 0
        while ti % 2 == 0:
            n += 1
            ti = int(math.ceil(0.5 * ti))
        n -= 1
        if d > 65 and n < VMGNLEV:
            ti = int(d / 2 ** (VMGNLEV + 1))
            if ti < 1:
                ti = 1
            d = ti * 2 ** (VMGNLEV + 1) + 1
            n = 4
        tnlev.append(n)
        tdime.append(d)
    return min(tnlev), tdime


def auto_chain(dime, cglen, fglen, ccenter, fcenter):
    """NOsh_setupCalcMGAUTO (nosh.c:1828-2101): list of (grid, glen, center)."""
    cgrid = [cglen[j] / (dime[j] - 1) for j in range(3)]
    fgrid = [fglen[j] / (dime[j] - 1) for j in range(3)]
    tn = []
    for j in range(3):
        if fgrid[j] / cgrid[j] < VREDFRAC:
            redfrac = fgrid[j] / cgrid[j]
            td = math.log(redfrac) / math.log(VREDFRAC)
            tn.append(int(math.ceil(td)) + 1)
        else:
            tn.append(2)
    nfocus = max(tn)
    redrat = [(fgrid[j] / cgrid[j]) ** (1.0 / (nfocus - 1.0)) for j in range(3)]
    out = []
    for ifocus in range(nfocus):
        if ifocus == 0:
            grid = list(cgrid)
            glen = list(cglen)
            center = list(ccenter)
        else:
            pg, pl, pc = out[-1]
            grid = [redrat[j] * pg[j] for j in range(3)]
            glen = [redrat[j] * pl[j] for j in range(3)]
            center = list(fcenter)
            for j in range(3):
                dofix = False
                minf = center[j] - 0.5 * glen[j]
                minc = pc[j] - 0.5 * pl[j]
                d = minc - minf
                if d >= VSMALL:
                    if ifocus == nfocus - 1:
                        raise ValueError("fine mesh fell off the coarser meshes")
                    center[j] += d
                    dofix = True
                maxf = center[j] + 0.5 * glen[j]
                maxc = pc[j] + 0.5 * pl[j]
                d = maxf - maxc
                if d >= VSMALL:
                    if ifocus == nfocus - 1 or dofix:
                        raise ValueError("fine mesh fell off the coarser meshes")
                    center[j] -= d
        out.append((grid, glen, center))
    return out


def kT_to_kJmol(e_kT, temp):
    return VUNIT_KB * temp * 1e-3 * VUNIT_NA * e_kT


# ------------------------------------------------------------------- driver

def make_params(dime, grid, center, nlev, pbe):
    p = OrcParams()
    for j in range(3):
        p.dime[j] = int(dime[j])
        p.grid[j] = float(grid[j])
        p.center[j] = float(center[j])
    p.nlev = int(nlev)
    p.pbetype = 1 if pbe.get("npbe") else 0
    ions = pbe.get("ions", [])
    p.nion = len(ions)
    for i, (q, c, r) in enumerate(ions):
        p.ionq[i], p.ionc[i], p.ionr[i] = q, c, r
    p.pdie = pbe["pdie"]
    p.sdie = pbe["sdie"]
    p.srad = pbe.get("srad", 1.4)
    p.swin = pbe.get("swin", 0.3)
    p.sdens = pbe.get("sdens", 10.0)
    p.temp = pbe.get("temp", 298.15)
    p.srfm = SRFM[pbe["srfm"]]
    p.chgm = CHGM[pbe["chgm"]]
    p.bcfl = BCFL[pbe.get("bcfl", "mdh")]
    p.etol = pbe.get("etol", 0.0)
    p.coarse_smoother = pbe.get("coarse_smoother", 0)
    p.deep_coarse = pbe.get("deep_coarse", 0)
    p.newton_itmax = pbe.get("newton_itmax", 0)
    return p


class OrcCalc:
    """One Vpmg-like object of the oracle."""

    def __init__(self, params, atoms, old=None, energy_flag=1):
        xyz, q, r = atoms
        self.atoms = atoms
        self.params = params
        self.n = params.dime[0] * params.dime[1] * params.dime[2]
        h = lib().orc_create(ctypes.byref(params), len(q), dptr(xyz), dptr(r), dptr(q),
                             old.handle if old is not None else None, energy_flag)
        if old is not None:
            old.handle = None   # consumed, like vpmg.c:396
        if not h:
            raise RuntimeError("orc_create failed")
        self.handle = h

    def fillco(self):
        assert lib().orc_fillco(self.handle) == 1

    def try_fillco(self):
        return lib().orc_fillco(self.handle)

    def set_map(self, kind, data, origin, spacing):
        """A usemap map: data (nz, ny, nx), first point origin, spacing h."""
        data = np.ascontiguousarray(data, dtype=np.float64)
        n = (ctypes.c_int * 3)(data.shape[2], data.shape[1], data.shape[0])
        h = np.asarray(spacing, dtype=np.float64)
        o = np.asarray(origin, dtype=np.float64)
        assert lib().orc_set_map(self.handle, MAPS[kind], n, dptr(h), dptr(o), dptr(data)) == 1

    def solve(self):
        assert lib().orc_solve(self.handle) == 1

    def set_part_domain(self):
        lib().orc_set_part_domain(self.handle)

    def energy(self, ext=1):
        return lib().orc_energy(self.handle, ext)

    def components(self, ext=1):
        L = lib()
        return {"qf": L.orc_qf_energy(self.handle, ext), "qm": L.orc_qm_energy(self.handle, ext),
                "diel": L.orc_diel_energy(self.handle, ext)}

    def force(self, atom):
        qf, ib, db = (np.zeros(3) for _ in range(3))
        rc = lib().orc_force(self.handle, atom, dptr(qf), dptr(ib), dptr(db))
        return rc, qf, ib, db

    def array(self, which):
        names = {"u": 0, "epsx": 1, "epsy": 2, "epsz": 3, "kappa": 4, "charge": 5, "pvec": 6}
        out = np.empty(self.n)
        lib().orc_get_array(self.handle, names[which], dptr(out))
        return out.reshape(self.params.dime[2], self.params.dime[1], self.params.dime[0])

    def boundary(self):
        d = self.params.dime
        gx = np.empty(2 * d[1] * d[2])
        gy = np.empty(2 * d[0] * d[2])
        gz = np.empty(2 * d[0] * d[1])
        lib().orc_get_boundary(self.handle, dptr(gx), dptr(gy), dptr(gz))
        return gx, gy, gz

    def stats(self):
        out = np.zeros(8)
        lib().orc_stats(self.handle, dptr(out))
        return {"iters": int(out[0]), "inner": int(out[1]), "nlev": int(out[2]),
                "relres": out[3], "zkappa2": out[4], "zmagic": out[5], "xkappa": out[6],
                "ionstr": out[7]}

    def __del__(self):
        if getattr(self, "handle", None):
            lib().orc_destroy(self.handle)
            self.handle = None


def map_value(data, origin, spacing, pt):
    """Vgrid_value of the oracle: the value at pt, or None off the map."""
    data = np.ascontiguousarray(data, dtype=np.float64)
    n = (ctypes.c_int * 3)(data.shape[2], data.shape[1], data.shape[0])
    h = np.asarray(spacing, dtype=np.float64)
    o = np.asarray(origin, dtype=np.float64)
    p = np.asarray(pt, dtype=np.float64)
    v = np.zeros(1)
    ok = lib().orc_map_value(n, dptr(h), dptr(o), dptr(data), dptr(p), dptr(v))
    return float(v[0]) if ok else None


def run_manual(atoms, dime, center, pbe, grid=None, glen=None, nlev_user=4):
    """One mg-manual ELEC block: returns (energy kJ/mol, calc)."""
    nlev, dime = check_nlev(dime, nlev_user)
    if grid is None:
        grid = [glen[j] / (dime[j] - 1) for j in range(3)]
    p = make_params(dime, grid, center, nlev, pbe)
    c = OrcCalc(p, atoms)
    c.fillco()
    c.solve()
    c.set_part_domain()
    return kT_to_kJmol(c.energy(1), p.temp), c


def run_auto(atoms, dime, cglen, fglen, ccenter, fcenter, pbe, nlev_user=4, keep_last=True):
    """One mg-auto ELEC block: returns (per-level energies kJ/mol, last calc)."""
    nlev, dime = check_nlev(dime, nlev_user)
    chain = auto_chain(dime, cglen, fglen, ccenter, fcenter)
    energies = []
    calc = None
    for i, (grid, glen, center) in enumerate(chain):
        pb = dict(pbe)
        if i > 0:
            pb["bcfl"] = "focus"
        p = make_params(dime, grid, center, nlev, pb)
        calc = OrcCalc(p, atoms, old=calc, energy_flag=1)
        calc.fillco()
        calc.solve()
        calc.set_part_domain()
        energies.append(kT_to_kJmol(calc.energy(1), p.temp))
    return energies, calc
