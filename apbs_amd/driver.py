"""mg-manual / mg-auto calculations on the MI355X solver.

Restates the calculation sequence of APBS's src/routines.c for one ELEC
block: initMG (routines.c:1314-1600: Vpbe_ctor, Vpmgp_ctor, Vpmg_ctor,
Vpmg_fillco), solveMG (:1635), setPartMG (:1676), energyMG (:1726) and forceMG
(:1803), with the focusing chain of NOsh_setupCalcMGAUTO
(src/generic/nosh.c:1828-2101) done by MGparm_setupAuto in the library.
"""
import ctypes

import numpy as np

from . import (BCFL_FOCUS, BCFL_MAP, BCFL_MDH, BCFL_SDH, BCFL_ZERO, MCT_AUTO, MCT_DUMMY,
               MCT_MANUAL, MCT_PARALLEL,
               NONLIN_LPBE, NONLIN_NPBE,
               PCE_NO, PCE_TOTAL, VCM_BSPL2, VCM_BSPL4, VCM_TRIL, VSM_MOL, VSM_MOLSMOOTH,
               VSM_SPLINE, VSM_SPLINE3, VSM_SPLINE4, VSOL_MG, VSOL_Newton, MGparm, Vpmg, lib)

SRFM = {"mol": VSM_MOL, "smol": VSM_MOLSMOOTH, "spl2": VSM_SPLINE, "spl3": VSM_SPLINE3,
        "spl4": VSM_SPLINE4}
CHGM = {"spl0": VCM_TRIL, "spl2": VCM_BSPL2, "spl4": VCM_BSPL4}
BCFL = {"zero": BCFL_ZERO, "sdh": BCFL_SDH, "mdh": BCFL_MDH, "focus": BCFL_FOCUS,
        "map": BCFL_MAP}

VUNIT_KB = 1.3806581e-23    # src/generic/vunit.h:75
VUNIT_NA = 6.0221367e+23    # src/generic/vunit.h:77


def kT_to_kJmol(e, temp):
    """routines.c:1758-1759"""
    return VUNIT_KB * temp * 1e-3 * VUNIT_NA * e


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Molecule:
    """A Valist built from arrays (xyz (n,3), charge (n,), radius (n,))."""

    def __init__(self, xyz, charge, radius):
        L = lib()
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        self.q = np.ascontiguousarray(charge, dtype=np.float64)
        self.r = np.ascontiguousarray(radius, dtype=np.float64)
        self.ptr = L.Valist_ctor()
        if not L.Valist_setAtoms(self.ptr, len(self.q), _dp(self.xyz), _dp(self.r), _dp(self.q)):
            raise ValueError("Valist_setAtoms failed")

    @classmethod
    def from_pqr(cls, path):
        """A molecule read by the library's Valist_readPQRFile
        (valist.c:606-723 in the reference)."""
        L = lib()
        ptr = L.Valist_ctor()
        if not L.Valist_readPQRFile(ptr, str(path).encode()):
            L.Valist_dtor(ctypes.byref(ptr))
            raise ValueError(f"cannot read PQR file {path}")
        n = L.Valist_getNumberAtoms(ptr)
        atoms = [L.Valist_getAtom(ptr, i).contents for i in range(n)]
        self = cls.__new__(cls)
        self.xyz = np.array([a.position[:] for a in atoms], dtype=np.float64).reshape(-1, 3)
        self.q = np.array([a.charge for a in atoms], dtype=np.float64)
        self.r = np.array([a.radius for a in atoms], dtype=np.float64)
        self.ptr = ptr
        return self

    def __len__(self):
        return len(self.q)

    @property
    def center(self):
        return np.array(self.ptr.contents.center[:])

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().Valist_dtor(ctypes.byref(self.ptr))
            self.ptr = None


class Map:
    """A Vgrid (Vgrid_ctor) over data (nz, ny, nx), x fastest in memory, with
    its first point at origin and spacing h per axis.  The array is kept
    alive with the map."""

    def __init__(self, data, origin, spacing):
        self.data = np.ascontiguousarray(data, dtype=np.float64)
        if self.data.ndim != 3:
            raise ValueError("a map is a 3-d array (nz, ny, nx)")
        nz, ny, nx = self.data.shape
        self.origin = tuple(float(v) for v in origin)
        self.spacing = tuple(float(v) for v in spacing)
        self.ptr = lib().Vgrid_ctor(nx, ny, nz, *self.spacing, *self.origin, _dp(self.data))
        if not self.ptr:
            raise ValueError("Vgrid_ctor failed")

    @classmethod
    def from_dx(cls, path):
        from . import dx
        data, origin, spacing = dx.read_dx(str(path))
        return cls(data, origin, spacing)

    def value(self, pt):
        """Vgrid_value: the trilinear value at pt, or None off the map."""
        p = np.ascontiguousarray(pt, dtype=np.float64)
        v = np.zeros(1)
        return float(v[0]) if lib().Vgrid_value(self.ptr, _dp(p), _dp(v)) else None

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().Vgrid_dtor(ctypes.byref(self.ptr))
            self.ptr = None


MAP_KEYS = ("dielx", "diely", "dielz", "kappa", "pot", "charge")


class Calc:
    """One Vpmg with its Vpbe / Vpmgp / MGparm (one mg-manual calculation)."""

    def __init__(self, mol, mgp, pbe, old=None, energy=True, precision=0, mixed_tol=-1.0,
                 mirror=True):
        L = lib()
        self.mol = mol
        self.pbe_opts = pbe
        ions = pbe.get("ions", [])
        nion = len(ions)
        self._ic = np.array([c for (_, c, _) in ions] + [0.0], dtype=np.float64)
        self._ir = np.array([r for (_, _, r) in ions] + [0.0], dtype=np.float64)
        self._iq = np.array([q for (q, _, _) in ions] + [0.0], dtype=np.float64)
        srfm = SRFM[pbe["srfm"]]
        sparm = pbe.get("swin", 0.3) if srfm == VSM_SPLINE else pbe.get("srad", 1.4)
        focus = 1 if old is not None else 0
        self.temp = pbe.get("temp", 298.15)
        self.pbe = L.Vpbe_ctor(mol.ptr, nion, _dp(self._ic), _dp(self._ir), _dp(self._iq),
                               self.temp, pbe["pdie"], pbe["sdie"], sparm, focus,
                               pbe.get("sdens", 10.0), 0.0, 0.0, 0.0, 0.0)
        if not self.pbe:
            raise ValueError("Vpbe_ctor failed")
        self.mgparm = mgp
        mgp.contents.nonlintype = NONLIN_NPBE if pbe.get("npbe") else NONLIN_LPBE
        mgp.contents.method = VSOL_Newton if pbe.get("npbe") else VSOL_MG
        if pbe.get("etol"):
            mgp.contents.etol = pbe["etol"]
            mgp.contents.setetol = 1
        self.pmgp = L.Vpmgp_ctor(mgp)
        bcfl = BCFL_FOCUS if focus else BCFL[pbe.get("bcfl", "mdh")]
        self.pmgp.contents.bcfl = bcfl
        self.pmgp.contents.xcent = mgp.contents.center[0]
        self.pmgp.contents.ycent = mgp.contents.center[1]
        self.pmgp.contents.zcent = mgp.contents.center[2]
        oldp = old.pmg if old is not None else None
        self.pmg = L.Vpmg_ctor(self.pmgp, self.pbe, focus, oldp, mgp,
                               PCE_TOTAL if energy else PCE_NO)
        if old is not None:
            old.pmg = None       # consumed by Vpmg_ctor (vpmg.c:396)
            old.release()
        if not self.pmg:
            raise RuntimeError("Vpmg_ctor failed")
        if precision:
            self.set_precision(precision, mixed_tol)
        L.Vpmg_setHostMirror(self.pmg, 1 if mirror else 0)
        self.srfm = srfm
        self.chgm = CHGM[pbe["chgm"]]

    def set_precision(self, precision, mixed_tol=-1.0):
        """Vpmg_setPrecision: 0 FP64, 1 mixed, 2 mixed with the MDH boundary
        from FP32 terms (takes effect at the next fillco)."""
        if not lib().Vpmg_setPrecision(self.pmg, int(precision), float(mixed_tol)):
            raise ValueError(f"precision {precision}: 0 FP64, 1 mixed, 2 mixed + FP32 MDH terms")

    def fillco(self):
        """Vpmg_fillco; pbe["maps"] may name Map objects under MAP_KEYS
        (the usemap statements of routines.c:1482-1561)."""
        maps = self.pbe_opts.get("maps") or {}
        unknown = set(maps) - set(MAP_KEYS)
        if unknown:
            raise ValueError(f"unknown map kinds {sorted(unknown)}")
        args = []
        for k in MAP_KEYS:
            m = maps.get(k)
            args += [0, None] if m is None else [1, m.ptr]
        ok = lib().Vpmg_fillco(self.pmg, self.srfm, self.pbe_opts.get("swin", 0.3), self.chgm,
                               *args)
        if not ok:
            raise RuntimeError("Vpmg_fillco failed")

    def solve(self):
        if not lib().Vpmg_solve(self.pmg):
            raise RuntimeError("Vpmg_solve failed")

    def set_part_domain(self):
        """setPartMG (routines.c:1676-1724): the whole grid, or for an
        mg-para level the processor's disjoint partition."""
        c = self.mgparm.contents
        if c.type == MCT_PARALLEL:
            lo = np.array([c.partDisjCenter[j] - 0.5 * c.partDisjLength[j] for j in range(3)])
            hi = np.array([c.partDisjCenter[j] + 0.5 * c.partDisjLength[j] for j in range(3)])
        else:
            lo = np.array([c.center[j] - 0.5 * c.glen[j] for j in range(3)])
            hi = np.array([c.center[j] + 0.5 * c.glen[j] for j in range(3)])
        bf = (ctypes.c_int * 6)(*c.partDisjOwnSide)
        lib().Vpmg_setPart(self.pmg, _dp(lo), _dp(hi), bf)

    def energy(self, ext=1):
        return lib().Vpmg_energy(self.pmg, ext)

    def energy_kJ(self, ext=1):
        return kT_to_kJmol(self.energy(ext), self.temp)

    def components(self, ext=1):
        L = lib()
        return {"qf": L.Vpmg_qfEnergy(self.pmg, ext), "qm": L.Vpmg_qmEnergy(self.pmg, ext),
                "diel": L.Vpmg_dielEnergy(self.pmg, ext)}

    def force(self, atom):
        L = lib()
        qf, ib, db = np.zeros(3), np.zeros(3), np.zeros(3)
        L.Vpmg_qfForce(self.pmg, _dp(qf), atom, self.chgm)
        rc1 = L.Vpmg_ibForce(self.pmg, _dp(ib), atom, self.srfm)
        rc2 = L.Vpmg_dbForce(self.pmg, _dp(db), atom, self.srfm)
        return rc1 and rc2, qf, ib, db

    def array(self, which):
        names = {"u": 0, "epsx": 1, "epsy": 2, "epsz": 3, "kappa": 4, "charge": 5, "pvec": 6}
        p = self.pmgp.contents
        out = np.empty(p.nx * p.ny * p.nz)
        if not lib().Vpmg_copyArray(self.pmg, names[which], _dp(out), out.size):
            raise RuntimeError("Vpmg_copyArray failed")
        return out.reshape(p.nz, p.ny, p.nx)

    def boundary(self):
        p = self.pmgp.contents
        gx = np.empty(2 * p.ny * p.nz)
        gy = np.empty(2 * p.nx * p.nz)
        gz = np.empty(2 * p.nx * p.ny)
        L = lib()
        L.Vpmg_copyArray(self.pmg, 7, _dp(gx), gx.size)
        L.Vpmg_copyArray(self.pmg, 8, _dp(gy), gy.size)
        L.Vpmg_copyArray(self.pmg, 9, _dp(gz), gz.size)
        return gx, gy, gz

    def stats(self):
        out = np.zeros(8)
        lib().Vpmg_getStats(self.pmg, _dp(out), 8)
        return {"iters": int(out[0]), "inner": int(out[1]), "nlev": int(out[2]),
                "relres": out[3], "t_solve": out[4], "t_fillco": out[5], "t_build": out[6],
                "outer": int(out[7])}

    def kernel_timing(self, on=True):
        """Start (zeroing the accumulators) or stop HIP-event kernel timing."""
        lib().Vpmg_setKernelTiming(self.pmg, 1 if on else 0)

    def kernel_times(self):
        out = np.zeros(8)
        lib().Vpmg_getKernelTiming(self.pmg, _dp(out), 8)
        return {"smoother_ms": float(out[0]), "smoother_launches": int(out[1]),
                "vcycle_ms": float(out[2]), "vcycles": int(out[3]), "word": int(out[4]),
                "points": int(out[5]), "words_per_point": float(out[6]),
                "vector_form": bool(out[7])}

    def release(self):
        L = lib()
        if getattr(self, "pmg", None):
            L.Vpmg_dtor(ctypes.byref(self.pmg))
            self.pmg = None
        if getattr(self, "pmgp", None):
            L.Vpmgp_dtor(ctypes.byref(self.pmgp))
            self.pmgp = None
        if getattr(self, "pbe", None):
            p = ctypes.c_void_p(self.pbe)
            L.Vpbe_dtor(ctypes.byref(p))
            self.pbe = None

    def __del__(self):
        self.release()


def solve_batch(calcs):
    """Vpmg_solveBatch over filled Calc objects: runs of calculations of one
    shape, method and precision share their V-cycle launches.  Returns the
    per-calculation Vpmg_solve results."""
    n = len(calcs)
    arr = (ctypes.POINTER(Vpmg) * n)(*[c.pmg for c in calcs])
    ok = (ctypes.c_int * n)()
    lib().Vpmg_solveBatch(arr, n, ok)
    return list(ok)


def _new_mgparm(kind):
    return lib().MGparm_ctor(kind)


def manual_mgparm(dime, center, grid=None, glen=None, chgm="spl2", nlev=4, dummy=False):
    """An mg-manual MGparm (mg-dummy with dummy=True: the same keywords, no
    nlev/dime adjustment, mgparm.c:205-270)."""
    L = lib()
    m = _new_mgparm(MCT_DUMMY if dummy else MCT_MANUAL)
    c = m.contents
    for j in range(3):
        c.dime[j] = int(dime[j])
        c.center[j] = float(center[j])
        if grid is not None:
            c.grid[j] = float(grid[j])
        if glen is not None:
            c.glen[j] = float(glen[j])
    c.setdime = 1
    c.nlev = nlev
    c.setgcent = 1
    c.setgrid = 1 if grid is not None else 0
    c.setglen = 1 if glen is not None else 0
    c.chgm = CHGM[chgm]
    c.setchgm = 1
    if not L.MGparm_check(m):
        raise ValueError("MGparm_check failed")
    return m


def run_manual(mol, dime, center, pbe, grid=None, glen=None, precision=0, mixed_tol=-1.0,
               mirror=True):
    """One mg-manual ELEC block: returns (energy in kJ/mol, Calc)."""
    m = manual_mgparm(dime, center, grid, glen, pbe["chgm"])
    c = Calc(mol, m, pbe, precision=precision, mixed_tol=mixed_tol, mirror=mirror)
    c.fillco()
    c.solve()
    c.set_part_domain()
    return c.energy_kJ(1), c


def auto_chain(dime, cglen, fglen, ccenter, fcenter, chgm="spl2"):
    """The focusing chain of one mg-auto block as checked MGparm objects."""
    L = lib()
    e = _new_mgparm(MCT_AUTO)
    c = e.contents
    for j in range(3):
        c.dime[j] = int(dime[j])
        c.cglen[j], c.fglen[j] = float(cglen[j]), float(fglen[j])
        c.ccenter[j], c.fcenter[j] = float(ccenter[j]), float(fcenter[j])
    c.setdime = c.setcglen = c.setfglen = c.setcgcent = c.setfgcent = 1
    c.chgm = CHGM[chgm]
    c.setchgm = 1
    if not L.MGparm_check(e):
        raise ValueError("MGparm_check failed")
    outs = [_new_mgparm(MCT_MANUAL) for _ in range(20)]
    arr = (ctypes.POINTER(MGparm) * 20)(*outs)
    nf = L.MGparm_setupAuto(e, arr, 20)
    if nf <= 0:
        raise ValueError("MGparm_setupAuto failed")
    for k in range(nf, 20):
        L.MGparm_dtor(ctypes.byref(outs[k]))
    L.MGparm_dtor(ctypes.byref(e))
    return outs[:nf]


def para_chain(dime, cglen, fglen, ccenter, fcenter, pdime, ofrac, rank, chgm="spl2"):
    """The focusing chain of processor `rank` of one mg-para block
    (NOsh_setupCalcMGPARA, nosh.c:2103-2394) as checked MGparm objects."""
    L = lib()
    e = _new_mgparm(MCT_PARALLEL)
    c = e.contents
    for j in range(3):
        c.dime[j] = int(dime[j])
        c.cglen[j], c.fglen[j] = float(cglen[j]), float(fglen[j])
        c.ccenter[j], c.fcenter[j] = float(ccenter[j]), float(fcenter[j])
        c.pdime[j] = int(pdime[j])
    c.ofrac = float(ofrac)
    c.setdime = c.setcglen = c.setfglen = c.setcgcent = c.setfgcent = 1
    c.setpdime = c.setofrac = 1
    setattr(c, "async", int(rank))     # a Python keyword, so not c.async
    c.setasync = 1
    c.chgm = CHGM[chgm]
    c.setchgm = 1
    if not L.MGparm_check(e):
        raise ValueError("MGparm_check failed")
    outs = [_new_mgparm(MCT_MANUAL) for _ in range(20)]
    arr = (ctypes.POINTER(MGparm) * 20)(*outs)
    nf = L.MGparm_setupPara(e, int(rank), arr, 20)
    if nf <= 0:
        raise ValueError("MGparm_setupPara failed")
    for k in range(nf, 20):
        L.MGparm_dtor(ctypes.byref(outs[k]))
    L.MGparm_dtor(ctypes.byref(e))
    return outs[:nf]


def run_auto(mol, dime, cglen, fglen, ccenter, fcenter, pbe, precision=0, mixed_tol=-1.0,
             energy=True):
    """One mg-auto ELEC block: returns (per-level energies kJ/mol, last Calc)."""
    chain = auto_chain(dime, cglen, fglen, ccenter, fcenter, pbe["chgm"])
    energies = []
    calc = None
    for m in chain:
        calc = Calc(mol, m, pbe, old=calc, energy=energy, precision=precision,
                    mixed_tol=mixed_tol)
        calc.fillco()
        calc.solve()
        calc.set_part_domain()
        energies.append(calc.energy_kJ(1) if energy else None)
    return energies, calc
