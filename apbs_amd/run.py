"""Run an APBS input file on the MI355X PMG solver.

    python -m apbs_amd.run input.in [--precision mixed]

Restates APBS's main loop (src/main.c:528-600 and the PRINT loop :931-960)
for the sections `apbs_amd.nosh` accepts: every ELEC section expands into its
calculations (one for mg-manual; the focusing chain of MGparm_setupAuto for
mg-auto; for mg-para the chain of one processor's partition from
MGparm_setupPara, with the `async` rank or, without one, every processor in
turn on this GPU, whose PRINT values are summed as an MPI run reduces them),
each run as initMG / solveMG / setPartMG / energyMG / forceMG /
writedataMG (src/routines.c:1314-1941, 2620-2990) through the library's Vpmg
C API.  A `bcfl focus` calculation takes over the previous calculation's
Vpmg (routines.c:1459-1477); any other calculation frees it.

The lines the reference prints for energies and forces keep its text and
number formats (energyMG routines.c:1726-1801, forceMG :1803-1941,
printElecEnergy :3117-3195, printElecForce :3549-3821), so the regression
scripts that read APBS output read this output the same way.
"""
import argparse
import ctypes
import os
import sys

import numpy as np

from . import VDT_CHARGE, VDT_DIELX, VDT_DIELY, VDT_DIELZ, VDT_KAPPA, VDT_POT, lib
from . import driver, dx, nosh

VUNIT_KB = driver.VUNIT_KB
VUNIT_NA = driver.VUNIT_NA

_WRITE = {"pot": (VDT_POT, "POTENTIAL (kT/e)"),
          "charge": (VDT_CHARGE, "CHARGE DISTRIBUTION (e)"),
          "dielx": (VDT_DIELX, "X-SHIFTED DIELECTRIC MAP"),
          "diely": (VDT_DIELY, "Y-SHIFTED DIELECTRIC MAP"),
          "dielz": (VDT_DIELZ, "Z-SHIFTED DIELECTRIC MAP"),
          "kappa": (VDT_KAPPA, "KAPPA MAP")}


class CalcResult:
    """What main.c keeps per calculation for the PRINT section."""

    def __init__(self, temp, calcenergy, calcforce):
        self.temp = temp
        self.calcenergy = calcenergy
        self.calcforce = calcforce
        self.energy = 0.0          # kT (totEnergy)
        self.forces = None         # (n, 3, 3): qf, ib, db per atom, or summed (1, 3, 3)


def _pbe_dict(e):
    return {"npbe": e.pbe == "npbe", "pdie": e.pdie, "sdie": e.sdie, "srfm": e.srfm,
            "chgm": e.chgm, "swin": e.swin if e.swin is not None else 0.3,
            "srad": e.srad if e.srad is not None else 0.0,
            "sdens": e.sdens if e.sdens is not None else 10.0, "temp": e.temp,
            "bcfl": e.bcfl, "ions": list(e.ions), "etol": e.etol}


def _mgparm_chain(e, centers, rank):
    """The MGparm objects of one ELEC section's calculations (for mg-para,
    those of processor `rank`)."""
    if e.method in ("mg-manual", "mg-dummy"):
        c = e.gcent.resolve(centers)
        return [driver.manual_mgparm(e.dime, c, grid=e.grid, glen=e.glen if e.grid is None else None,
                                     chgm=e.chgm, nlev=e.nlev or 4,
                                     dummy=e.method == "mg-dummy")]
    if e.method == "mg-para":
        return driver.para_chain(e.dime, e.cglen, e.fglen, e.cgcent.resolve(centers),
                                 e.fgcent.resolve(centers), e.pdime, e.ofrac, rank, e.chgm)
    return driver.auto_chain(e.dime, e.cglen, e.fglen, e.cgcent.resolve(centers),
                             e.fgcent.resolve(centers), e.chgm)


def processor_ranks(inp):
    """(ranks to run, processor count) for the mg-para sections: the `async`
    rank when the sections name one (the reference's asynchronous mode, one
    processor per run), else every rank of the pdime grid in turn."""
    para = [e for e in inp.elecs if e.method == "mg-para"]
    if not para:
        return [0], 1
    nproc = para[0].pdime[0] * para[0].pdime[1] * para[0].pdime[2]
    asyncs = {e.async_rank for e in para}
    if asyncs == {None}:
        return list(range(nproc)), nproc
    return [asyncs.pop()], nproc


def _kj(temp, e):
    """kT -> kJ/mol in energyMG's order (routines.c:1758)."""
    return VUNIT_KB * temp * 1e-3 * VUNIT_NA * e


class Runner:
    def __init__(self, inp, out=sys.stdout, precision=0, mixed_tol=-1.0, outdir=None):
        self.inp = inp
        self.out = out
        self.outdir = outdir      # where `write` files go (default: the input's directory)
        self.precision = precision
        self.mixed_tol = mixed_tol
        self.results = []        # per calculation (of the processor being run)
        self.elec2calc = []      # ELEC section -> index of its last calculation
        self.ranks = [0]
        self.per_rank = []
        self.mols = []
        self.maps = {k: [] for k in nosh.MAP_KINDS}

    def p(self, s=""):
        self.out.write(s + "\n")

    def run(self):
        why = nosh.check_runnable(self.inp)
        if why:
            raise nosh.InputError("this input cannot run on the PMG path:\n  " + "\n  ".join(why))
        for m in self.inp.molecules:
            mol = driver.Molecule.from_pqr(self.inp.path(m.path))
            self.p(f"Got {len(mol)} atoms from {m.path}")
            self.mols.append(mol)
        self.maps = self._load_maps()
        ranks, nproc = processor_ranks(self.inp)
        self.ranks = ranks
        self.per_rank = []     # (results, elec2calc) per processor run
        for r in ranks:
            if nproc > 1:
                self.p("========================================")
                self.p(f"PROCESSOR {r} OF {nproc}")
            self.results, self.elec2calc = [], []
            self._calculations(r)
            self.per_rank.append((self.results, self.elec2calc))
        self.p("----------------------------------------")
        self.p("PRINT STATEMENTS")
        for pr in self.inp.prints:
            if pr.what == "elecenergy":
                self._print_energy(pr)
            elif pr.what == "elecforce":
                self._print_force(pr)
        self.p("----------------------------------------")
        return self.per_rank

    def _calculations(self, rank):
        """main.c's calculation loop for one processor."""
        centers = [m.center for m in self.mols]
        prev = None
        icalc = 0
        for e in self.inp.elecs:
            chain = _mgparm_chain(e, centers, rank)
            pbe = _pbe_dict(e)
            pbe["maps"] = self._maps_of(e)
            mol = self.mols[e.mol - 1]
            for level, mgp in enumerate(chain):
                focus = level > 0 or e.bcfl == "focus"
                finest = level == len(chain) - 1
                self.p("----------------------------------------")
                label = f" ({e.name})" if e.name else ""
                self.p(f"CALCULATION #{icalc + 1}{label}: MULTIGRID")
                if not focus and prev is not None:
                    prev.release()
                    prev = None
                calc = driver.Calc(mol, mgp, pbe, old=prev if focus else None,
                                   energy=e.calcenergy != "no", precision=self.precision,
                                   mixed_tol=self.mixed_tol)
                prev = None
                calc.fillco()
                if e.method == "mg-dummy":
                    # solveMG (routines.c:1653-1668): no solve, the solution
                    # stays the zero array Vpmg_ctor starts it at
                    self.p("  Skipping solve for mg-dummy run; zeroing solution array")
                else:
                    calc.solve()
                calc.set_part_domain()
                res = CalcResult(e.temp, e.calcenergy, e.calcforce)
                self._energy(calc, e, res)
                self._forces(calc, e, res, mol)
                if finest:    # only the finest level of a chain writes (nosh.c:2082-2083)
                    self._write(calc, e)
                self.results.append(res)
                prev = calc
                icalc += 1
            self.elec2calc.append(icalc - 1)
        if prev is not None:
            prev.release()

    def _load_maps(self):
        """The READ maps in input order per kind (loadDielMaps, loadKappaMaps,
        loadChargeMaps, loadPotMaps: routines.c:278-1040): a dielectric entry
        is the (x, y, z)-shifted triple."""
        maps = {k: [] for k in nosh.MAP_KINDS}
        for kind, fmt, paths in self.inp.maps:
            if kind not in maps:
                continue
            loaded = [driver.Map.from_dx(self.inp.path(pth)) for pth in paths]
            for pth, m in zip(paths, loaded):
                n = m.data.shape
                self.p(f"Got {kind} map {pth}: {n[2]} x {n[1]} x {n[0]} points")
            maps[kind].append(loaded if kind == "diel" else loaded[0])
        return maps

    def _maps_of(self, e):
        """Vpmg_fillco's map arguments for one ELEC section (its usemap
        statements, 1-based ids per kind)."""
        out = {}
        for kind, ident in e.usemaps:
            m = self.maps[kind][ident - 1]
            if kind == "diel":
                out["dielx"], out["diely"], out["dielz"] = m
            else:
                out[kind] = m
        return out

    # ------------------------------------------------------------ per calc
    def _energy(self, calc, e, res):
        """energyMG (routines.c:1726-1801)."""
        if e.calcenergy == "no":
            return
        T = e.temp
        res.energy = calc.energy(1)
        self.p("  Total electrostatic energy = %1.12E kJ/mol" % _kj(T, res.energy))
        if e.calcenergy == "comps":
            L = lib()
            qf = L.Vpmg_qfEnergy(calc.pmg, 1)
            qm = L.Vpmg_qmEnergy(calc.pmg, 1)
            diel = L.Vpmg_dielEnergy(calc.pmg, 1)
            self.p("  Fixed charge energy = %g kJ/mol" % (0.5 * _kj(T, qf)))
            self.p("  Mobile charge energy = %g kJ/mol" % _kj(T, qm))
            self.p("  Dielectric energy = %g kJ/mol" % _kj(T, diel))
            self.p("  Per-atom energies:")
            alist = calc.mol.ptr
            for i in range(L.Valist_getNumberAtoms(alist)):
                te = L.Vpmg_qfAtomEnergy(calc.pmg, L.Valist_getAtom(alist, i))
                self.p("      Atom %d:  %1.12E kJ/mol" % (i, 0.5 * _kj(T, te)))

    def _forces(self, calc, e, res, mol):
        """forceMG (routines.c:1803-1941)."""
        if e.calcforce == "no":
            return
        self.p("  Calculating forces...")
        n = len(mol)
        per = np.zeros((n, 3, 3))
        for j in range(n):
            ok, qf, ib, db = calc.force(j)
            if not ok:
                raise RuntimeError(f"force on atom {j} failed (atom too close to the boundary, "
                                   "or a surface method without boundary forces)")
            per[j, 0], per[j, 1], per[j, 2] = qf, ib, db
        T = e.temp
        if e.calcforce == "total":
            tot = np.zeros((3, 3))
            for j in range(n):       # atom-order accumulation (routines.c:1850-1856)
                tot = tot + per[j]
            res.forces = tot[None]
            self.p(f"  Printing net forces for molecule {e.mol} (kJ/mol/A)")
            self.p("  Legend:")
            self.p("    qf  -- fixed charge force")
            self.p("    db  -- dielectric boundary force")
            self.p("    ib  -- ionic boundary force")
            for k, lab in ((0, "qf"), (1, "ib"), (2, "db")):
                self.p("  %s  %4.3e  %4.3e  %4.3e" % ((lab,) + tuple(_kj(T, v) for v in tot[k])))
        else:
            res.forces = per
            self.p(f"  Printing per-atom forces for molecule {e.mol} (kJ/mol/A)")
            self.p("  Legend:")
            self.p("    tot n -- total force for atom n")
            self.p("    qf  n -- fixed charge force for atom n")
            self.p("    db  n -- dielectric boundary force for atom n")
            self.p("    ib  n -- ionic boundary force for atom n")
            for j in range(n):
                s = per[j, 0] + per[j, 1] + per[j, 2]
                self.p("mgF  tot %d  %4.3e  %4.3e  %4.3e" % ((j,) + tuple(_kj(T, v) for v in s)))
                for k, lab in ((0, "qf"), (1, "ib"), (2, "db")):
                    self.p("mgF  %s  %d  %4.3e  %4.3e  %4.3e"
                           % ((lab, j) + tuple(_kj(T, v) for v in per[j, k])))

    def _write(self, calc, e):
        """writedataMG (routines.c:2620-2990) for the dx / gz formats."""
        if not e.writes:
            return
        L = lib()
        pp = calc.pmgp.contents
        nx, ny, nz = pp.nx, pp.ny, pp.nz
        h = (pp.hx, pp.hy, pp.hzed)
        cen = (pp.xcent, pp.ycent, pp.zcent)
        for what, fmt, stem in e.writes:
            vdt, title = _WRITE[what]
            # the shifted dielectric maps sit half a spacing up their axis
            # (routines.c:2812-2878)
            shift = [0.5 * h[d] if what == ("dielx", "diely", "dielz")[d] else 0.0
                     for d in range(3)]
            origin = tuple(cen[d] + shift[d] - 0.5 * ((nx, ny, nz)[d] - 1) * h[d]
                           for d in range(3))
            buf = np.empty(nx * ny * nz)
            if not L.Vpmg_fillArray(calc.pmg, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                    vdt, 0.0, 0, None):
                raise RuntimeError(f"Vpmg_fillArray failed for {what}")
            name = stem + (".dx" if fmt == "dx" else ".dx.gz")
            path = os.path.join(self.outdir, name) if self.outdir else self.inp.path(name)
            self.p(f"  Writing {what} to {path}")
            dx.write_dx(path, buf.reshape(nz, ny, nx), origin, h, title, gz=(fmt == "gz"))

    # ------------------------------------------------------------ PRINT
    def _header(self, pr, word):
        ids = [nosh.resolve_id(i, self.inp.elecs, "ELEC") for i in pr.ids]

        def lab(k):
            nm = self.inp.elecs[k].name
            return f"{k + 1} ({nm}) " if nm else f"{k + 1} "
        s = f"print {word} " + lab(ids[0])
        for op, k in zip(pr.ops, ids[1:]):
            s += f"{op} " + lab(k)
        self.p("\n" + s + "end")
        return ids

    def _print_energy(self, pr):
        """printElecEnergy (routines.c:3117-3195): the signed sum of the
        sections' energies on each processor, then their sum over the
        processors (Vcom_reduce)."""
        ids = self._header(pr, "energy")
        total = 0.0
        for r, (results, elec2calc) in zip(self.ranks, self.per_rank):
            res0 = results[elec2calc[ids[0]]]
            if res0.calcenergy == "no":
                self.p(f"  Didn't calculate energy in Calculation #{elec2calc[ids[0]] + 1}")
                return
            e = VUNIT_KB * 1e-3 * VUNIT_NA * res0.temp * res0.energy
            for op, k in zip(pr.ops, ids[1:]):
                res = results[elec2calc[k]]
                scalar = 1.0 if op == "+" else -1.0
                e += scalar * VUNIT_KB * 1e-3 * VUNIT_NA * res.temp * res.energy
            self.p("  Local net energy (PE %d) = %1.12E kJ/mol" % (r, e))
            total = total + e
        self.p("  Global net ELEC energy = %1.12E kJ/mol" % total)

    def _print_force(self, pr):
        """printElecForce (routines.c:3549-3821)."""
        ids = self._header(pr, "force")
        f = None
        for results, elec2calc in self.per_rank:
            res0 = results[elec2calc[ids[0]]]
            if res0.calcforce == "no":
                self.p(f"  Didn't calculate force in calculation #{elec2calc[ids[0]] + 1}")
                return
            lf = VUNIT_KB * 1e-3 * VUNIT_NA * res0.temp * res0.forces
            for op, k in zip(pr.ops, ids[1:]):
                r = results[elec2calc[k]]
                if r.calcforce != res0.calcforce or r.forces.shape != res0.forces.shape:
                    self.p("  Inconsistent calcforce declarations or force counts")
                    return
                scalar = 1.0 if op == "+" else -1.0
                lf = lf + scalar * VUNIT_KB * 1e-3 * VUNIT_NA * r.temp * r.forces
            f = lf if f is None else f + lf
        if res0.calcforce == "total":
            self.p("  Printing net forces (kJ/mol/A).")
            self.p("  Legend:")
            self.p("    tot -- Total force")
            self.p("    qf  -- Fixed charge force")
            self.p("    db  -- Dielectric boundary force")
            self.p("    ib  -- Ionic boundary force")
            g = f[0]
            tot = g[0] + g[1] + g[2]
            self.p("  tot %1.12E  %1.12E  %1.12E" % tuple(tot))
            self.p("  qf  %1.12E  %1.12E  %1.12E" % tuple(g[0]))
            self.p("  ib  %1.12E  %1.12E  %1.12E" % tuple(g[1]))
            self.p("  db  %1.12E  %1.12E  %1.12E" % tuple(g[2]))
        else:
            self.p("  Printing per-atom forces (kJ/mol/A).")
            self.p("  Legend:")
            self.p("    tot n -- Total force for atom n")
            self.p("    qf  n -- Fixed charge force for atom n")
            self.p("    db  n -- Dielectric boundary force for atom n")
            self.p("    ib  n -- Ionic boundary force for atom n")
            self.p("    tot all -- Total force for system")
            total = np.zeros(3)
            for j, g in enumerate(f):
                self.p("  qf  %d  %1.12E  %1.12E  %1.12E" % ((j,) + tuple(g[0])))
                self.p("  ib  %d  %1.12E  %1.12E  %1.12E" % ((j,) + tuple(g[1])))
                self.p("  db  %d  %1.12E  %1.12E  %1.12E" % ((j,) + tuple(g[2])))
                t = g[2] + g[1] + g[0]
                self.p("  tot %d  %1.12E  %1.12E  %1.12E" % ((j,) + tuple(t)))
                total = total + t
            self.p("  tot all %1.12E  %1.12E  %1.12E" % tuple(total))


def run_file(path, out=sys.stdout, precision=0, mixed_tol=-1.0):
    inp = nosh.parse_file(path)
    r = Runner(inp, out=out, precision=precision, mixed_tol=mixed_tol)
    r.run()
    return r


def main(argv=None):
    ap = argparse.ArgumentParser(description="Run an APBS input file (mg-manual / mg-auto "
                                             "ELEC sections) on the MI355X PMG solver.")
    ap.add_argument("input")
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64")
    ap.add_argument("--mixed-tol", type=float, default=-1.0)
    a = ap.parse_args(argv)
    try:
        run_file(a.input, precision=1 if a.precision == "mixed" else 0, mixed_tol=a.mixed_tol)
    except nosh.InputError as err:
        print(f"apbs_amd.run: {err}", file=sys.stderr)
        return 2
    return 0


if __name__ == "__main__":
    sys.exit(main())
