"""APBS input files: the READ / ELEC / APOLAR / PRINT / QUIT sections.

Follows the grammar of APBS's NOsh parser (NOsh_parseInput,
src/generic/nosh.c:517-600; NOsh_parseREAD :974-1023 with its READ_* helpers
:602-972; NOsh_parsePRINT :1025-1189; NOsh_parseELEC :1191-1322;
NOsh_parseAPOLAR :1324) and of the MG and PBE keyword parsers it delegates to
(MGparm_parseToken, src/generic/mgparm.c:916; PBEparm_parseToken,
src/generic/pbeparm.c:1215): whitespace separated, case-insensitive keywords,
`#` comments to the end of the line, each section closed by `end`.

Only the parts the PMG path runs are accepted for execution (READ `mol pqr`;
ELEC `mg-manual` / `mg-auto`; PRINT `elecEnergy` / `elecForce`); other
sections and keywords are parsed and reported, so an input using them fails
with a message that names the keyword instead of a wrong result.  APOLAR
sections are parsed and skipped: their print statements produce no output.
"""
import os
from dataclasses import dataclass, field


class InputError(ValueError):
    """A malformed input file, or a keyword this build does not run."""


# ------------------------------------------------------------------ tokens

def tokenize(text):
    """Whitespace tokens with `#` comments removed (NOsh reads with
    Vio_scanf("%s") from a socket whose comment character is '#')."""
    out = []
    for line in text.splitlines():
        line = line.split("#", 1)[0]
        out.extend(line.split())
    return out


class _Tokens:
    def __init__(self, toks):
        self.toks = toks
        self.i = 0

    def peek(self):
        return self.toks[self.i] if self.i < len(self.toks) else None

    def next(self, what="a token"):
        if self.i >= len(self.toks):
            raise InputError(f"unexpected end of input while reading {what}")
        t = self.toks[self.i]
        self.i += 1
        return t

    def float(self, what):
        t = self.next(what)
        try:
            return float(t)
        except ValueError:
            raise InputError(f"expected a number for {what}, got {t!r}") from None

    def int(self, what):
        t = self.next(what)
        try:
            return int(t)
        except ValueError:
            raise InputError(f"expected an integer for {what}, got {t!r}") from None

    def floats(self, n, what):
        return tuple(self.float(what) for _ in range(n))

    def ints(self, n, what):
        return tuple(self.int(what) for _ in range(n))


# ------------------------------------------------------------------ objects

@dataclass
class Molecule:
    fmt: str
    path: str


@dataclass
class Center:
    """A grid centre: the centre of molecule `mol` (1-based) or coordinates."""
    mol: int = 0
    xyz: tuple = None

    def resolve(self, centers):
        if self.xyz is not None:
            return tuple(self.xyz)
        if not 1 <= self.mol <= len(centers):
            raise InputError(f"centre on molecule {self.mol}, but {len(centers)} were read")
        return tuple(centers[self.mol - 1])


@dataclass
class Elec:
    """One ELEC section (its MGparm and PBEparm keywords)."""
    name: str = ""
    method: str = ""
    dime: tuple = None
    nlev: int = None
    grid: tuple = None
    glen: tuple = None
    gcent: Center = None
    cglen: tuple = None
    fglen: tuple = None
    cgcent: Center = None
    fgcent: Center = None
    chgm: str = None
    etol: float = None
    pdime: tuple = None
    ofrac: float = None
    async_rank: int = None
    useaqua: bool = False
    mol: int = None
    pbe: str = None           # lpbe / npbe / lrpbe / nrpbe / smpbe
    bcfl: str = None
    ions: list = field(default_factory=list)   # (charge, conc, radius)
    pdie: float = None
    sdie: float = None
    srfm: str = None
    srad: float = None
    swin: float = None
    sdens: float = None
    temp: float = None
    calcenergy: str = "no"
    calcforce: str = "no"
    writes: list = field(default_factory=list)  # (what, format, stem)
    writemat: list = field(default_factory=list)
    usemaps: list = field(default_factory=list)
    unsupported: list = field(default_factory=list)


@dataclass
class Apolar:
    name: str = ""
    tokens: list = field(default_factory=list)


@dataclass
class Print:
    what: str                       # elecenergy / elecforce / apolenergy / apolforce
    ids: list = field(default_factory=list)   # section names or 1-based numbers (str)
    ops: list = field(default_factory=list)   # '+' / '-' between ids


@dataclass
class Input:
    molecules: list = field(default_factory=list)
    maps: list = field(default_factory=list)      # (kind, fmt, path) from READ
    parms: list = field(default_factory=list)
    elecs: list = field(default_factory=list)
    apolars: list = field(default_factory=list)
    prints: list = field(default_factory=list)
    base: str = "."

    def path(self, p):
        return p if os.path.isabs(p) else os.path.join(self.base, p)


# ------------------------------------------------------------------ parser

def parse(text, base="."):
    """Parse an APBS input; `base` is the directory relative paths are read
    from (APBS reads them from its working directory)."""
    t = _Tokens(tokenize(text))
    inp = Input(base=base)
    while True:
        tok = t.peek()
        if tok is None:
            raise InputError("input ended without QUIT")
        tok = t.next().lower()
        if tok == "quit":
            break
        if tok == "read":
            _parse_read(t, inp)
        elif tok == "elec":
            inp.elecs.append(_parse_elec(t))
        elif tok == "apolar":
            inp.apolars.append(_parse_apolar(t))
        elif tok == "print":
            inp.prints.append(_parse_print(t))
        else:
            raise InputError(f"unknown section keyword {tok!r}")
    return inp


def parse_file(path):
    with open(path) as fh:
        return parse(fh.read(), base=os.path.dirname(os.path.abspath(path)))


def _parse_read(t, inp):
    """NOsh_parseREAD (nosh.c:974-1023)."""
    while True:
        tok = t.next("READ").lower()
        if tok == "end":
            return
        if tok == "mol":
            fmt = t.next("mol format").lower()
            inp.molecules.append(Molecule(fmt, t.next("mol path")))
        elif tok == "parm":
            inp.parms.append((t.next("parm format").lower(), t.next("parm path")))
        elif tok in ("diel", "kappa", "charge", "pot"):
            fmt = t.next(f"{tok} format").lower()
            n = 3 if tok == "diel" else 1
            paths = [t.next(f"{tok} path") for _ in range(n)]
            inp.maps.append((tok, fmt, paths))
        elif tok == "mesh":
            inp.maps.append((tok, t.next("mesh format").lower(), [t.next("mesh path")]))
        else:
            raise InputError(f"unknown READ keyword {tok!r}")


_MG_METHODS = ("mg-manual", "mg-auto", "mg-para", "mg-dummy")
_OTHER_METHODS = ("fe-manual", "geoflow", "geoflow-manual", "geoflow-auto", "tabi", "bem-manual",
                  "pbam", "pbam-auto", "pbsam", "pbsam-auto", "pygbe")


def _center(t, what):
    tok = t.next(what)
    if tok.lower() == "mol":
        return Center(mol=t.int(what + " molecule"))
    try:
        x = float(tok)
    except ValueError:
        raise InputError(f"bad {what} {tok!r}") from None
    return Center(xyz=(x, t.float(what), t.float(what)))


def _parse_elec(t):
    """NOsh_parseELEC (nosh.c:1191-1322) with MGparm_parseToken
    (mgparm.c:916) and PBEparm_parseToken (pbeparm.c:1215)."""
    e = Elec()
    tok = t.next("ELEC").lower()
    if tok == "name":
        e.name = t.next("ELEC name")
        tok = t.next("ELEC method").lower()
    if tok in _MG_METHODS:
        e.method = tok
    elif tok in _OTHER_METHODS:
        e.method = tok
        e.unsupported.append(tok)
        while t.next("ELEC").lower() != "end":
            pass
        return e
    else:
        raise InputError(f"unknown ELEC method {tok!r}")
    while True:
        tok = t.next("ELEC").lower()
        if tok == "end":
            return e
        # MGparm keywords (MGparm_parseToken, mgparm.c:916)
        if tok == "dime":
            e.dime = t.ints(3, "dime")
        elif tok == "nlev":
            e.nlev = t.int("nlev")
        elif tok == "chgm":
            e.chgm = t.next("chgm").lower()
        elif tok == "etol":
            e.etol = t.float("etol")
        elif tok == "grid":
            e.grid = t.floats(3, "grid")
        elif tok == "glen":
            e.glen = t.floats(3, "glen")
        elif tok == "gcent":
            e.gcent = _center(t, "gcent")
        elif tok == "cglen":
            e.cglen = t.floats(3, "cglen")
        elif tok == "fglen":
            e.fglen = t.floats(3, "fglen")
        elif tok == "cgcent":
            e.cgcent = _center(t, "cgcent")
        elif tok == "fgcent":
            e.fgcent = _center(t, "fgcent")
        elif tok == "pdime":
            e.pdime = t.ints(3, "pdime")
        elif tok == "ofrac":
            e.ofrac = t.float("ofrac")
        elif tok == "async":
            e.async_rank = t.int("async")
        elif tok == "useaqua":
            e.useaqua = True
        # PBEparm keywords (PBEparm_parseToken, pbeparm.c:1215)
        elif tok == "mol":
            e.mol = t.int("mol")
        elif tok in ("lpbe", "npbe", "lrpbe", "nrpbe", "smpbe"):
            e.pbe = tok
            if tok == "smpbe":   # smpbe vol V size S (PBEparm_parseSMPBE, pbeparm.c:415)
                for _ in range(2):
                    t.next("smpbe key")
                    t.float("smpbe value")
        elif tok == "bcfl":
            e.bcfl = _bcfl(t.next("bcfl").lower())
        elif tok == "ion":
            e.ions.append(_parse_ion(t))
        elif tok == "pdie":
            e.pdie = t.float("pdie")
        elif tok == "sdie":
            e.sdie = t.float("sdie")
        elif tok == "sdens":
            e.sdens = t.float("sdens")
        elif tok == "srfm":
            e.srfm = t.next("srfm").lower()
        elif tok == "srad":
            e.srad = t.float("srad")
        elif tok == "swin":
            e.swin = t.float("swin")
        elif tok == "temp":
            e.temp = t.float("temp")
        elif tok == "calcenergy":
            e.calcenergy = _choice(t.next("calcenergy").lower(), ("no", "total", "comps"),
                                   "calcenergy")
        elif tok == "calcforce":
            e.calcforce = _choice(t.next("calcforce").lower(), ("no", "total", "comps"),
                                  "calcforce")
        elif tok == "write":
            what = t.next("write type").lower()
            fmt = t.next("write format").lower()
            e.writes.append((what, fmt, t.next("write stem")))
        elif tok == "writemat":
            e.writemat.append((t.next("writemat type").lower(), t.next("writemat stem")))
        elif tok == "usemap":
            e.usemaps.append((t.next("usemap type").lower(), t.int("usemap id")))
        elif tok in ("zmem", "lmem", "mdie", "memv"):   # bcfl mem parameters
            t.float(tok)
            e.unsupported.append(tok)
        else:
            raise InputError(f"unknown ELEC keyword {tok!r}")


# the deprecated integer form of bcfl (Vbcfl, src/generic/vhal.h)
_BCFL_INT = {0: "zero", 1: "sdh", 2: "mdh", 4: "focus", 5: "mem", 6: "map"}


def _bcfl(tok):
    """PBEparm_parseBCFL (pbeparm.c:474-550): a word, or a deprecated int."""
    try:
        k = int(tok)
    except ValueError:
        if tok not in ("zero", "sdh", "mdh", "focus", "mem", "map"):
            raise InputError(f"unknown bcfl {tok!r}") from None
        return tok
    if k not in _BCFL_INT:
        raise InputError(f"unknown bcfl {k}")
    return _BCFL_INT[k]


def _choice(v, options, what):
    if v not in options:
        raise InputError(f"{what} must be one of {', '.join(options)}, got {v!r}")
    return v


def _parse_ion(t):
    """PBEparm_parseION (pbeparm.c:555-640): `ion charge q conc c radius r`
    in any key order, or the deprecated positional `ion q c r`."""
    first = t.peek()
    try:
        q = float(first)
    except (TypeError, ValueError):
        vals = {}
        for _ in range(3):
            key = t.next("ion key").lower()
            if key not in ("charge", "conc", "radius"):
                raise InputError(f"illegal ION key {key!r}")
            vals[key] = t.float(f"ion {key}")
        if len(vals) != 3:
            raise InputError("ION needs charge, conc and radius")
        return (vals["charge"], vals["conc"], vals["radius"])
    t.next()
    return (q, t.float("ion conc"), t.float("ion radius"))


def _parse_apolar(t):
    a = Apolar()
    tok = t.next("APOLAR")
    if tok.lower() == "name":
        a.name = t.next("APOLAR name")
    else:
        a.tokens.append(tok)
        if tok.lower() == "end":
            return a
    while True:
        tok = t.next("APOLAR")
        if tok.lower() == "end":
            return a
        a.tokens.append(tok)


def _parse_print(t):
    """NOsh_parsePRINT (nosh.c:1025-1189): `print what id (+|- id)* end`."""
    what = t.next("PRINT").lower()
    if what == "energy":
        what = "elecenergy"
    elif what == "force":
        what = "elecforce"
    if what not in ("elecenergy", "elecforce", "apolenergy", "apolforce"):
        raise InputError(f"unknown PRINT keyword {what!r}")
    p = Print(what)
    p.ids.append(t.next("PRINT id"))
    while True:
        tok = t.next("PRINT")
        if tok.lower() == "end":
            return p
        if tok not in ("+", "-"):
            raise InputError(f"PRINT expects + or - between ids, got {tok!r}")
        p.ops.append(tok)
        p.ids.append(t.next("PRINT id"))


def resolve_id(ident, sections, kind):
    """Index (0-based) of a PRINT id among the ELEC (or APOLAR) sections: an
    all-digit id is a 1-based section number, anything else a section name
    (NOsh_parsePRINT, nosh.c:1098-1160)."""
    if ident.isdigit():
        k = int(ident)
        if not 1 <= k <= len(sections):
            raise InputError(f"PRINT refers to {kind} section {k}, but {len(sections)} exist")
        return k - 1
    for i, s in enumerate(sections):
        if s.name == ident:
            return i
    raise InputError(f"PRINT refers to unknown {kind} section {ident!r}")


def check_runnable(inp):
    """Reasons this input cannot run on the PMG path (empty when it can)."""
    why = []
    for m in inp.molecules:
        if m.fmt != "pqr":
            why.append(f"READ mol {m.fmt}: only PQR molecules are read")
    nmaps = {}
    for kind, fmt, _ in inp.maps:
        if kind not in MAP_KINDS:
            why.append(f"READ {kind}: finite-element meshes are not used by the PMG path")
        elif fmt not in ("dx", "gz"):
            why.append(f"READ {kind} {fmt}: maps are read as OpenDX (dx, gz)")
        nmaps[kind] = nmaps.get(kind, 0) + 1
    for i, e in enumerate(inp.elecs):
        tag = e.name or str(i + 1)
        if e.method not in ("mg-manual", "mg-auto", "mg-para", "mg-dummy"):
            why.append(f"ELEC {tag}: method {e.method} is not run by this build")
            continue
        if e.unsupported:
            why.append(f"ELEC {tag}: {', '.join(e.unsupported)} not supported")
        if e.pbe not in ("lpbe", "npbe"):
            why.append(f"ELEC {tag}: {e.pbe or 'no PBE type'} (lpbe and npbe are run)")
        if e.bcfl not in ("zero", "sdh", "mdh", "focus", "map"):
            why.append(f"ELEC {tag}: bcfl {e.bcfl} (zero, sdh, mdh, map and focus are run)")
        if e.chgm not in ("spl0", "spl2", "spl4"):
            why.append(f"ELEC {tag}: chgm {e.chgm} (spl0, spl2 and spl4 are run)")
        if e.srfm not in ("mol", "smol", "spl2", "spl3", "spl4"):
            why.append(f"ELEC {tag}: srfm {e.srfm} (mol, smol, spl2, spl3 and spl4 are run)")
        used = set()
        for kind, ident in e.usemaps:     # routines.c:1482-1561
            if kind not in MAP_KINDS:
                why.append(f"ELEC {tag}: usemap {kind} (diel, kappa, charge and pot are read)")
            elif not 1 <= ident <= nmaps.get(kind, 0):
                why.append(f"ELEC {tag}: {ident} is not a valid {kind} map ID")
            used.add(kind)
        if e.bcfl == "map" and "pot" not in used:
            why.append(f"ELEC {tag}: bcfl map needs a potential map (usemap pot)")
        salt = any(c > 0 and q != 0 for (q, c, _) in e.ions)
        if "kappa" in used and "diel" not in used:
            why.append(f"ELEC {tag}: a kappa map is used only with dielectric maps (usemap diel)")
        if "diel" in used and salt and "kappa" not in used:
            why.append(f"ELEC {tag}: dielectric maps with salt need a kappa map (usemap kappa)")
        for what, fmt, _ in e.writes:
            if fmt not in ("dx", "gz"):
                why.append(f"ELEC {tag}: write {what} {fmt} (dx and gz output are written)")
            elif what not in WRITABLE:
                why.append(f"ELEC {tag}: write {what} (writable: {', '.join(WRITABLE)})")
        missing = [k for k in ("dime", "mol", "pdie", "sdie", "srfm", "chgm", "temp", "bcfl")
                   if getattr(e, k) is None]
        if e.srfm in ("mol", "smol") and e.srad is None:
            missing.append("srad")
        if e.srfm in ("mol", "smol") and e.sdens is None and (e.srad or 0) > 1e-9:
            missing.append("sdens")
        if e.srfm in ("spl2", "spl3", "spl4") and e.swin is None:
            missing.append("swin")
        if e.method in ("mg-manual", "mg-dummy") and e.grid is None and e.glen is None:
            missing.append("grid or glen")
        if e.method in ("mg-manual", "mg-dummy") and e.gcent is None:
            missing.append("gcent")
        if e.method in ("mg-auto", "mg-para"):
            missing += [k for k in ("cglen", "fglen", "cgcent", "fgcent") if getattr(e, k) is None]
        if e.method == "mg-para":
            missing += [k for k in ("pdime", "ofrac") if getattr(e, k) is None]
        if missing:
            why.append(f"ELEC {tag}: not set: {', '.join(missing)}")
        if e.bcfl == "focus" and i == 0:
            why.append(f"ELEC {tag}: can't focus the first calculation")
    para = [e for e in inp.elecs if e.method == "mg-para" and e.pdime]
    nprocs = {e.pdime[0] * e.pdime[1] * e.pdime[2] for e in para}
    if len(nprocs) > 1:
        why.append("mg-para sections with different processor counts (pdime)")
    asyncs = {e.async_rank for e in para}
    if len(asyncs) > 1:
        why.append("mg-para sections with different async ranks")
    elif para and None not in asyncs and len(nprocs) == 1:
        rank = asyncs.pop()
        if not 0 <= rank < nprocs.pop():
            why.append(f"async {rank} is outside the pdime processor grid")
    if para and len(inp.elecs) != len(para):
        why.append("mg-para sections mixed with other ELEC methods")
    for p in inp.prints:
        if p.what.startswith("elec"):
            for ident in p.ids:
                try:
                    resolve_id(ident, inp.elecs, "ELEC")
                except InputError as err:
                    why.append(str(err))
    return why


# map kinds READ and usemap take (NOsh_parseREAD_*, PBEparm_parseUSEMAP)
MAP_KINDS = ("diel", "kappa", "charge", "pot")

# data arrays `write` can produce (Vpmg_copyArray names)
WRITABLE = ("pot", "charge", "dielx", "diely", "dielz", "kappa")
