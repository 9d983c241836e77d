"""OpenDX scalar grids: the `write <type> dx|gz` output of an ELEC section
and a reader for the same files.

Layout of APBS's Vgrid_writeDX (src/mg/vgrid.c:1206-1456, the non-partition
branch): four comment lines, `object 1 class gridpositions` with the origin
and three axis deltas, `object 2 class gridconnections`, `object 3 class
array ... data follows` with the values three to a line as "%12.6e " with the
z index fastest (loop order i, j, k over the x-fastest array), then the
attribute and field trailer.  `gz` is the same text gzip-compressed
(Vgrid_writeGZ, vgrid.c:1011).
"""
import gzip

import numpy as np

_FMT3 = "%12.6e %12.6e %12.6e"


def write_dx(path, data, origin, spacing, title="", gz=False, source="apbs-pmg-mi355x"):
    """data: array of shape (nz, ny, nx) (x fastest in memory, as Vpmg keeps
    its grids); origin (xmin, ymin, zmin); spacing (hx, hy, hz)."""
    a = np.asarray(data, dtype=np.float64)
    if a.ndim != 3:
        raise ValueError("data must have shape (nz, ny, nx)")
    nz, ny, nx = a.shape
    lines = [f"# Data from {source}", "# ", f"# {title}", "# ",
             f"object 1 class gridpositions counts {nx} {ny} {nz}",
             "origin " + _FMT3 % tuple(origin),
             "delta " + _FMT3 % (spacing[0], 0.0, 0.0),
             "delta " + _FMT3 % (0.0, spacing[1], 0.0),
             "delta " + _FMT3 % (0.0, 0.0, spacing[2]),
             f"object 2 class gridconnections counts {nx} {ny} {nz}",
             f"object 3 class array type double rank 0 items {nx * ny * nz} data follows"]
    # z fastest: the (nz, ny, nx) array transposed to (nx, ny, nz), flattened
    vals = np.ascontiguousarray(a.transpose(2, 1, 0)).ravel()
    body = []
    full = (vals.size // 3) * 3
    for r in range(0, full, 3):
        body.append("%12.6e %12.6e %12.6e " % (vals[r], vals[r + 1], vals[r + 2]))
    if full < vals.size:
        body.append("".join("%12.6e " % v for v in vals[full:]))
    tail = ['attribute "dep" string "positions"',
            'object "regular positions regular connections" class field',
            'component "positions" value 1', 'component "connections" value 2',
            'component "data" value 3']
    text = "\n".join(lines + body + tail) + "\n"
    if gz:
        with gzip.open(path, "wt") as fh:
            fh.write(text)
    else:
        with open(path, "w") as fh:
            fh.write(text)


def read_dx(path):
    """Read a scalar OpenDX grid (Vgrid_readDX, vgrid.c:586-808): returns
    (data (nz, ny, nx), origin, spacing)."""
    with open(path, "rb") as fh:       # gzip-compressed DX (VDF_GZ) or plain text
        zipped = fh.read(2) == b"\x1f\x8b"
    opener = gzip.open if zipped else open
    with opener(path, "rt") as fh:
        text = fh.read()
    counts = origin = None
    deltas = []
    items = None
    toks = []
    it = iter(text.splitlines())
    for line in it:
        s = line.strip()
        if not s or s.startswith("#"):
            continue
        w = s.split()
        if w[0] == "object" and "gridpositions" in w:
            counts = tuple(int(v) for v in w[-3:])
        elif w[0] == "origin":
            origin = tuple(float(v) for v in w[1:4])
        elif w[0] == "delta":
            deltas.append(tuple(float(v) for v in w[1:4]))
        elif w[0] == "object" and "array" in w:
            items = int(w[w.index("items") + 1])
            for body in it:
                b = body.strip()
                if not b:
                    continue
                if b[0].isalpha():
                    break
                toks.extend(b.split())
                if len(toks) >= items:
                    break
            break
    if counts is None or origin is None or len(deltas) != 3 or items is None:
        raise ValueError(f"{path}: not a scalar OpenDX grid")
    nx, ny, nz = counts
    if items != nx * ny * nz or len(toks) < items:
        raise ValueError(f"{path}: {len(toks)} values for {nx}x{ny}x{nz}")
    spacing = (deltas[0][0], deltas[1][1], deltas[2][2])
    vals = np.array(toks[:items], dtype=np.float64).reshape(nx, ny, nz)
    return np.ascontiguousarray(vals.transpose(2, 1, 0)), origin, spacing
