"""Copy the reference's regression inputs into tests/golden/examples/ and
write expected.json from its tests/test_cases.cfg.

    python tests/golden/make_examples.py /path/to/apbs-reference

The copied files are data: the example input files (`*.in`), their molecules
(`*.pqr`) and force fixtures, unchanged, so the inputs run as written.
expected.json maps "<example>/<input stem>" to the list of values the
reference's tester reads from APBS output for that input (test_cases.cfg;
`*` entries, which it skips, become null; "forces" marks an input whose
printed forces are compared with the example's polarforces file).
"""
import json
import os
import re
import shutil
import sys

# example directory -> input stems run by tests/test_gpu_examples.py
EXAMPLES = {
    "born": ["apbs-mol-auto", "apbs-smol-auto", "apbs-forces", "apbs-mol-parallel",
             "apbs-smol-parallel"],
    "solv": ["apbs-mol", "apbs-smol"],
    "ion-protein": ["apbs-mol-pdiel2", "apbs-smol-pdiel2", "apbs-mol-pdiel12",
                    "apbs-smol-pdiel12"],
    "hca-bind": ["apbs-mol", "apbs-smol"],
    "pka-lig": ["apbs-mol-vdw", "apbs-smol-vdw", "apbs-mol-surf", "apbs-smol-surf"],
    "FKBP": ["1d7h-dmso-mol", "1d7h-dmso-smol", "1d7i-dss-mol", "1d7i-dss-smol"],
    "ionize": ["apbs-mol", "apbs-smol"],
    "point-pmf": ["complex-0_1", "complex-0_2", "complex-0_3", "complex-0_4"],
    "actin-dimer": ["apbs-mol-auto", "apbs-smol-auto", "apbs-mol-parallel", "apbs-smol-parallel"],
}
# test_cases.cfg section -> example directory
SECTIONS = {"born": "born", "solv": "solv", "ion-protein": "ion-protein", "hca-bind": "hca-bind",
            "pka-lig": "pka-lig", "FKBP": "FKBP", "ionize": "ionize", "point-pmf": "point-pmf",
            "actin-dimer-auto": "actin-dimer", "actin-dimer-parallel": "actin-dimer"}
# sections test_cases.cfg keeps commented out as slow; their values are still read
COMMENTED = {"actin-dimer-parallel"}
EXTRA = {"born": ["polarforces"]}


def referenced_files(inp_text):
    """Molecule files named by `mol pqr <file>` in a READ section."""
    return re.findall(r"^\s*mol\s+pqr\s+(\S+)", inp_text, flags=re.M | re.I)


def main(ref):
    here = os.path.dirname(os.path.abspath(__file__))
    out = os.path.join(here, "examples")
    expected = {}
    section = None
    with open(os.path.join(ref, "tests", "test_cases.cfg")) as fh:
        for line in fh:
            s = line.strip()
            m = re.match(r"#?\s*\[(.+)\]", s)
            if m:
                section = m.group(1)
                if s.startswith("#") and section not in COMMENTED:
                    section = None
                continue
            if s.startswith("#") and section in COMMENTED:
                s = s.lstrip("#").strip()
            if not s or s.startswith("#"):
                continue
            key, _, val = s.partition(":")
            key = key.strip()
            ex = SECTIONS.get(section)
            if ex is None or key not in EXAMPLES[ex]:
                continue
            if val.strip() == "forces":     # compared with the example's polarforces file
                expected[f"{ex}/{key}"] = "forces"
            else:
                expected[f"{ex}/{key}"] = [None if v == "*" else float(v) for v in val.split()]
    for ex, stems in EXAMPLES.items():
        src = os.path.join(ref, "examples", ex)
        dst = os.path.join(out, ex)
        os.makedirs(dst, exist_ok=True)
        files = set(EXTRA.get(ex, []))
        for st in stems:
            files.add(st + ".in")
            with open(os.path.join(src, st + ".in")) as fh:
                files.update(referenced_files(fh.read()))
        for f in sorted(files):
            shutil.copyfile(os.path.join(src, f), os.path.join(dst, f))
    with open(os.path.join(out, "expected.json"), "w") as fh:
        json.dump(expected, fh, indent=1, sort_keys=True)
    missing = [f"{ex}/{st}" for ex, sts in EXAMPLES.items() for st in sts
               if f"{ex}/{st}" not in expected]
    print(f"{len(expected)} expected entries; no expected values for: {missing}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
