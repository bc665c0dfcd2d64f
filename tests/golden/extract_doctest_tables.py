"""Writes doctest_tables.json: the known-answer tables printed by the doctests of
the reference's Python wrappers (Corrfunc/theory/DD.py:166-202,
DDrppi.py:183-249, DDsmu.py:199-265, xi.py:148-182, wp.py:437-472).

Only the numeric rows are kept, together with the call that produced them
(np.random.seed(42); X, Y, Z = uniform(0, 420, N); bins = theory/tests/bins,
copied here as `bins`). Run in a container that has the reference sources:
    python tests/golden/extract_doctest_tables.py /root/reference
"""
import json
import os
import re
import sys

SPECS = {
    "DD": dict(file="Corrfunc/theory/DD.py", N=10000, call="DD", first_row=0,
               cols=["rmin", "rmax", "ravg", "npairs", "weightavg"], output_ravg=True),
    "DDrppi": dict(file="Corrfunc/theory/DDrppi.py", N=10000, call="DDrppi", first_row=519, pimax=40.0,
                   cols=["rmin", "rmax", "rpavg", "pimax", "npairs", "weightavg"], output_ravg=True),
    "DDsmu": dict(file="Corrfunc/theory/DDsmu.py", N=10000, call="DDsmu", first_row=100, mu_max=1.0, nmu_bins=10,
                  cols=["smin", "smax", "savg", "mu_max", "npairs", "weightavg"], output_ravg=True),
    "xi": dict(file="Corrfunc/theory/xi.py", N=100000, call="xi", first_row=0,
               cols=["rmin", "rmax", "ravg", "xi", "npairs", "weightavg"], output_ravg=True),
    "wp": dict(file="Corrfunc/theory/wp.py", N=10000, call="wp", first_row=0, pimax=40.0,
               cols=["rmin", "rmax", "rpavg", "wp", "npairs", "weightavg"], output_ravg=False),
}

ROW = re.compile(r"^\s+(-?\d+\.\d+\s+){2,}.*$")


def main(ref):
    out = {}
    for name, spec in SPECS.items():
        text = open(os.path.join(ref, spec["file"])).read()
        doc = [part for part in text.split('"""')
               if ">>> results = %s(" % spec["call"] in part.replace("\n", " ")][0]
        rows = []
        for line in doc.splitlines():
            if ROW.match(line) and ">>>" not in line and "..." not in line:
                vals = line.split()
                if len(vals) == len(spec["cols"]):
                    rows.append([float(v) if "." in v else int(v) for v in vals])
        d = {k: v for k, v in spec.items() if k != "file"}
        d.update(boxsize=420.0, seed=42, weights="ones", rows=rows)
        out[name] = d
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "doctest_tables.json"), "w") as f:
        json.dump(out, f, indent=1)
    for k, v in out.items():
        print(k, len(v["rows"]), "rows")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
