"""Writes mocks_tables.json: known answers for the on-sky (mocks) pair counters.

Sources, all in the reference's tree:
  * mocks/tests/Mr19_mock.DD -- DD(rp, pi) of the Mr19 north-only mock
    (mocks/tests/data/Mr19_mock_northonly.rdcz.*: RA, DEC, cz, weight) with
    cosmology 1, pimax 40, the bins of mocks/tests/bins, pair_product weights
    and average rp (mocks/tests/tests_mocks.c:46-131, 455-470). Columns: npairs,
    rpavg, log10(rp_max), pi_max, weightavg.
  * mocks/tests/bins -- those bins.
  * the doctest table of Corrfunc/mocks/DDrppi_mocks.py:241-317 (uniform points
    in a 420 box seen from its centre, is_comoving_dist=True).
Run in a container that has the reference sources:
    python tests/golden/make_mocks_fixture.py /root/reference
"""
import json
import os
import re
import sys


def main(ref):
    dd = [[float(v) for v in line.split()] for line in open(os.path.join(ref, "mocks/tests/Mr19_mock.DD"))
          if line.strip()]
    bins = [[float(v) for v in line.split()] for line in open(os.path.join(ref, "mocks/tests/bins")) if line.strip()]
    text = open(os.path.join(ref, "Corrfunc/mocks/DDrppi_mocks.py")).read()
    doc = [p for p in text.split('"""') if ">>> results = DDrppi_mocks(" in p][0]
    rows = []
    for line in doc.splitlines():
        if re.match(r"^\s+\d+\.\d+\s+\d+\.\d+\s+\d+\.\d+\s+\d+\.\d+\s+\d+\s+\d+\.\d+\s*$", line):
            v = line.split()
            rows.append([float(v[0]), float(v[1]), float(v[2]), float(v[3]), int(v[4]), float(v[5])])
    out = dict(mr19_dd=dict(rows=dd, cols=["npairs", "rpavg", "log10_rpmax", "pimax", "weightavg"],
                            input="mocks/tests/data/Mr19_mock_northonly.rdcz.ff", cosmology=1, pimax=40.0,
                            autocorr=1, weights="pair_product (4th column)", need_avg_sep=True),
               bins=bins,
               ddrppi_doctest=dict(rows=rows, first_row=519, N=100000, boxsize=420.0, seed=42, pimax=40.0,
                                   cosmology=1, is_comoving_dist=True, weights="ones",
                                   cols=["rmin", "rmax", "rpavg", "pimax", "npairs", "weightavg"]))
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "mocks_tables.json"), "w") as f:
        json.dump(out, f, indent=0)
    print(len(dd), "Mr19 DD rows;", len(rows), "doctest rows;", len(bins), "bins")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
