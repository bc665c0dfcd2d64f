"""Writes io_mr19_mock.json: known answers for the catalog readers.

Sources, both in the reference's tree:
  * the doctest table of read_ascii_catalog (Corrfunc/io.py:161-183): the
    first 20 (ra, dec, cz) rows of the Mr19 north-only mock, printed with
    "{:10.5f}";
  * mocks/tests/data/Mr19_mock_northonly.rdcz.ff, the same catalog as a
    fast-food file: its header records and the first 20 values of each of its
    four columns (ra, dec, cz, weight), stored as exact hex doubles.
Run in a container that has the reference sources:
    python tests/golden/make_io_fixture.py /root/reference
"""
import json
import os
import re
import struct
import sys

import numpy as np

NROWS = 20


def doctest_rows(ref):
    text = open(os.path.join(ref, "Corrfunc/io.py")).read()
    doc = [p for p in text.split('"""') if "read_ascii_catalog(filename)" in p][0]
    rows = []
    for line in doc.splitlines():
        if re.match(r"^\s+-?\d+\.\d+\s+-?\d+\.\d+\s+-?\d+\.\d+\s*$", line):
            rows.append([float(v) for v in line.split()])
    return rows


def fastfood_head(path):
    with open(path, "rb") as f:
        b = f.read()
    off = 0

    def record():
        nonlocal off
        n = struct.unpack_from("<i", b, off)[0]
        payload = b[off + 4:off + 4 + n]
        n2 = struct.unpack_from("<i", b, off + 4 + n)[0]
        assert n2 == n
        off += 8 + n
        return payload

    idat = list(struct.unpack("<5i", record()))
    fdat = [float(v) for v in np.frombuffer(record(), "<f4")]
    znow = float(np.frombuffer(record(), "<f4")[0])
    ngal = idat[1]
    cols = []
    while off < len(b):
        payload = record()
        dt = "<f4" if len(payload) // ngal == 4 else "<f8"
        cols.append(np.frombuffer(payload, dt))
    return dict(idat=idat, fdat=fdat, znow=znow, ngal=ngal, ncols=len(cols),
                itemsize=int(cols[0].dtype.itemsize),
                head=[[float(v).hex() for v in c[:NROWS]] for c in cols])


def main(ref):
    out = dict(doctest_rows=doctest_rows(ref),
               file="mocks/tests/data/Mr19_mock_northonly.rdcz.ff",
               **fastfood_head(os.path.join(ref, "mocks/tests/data/Mr19_mock_northonly.rdcz.ff")))
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "io_mr19_mock.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(len(out["doctest_rows"]), "doctest rows;", out["ngal"], "galaxies,", out["ncols"], "columns")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
