"""The tuned kernel sizes its LDS histogram copies by the device's LDS limit
per workgroup (160 KiB on gfx950; above 64 KiB the launch raises the kernel's
dynamic-LDS attribute). The limit is read once per process, so the 64 KiB
shape runs in a child process (CORRFUNC_AMD_LDS_MAX=65536); both shapes must
give the oracle's counts, and the default must use at least as many copies."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from corrfunc_amd import _abi, _lib
from helpers import assert_same, clustered, mode_kwargs, run, weights

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
L = 160.0

CHILD = r"""
import json, sys
import numpy as np
sys.path.insert(0, sys.argv[1]); sys.path.insert(0, sys.argv[2])
from corrfunc_amd import _abi, _lib
from helpers import clustered, mode_kwargs, run, weights
from test_gpu_lds_limit import case
mode, out = sys.argv[3], sys.argv[4]
kw, o = case(mode)
_lib.set_engine(_lib.ENGINE_TUNED)
r = run(_lib.lib(), mode, options=_abi.make_options(**o), **kw)
np.savez(out, **{k: np.asarray(v) for k, v in r.items()})
print(json.dumps({"hist_copies": _lib.last_stats()["hist_copies"]}))
"""


def case(mode):
    """615 (rp, pi) slots or 31 x 41 (s, mu) slots: more copies fit in 160 KiB than in 64."""
    X = clustered(30000, L, 301, nparent=30, sigma=2.0, dtype=np.float64)
    W = weights(30000, 302, np.float64)
    if mode == "countpairs_rp_pi":
        bins = np.logspace(np.log10(0.1), np.log10(30.0), 15)
    else:
        bins = np.linspace(0.0, 30.0, 31)
    kw = mode_kwargs(mode, X, W, autocorr=1, bins=bins, L=L, pimax=40.0, mu_max=1.0, nmu_bins=40)
    return kw, dict(periodic=True, boxsize=L, need_avg_sep=True, isa="avx512f")


@pytest.mark.parametrize("mode", ["countpairs_rp_pi", "countpairs_s_mu"])
def test_lds_limit_shapes_keep_counts(engine, oracle, mode, tmp_path):
    kw, o = case(mode)
    os.environ["CORRFUNC_AMD_DENSE"] = "0"
    _lib.set_engine(_lib.ENGINE_TUNED)
    got = run(engine, mode, options=_abi.make_options(**o), **kw)
    r_default = _lib.last_stats()["hist_copies"]
    want = run(oracle, mode, "oracle_", options=_abi.make_options(**o), **kw)
    assert_same(mode, got, want, what="device LDS limit")
    assert want["npairs"].sum() > 0

    out = str(tmp_path / "child.npz")
    env = dict(os.environ, CORRFUNC_AMD_LDS_MAX="65536")
    p = subprocess.run([sys.executable, "-c", CHILD, ROOT, os.path.join(ROOT, "tests"), mode, out], env=env,
                       capture_output=True, text=True, timeout=180)
    assert p.returncode == 0, p.stderr[-2000:]
    r_64k = json.loads(p.stdout.strip().splitlines()[-1])["hist_copies"]
    child = dict(np.load(out))
    assert_same(mode, child, want, what="64 KiB LDS limit")
    assert r_64k >= 1 and r_default >= r_64k, (r_default, r_64k)
