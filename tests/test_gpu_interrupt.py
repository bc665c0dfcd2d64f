"""Interrupt semantics of the counting entry points, as the reference's
(utils/macros.h:145-167; theory/DD/countpairs_impl.c.src:31-37, 191-193 and
554-566): SIGTERM, SIGINT or SIGHUP arriving during a call is caught by the
library, the call returns EXIT_FAILURE, and the caller's own handlers are in
place again afterwards. The signal is raised from the host inside the call
(CORRFUNC_AMD_TEST_SIGNAL), so every kernel still runs to its normal end."""
import os
import signal

import numpy as np
import pytest

from corrfunc_amd import _abi, _lib
from helpers import assert_same, run, uniform

pytestmark = pytest.mark.gpu

L = 200.0
BINS = np.logspace(-1, np.log10(15.0), 11)


@pytest.mark.parametrize("signo", [signal.SIGTERM, signal.SIGINT, signal.SIGHUP], ids=["TERM", "INT", "HUP"])
@pytest.mark.parametrize("mode", ["countpairs", "countpairs_xi"])
def test_signal_aborts_the_call_and_restores_handlers(engine, oracle, monkeypatch, signo, mode):
    _lib.set_engine(_lib.ENGINE_TUNED)
    X = uniform(20000, L, 31)
    kw = dict(nthreads=2, binfile=BINS, X1=X[0], Y1=X[1], Z1=X[2])
    kw.update(boxsize=L) if mode == "countpairs_xi" else kw.update(autocorr=1)
    o = dict(periodic=True, boxsize=L)
    hits = []
    prev = signal.signal(signo, lambda s, f: hits.append(s))
    try:
        monkeypatch.setenv("CORRFUNC_AMD_TEST_SIGNAL", str(int(signo)))
        assert run(engine, mode, options=_abi.make_options(**o), **kw) is None
        # the library, not the caller's handler, took the signal
        assert hits == []
        monkeypatch.delenv("CORRFUNC_AMD_TEST_SIGNAL")
        # the next call starts uninterrupted and counts correctly
        got = run(engine, mode, options=_abi.make_options(**o), **kw)
        want = run(oracle, mode, "oracle_", options=_abi.make_options(**o), **kw)
        assert_same(mode, got, want)
        # and the caller's handler is back
        os.kill(os.getpid(), signo)
        assert hits == [signo]
    finally:
        signal.signal(signo, prev)


def test_signal_aborts_a_batched_call(engine, monkeypatch):
    _lib.set_engine(_lib.ENGINE_TUNED)
    cats = [uniform(8000, L, s) for s in (41, 42)]
    o = _abi.make_options(periodic=True, boxsize=L)
    monkeypatch.setenv("CORRFUNC_AMD_TEST_SIGNAL", str(int(signal.SIGTERM)))
    prev = signal.signal(signal.SIGTERM, signal.SIG_IGN)
    try:
        with pytest.raises(RuntimeError):
            _lib.countpairs_batch(cats, BINS, options=o)
    finally:
        signal.signal(signal.SIGTERM, prev)
    monkeypatch.delenv("CORRFUNC_AMD_TEST_SIGNAL")
    outs = _lib.countpairs_batch(cats, BINS, options=o)
    assert len(outs) == 2 and all(r["npairs"].sum() > 0 for r in outs)
