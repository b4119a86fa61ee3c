"""The learned start's fallback rule (GuessPolicy in incrmt_amd/csrc/mg.hh),
exercised on the CPU through incrmt_guess_policy_run with faults injected
into the residual sequence: a start whose relative residual after the first
V-cycle exceeds the tolerance, or is not finite (a network that produced
NaN or infinity), is rejected and the solve falls back to the quadratic
start with FP64 V-cycles; eight rejections in a row switch the learned start
off for 32 solves, after which one probe decides whether it comes back."""
import numpy as np

import incrmt_amd as ia

TOL = 1e-4


def test_accepts_small_residuals():
    assert ia.guess_policy_run([1e-6, 5e-5, 1e-4], TOL).tolist() == [1, 1, 1]


def test_rejects_large_and_non_finite_residuals():
    d = ia.guess_policy_run([1e-3, np.nan, np.inf, -np.inf, 1e-5], TOL).tolist()
    assert d == [0, 0, 0, 0, 1]


def test_breaker_opens_after_eight_misses_and_probes_after_32():
    seq = [1e-5] * 2 + [1e-2] * 8 + [1e-5] * 40
    d = ia.guess_policy_run(seq, TOL)
    assert d[:2].tolist() == [1, 1]
    assert d[2:10].tolist() == [0] * 8                # eight fallbacks
    off = np.flatnonzero(d == -1)
    assert off[0] == 10 and len(off) == 31            # learned start skipped meanwhile
    assert d[41] == 1 and np.all(d[41:] == 1)         # the probe passes: back on


def test_failed_probe_switches_off_again():
    seq = [1e-2] * 8 + [1e-5] * 31 + [1e-2] + [1e-5] * 40
    d = ia.guess_policy_run(seq, TOL)
    assert d[39] == 0                                 # the probe fails
    assert np.all(d[40:71] == -1) and d[71] == 1      # another 31 skipped, then the next probe passes


def test_disabled_by_zero_tolerance():
    # the host forces every fallback with a zero tolerance (the GPU tests
    # inject the fault this way)
    d = ia.guess_policy_run([1e-12] * 10, 0.0)
    assert d[:8].tolist() == [0] * 8 and np.all(d[8:] == -1)
