"""Full-size parity of the 256-circuit AC sweep (uneven RLC ladders of
1k-64k sections, 101 log-spaced frequencies from 1 Hz to 1 GHz: 25,856
complex systems) against the oracle's FP64 sparse LU (ac::run,
src/analysis/ac.rs:56-70, with a direct linear solve): every circuit is
checked at three frequencies (rotating, so all 101 are covered) plus the
sweep's two ends, from the solutions of one device sweep."""
import concurrent.futures as cf
import os

import numpy as np
import pytest

import oracle
from ohmnivore_amd import circuit as ir
from ohmnivore_amd import solver

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def be():
    return solver.HipBackend()


def close(x, ref, rel=1e-6, abs_=1e-9):
    err = np.abs(x - ref)
    return bool((err <= np.maximum(rel * np.abs(ref), abs_)).all()), float(err.max(initial=0.0))


def test_rlc_sweep_256_every_circuit(be):
    ckts, freqs = ir.rlc_sweep(n_circuits=256, seed=4)
    assert freqs.shape[0] == 101
    mnas = [solver.compile(c) for c in ckts]
    del ckts
    plan = solver.AcPlan(be, mnas)
    r = plan.solve(freqs)
    worst = max(s.residual_norm for sts in r.stats for s in sts)
    assert all(s.status == 0 for sts in r.stats for s in sts)
    assert worst < 1e-12, worst

    picks = {c: sorted({c % 101, (c + 34) % 101, (c + 67) % 101} |
                       ({0, 100} if c % 16 == 0 else set())) for c in range(len(mnas))}
    gpu = {c: np.stack([plan.x(c, f, 1)[0] for f in fs]) for c, fs in picks.items()}

    def check(c):
        ref = oracle.ac_sweep(mnas[c], freqs[picks[c]])
        return c, close(gpu[c], ref)

    threads = min(16, os.cpu_count() or 1)
    bad = []
    with cf.ThreadPoolExecutor(threads) as ex:
        for c, (ok, err) in ex.map(check, range(len(mnas))):
            if not ok:
                bad.append((c, err))
    assert not bad, bad[:10]
    del plan
