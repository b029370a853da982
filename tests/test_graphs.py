"""Recorded HIP-graph iterations of SequentialCTMRG equal eager iterations bit
for bit (GPU).

Two CTMRGs start from the same simple-update tensors (SU(2) Heisenberg, D = 8,
beta = 4, chi = 256, the benchmark workload): one runs plain eager moves
(config["ctmrg_graphs"] off), the other records its iterations as graphs once
the block structure has settled and replays them afterwards. After every
iteration all C and T buffers of the two environments must be identical (same
launches on the same data), the recorded run must have recorded once and then
replayed. Every replay is checked, with eager work and host synchronisations
between the replays (the other run's iterations, the comparisons).

Then the recorded cuts of the third move are altered, so that the next replay
finds different cuts there: the graphs must be dropped, the environment reset
to the state recorded at that move's start and the remaining moves run
eagerly, still equal to the eager run; later iterations record again.
"""

import pytest
import torch

pytestmark = pytest.mark.gpu

_NAMES = ("C1", "C2", "C3", "C4", "T1", "T2", "T3", "T4")


def _buffers(ctm):
    e = ctm._env
    return [t._data for k in _NAMES for t in e._site[k]]


def _same(a, b):
    ba, bb = _buffers(a), _buffers(b)
    return len(ba) == len(bb) and all(x.shape == y.shape and torch.equal(x, y) for x, y in zip(ba, bb, strict=True))


def _iterate(ctm, graphs):
    from frostspin_amd.config import config

    old = config["ctmrg_graphs"]
    config["ctmrg_graphs"] = graphs
    try:
        ctm.iterate()
    finally:
        config["ctmrg_graphs"] = old


def test_graph_replay_equals_eager(gpu):
    from frostspin_amd.ctmrg import models
    from frostspin_amd.symmetric_tensor.tools import config

    checks = config["debug_checks"]
    config["debug_checks"] = False
    try:
        h = models.heisenberg_bond("SU2")
        _, ref = models.thermal_ctmrg(h, 8, 4.0, 256)
        _, rec = models.thermal_ctmrg(h, 8, 4.0, 256)
        assert _same(ref, rec)
        it = 0
        replayed_checked = 0
        while it < 120:
            _iterate(ref, False)
            _iterate(rec, True)
            it += 1
            assert _same(ref, rec), f"iteration {it}: recorded/replayed run differs from the eager run"
            runner = rec._graphs
            if runner.stats["replayed"] >= 4:
                replayed_checked = runner.stats["replayed"]
                break
        print(f"graph stats after {it} iterations: {rec._graphs.stats}")
        assert replayed_checked >= 4, f"no settled replays within {it} iterations: {rec._graphs.stats}"

        # a replay whose decision differs at move 2: dropped, eager from that move on
        runner = rec._graphs
        d = runner.decisions[2]
        cuts = d.cuts.copy()
        cuts[cuts.nonzero()[0][0]] += 1
        d.cuts = cuts
        dropped = runner.stats["dropped"]
        _iterate(ref, False)
        _iterate(rec, True)
        assert runner.stats["dropped"] == dropped + 1 and runner.graphs is None
        assert _same(ref, rec), "eager continuation after a dropped replay differs"
        for _ in range(runner.settle + 3):
            _iterate(ref, False)
            _iterate(rec, True)
            assert _same(ref, rec), f"after re-recording: {runner.stats}"
        assert runner.stats["recorded"] >= 2, runner.stats
        torch.cuda.synchronize()
    finally:
        config["debug_checks"] = checks
