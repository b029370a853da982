"""Host logic of the CTMRG graph recorder that runs without a GPU.

* Synchronous host copies (`_ops.upload`, `_ops.download`) refuse to run while
  a recorder captures (NotRecordable, raised before any device call) and run
  normally otherwise.
* Native plans dropped while a capture is open are destroyed only after the
  last capture has ended.
* A projector decision replays as "same" only when the re-evaluated cuts equal
  the recorded ones; it then writes the new s^-1/2 weights into the pinned
  buffer that the recorded upload reads, and refuses when the re-evaluation
  asks for work the recording does not hold.
"""

import numpy as np
import pytest
import torch

from frostspin_amd import _ops
from frostspin_amd.ctmrg.ctm_renormalize import _CutDecision, _inv_sqrt_kept


class _FakeRecorder:
    def __init__(self, capturing):
        self.capturing = capturing
        self.kept = []

    def keep(self, obj):
        self.kept.append(obj)


def test_sync_copies_refused_while_capturing(monkeypatch):
    monkeypatch.setattr(_ops.GraphHooks, "active", _FakeRecorder(capturing=True))
    with pytest.raises(_ops.NotRecordable):
        _ops.upload(np.arange(3.0))
    with pytest.raises(_ops.NotRecordable):
        _ops.download(torch.zeros(3, dtype=torch.float64))
    # a recorder between its segments (not capturing) does not refuse
    monkeypatch.setattr(_ops.GraphHooks, "active", _FakeRecorder(capturing=False))
    assert np.array_equal(_ops.download(torch.arange(3, dtype=torch.float64)), np.arange(3.0))


def test_plan_destruction_deferred_past_captures():
    destroyed = []
    _ops.capture_started()
    _ops.capture_started()  # two captures open (two threads)
    _ops._destroy(destroyed.append, 1)
    _ops._destroy(destroyed.append, 2)
    assert destroyed == []
    _ops.capture_ended()
    assert destroyed == []  # one capture still open
    _ops.capture_ended()
    assert destroyed == [1, 2]
    _ops._destroy(destroyed.append, 3)  # none open: at once
    assert destroyed == [1, 2, 3]


def _decision(recorded_cuts, evaluate):
    d = _CutDecision()
    d.cuts = np.asarray(recorded_cuts, dtype=np.int64)
    d.non_empty = d.cuts.nonzero()[0]
    d.evaluate = evaluate
    n = int(d.cuts[d.non_empty].sum())
    d.staging.append(torch.full((n,), -1.0, dtype=torch.float64))
    return d


def test_decision_replay_writes_weights_when_cuts_agree():
    s_of = [np.array([4.0, 1.0, 0.25]), np.array([9.0, 0.5]), np.array([2.0])]
    cuts = np.array([2, 0, 1])
    d = _decision(cuts, lambda replay: (cuts.copy(), s_of, None, None))
    assert d.replay()
    assert np.array_equal(d.staging[0].numpy(), _inv_sqrt_kept(s_of, cuts, cuts.nonzero()[0]))
    assert np.array_equal(d.staging[0].numpy(), [0.5, 1.0, 1.0 / np.sqrt(2.0)])


def test_decision_replay_refuses_other_cuts_or_redo():
    s_of = [np.array([4.0, 1.0]), np.array([9.0])]
    d = _decision([2, 1], lambda replay: (np.array([1, 1]), s_of, None, None))
    assert not d.replay()
    assert (d.staging[0].numpy() == -1.0).all()  # nothing written
    d = _decision([2, 1], lambda replay: None)  # the eager path would redo a decomposition
    assert not d.replay()
