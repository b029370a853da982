"""Host logic of the CTMRG graph recorder that runs without a GPU.

* Synchronous host copies (`_ops.upload`, `_ops.download`) refuse to run while
  a recorder captures (NotRecordable, raised before any device call) and run
  normally otherwise.
* Native plans dropped while a capture is open are destroyed only after the
  last capture has ended.
* A projector decision replays as "same" only when the re-evaluated cuts equal
  the recorded ones, and refuses when the re-evaluation asks for work the
  recording does not hold; the s^-1/2 weights are gathered from the value
  buffers in sector order.
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
    return d


def test_decision_replay_compares_cuts():
    s_of = [np.array([4.0, 1.0]), np.array([9.0])]
    assert _decision([2, 1], lambda replay: (np.array([2, 1]), s_of, None, None, None)).replay()
    assert not _decision([2, 1], lambda replay: (np.array([1, 1]), s_of, None, None, None)).replay()
    # the eager path would redo a decomposition: the recorded launches do not hold it
    assert not _decision([2, 1], lambda replay: None).replay()


def test_inv_sqrt_weights_gathered_from_value_buffers(emulated_device):
    """s^-1/2 of the kept values, packed in sector order, gathered from two
    value buffers (randomized and dense sectors interleaved); within one ulp
    of the host formula (the CPU tensor library's vectorised square root is
    not always correctly rounded)."""
    a = torch.tensor([4.0, 1.0, 0.25, 16.0], dtype=torch.float64)
    b = torch.tensor([9.0, 0.5, 2.0], dtype=torch.float64)
    S_of = [(a, 0), (b, 0), (a, 3), (b, 2)]
    cuts = np.array([2, 1, 0, 1])
    w = _inv_sqrt_kept(S_of, cuts, cuts.nonzero()[0]).numpy()
    want = 1.0 / np.sqrt(np.array([4.0, 1.0, 9.0, 2.0]))
    np.testing.assert_allclose(w, want, rtol=2.0**-52, atol=0.0)
    assert _inv_sqrt_kept(S_of, np.zeros(4, np.int64), np.zeros(0, np.int64)).numel() == 0
