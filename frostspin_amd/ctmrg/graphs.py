"""Replay of settled CTMRG iterations as recorded HIP graphs.

Once the block structure of an environment stops changing (no permute or
structural plan built for `settle` consecutive iterations, and every move's
cuts equal to the previous iteration's), an iteration of SequentialCTMRG is a
fixed sequence of launches around four host waits (the singular values of each
move's projectors, sequential_ctmrg.py:38-45 of the reference). The recorder
then captures it, one move at a time, into eight graphs (torch.cuda.CUDAGraph
on a stream of its own, one memory pool): for move m, the launches before its
wait (A_m) and the launches after it (B_m). Every recorded graph runs right
after its capture, so the recording iteration computes what an eager one does.

Later iterations replay the graphs: after A_m the host re-evaluates the move's
decision (singular values read back into the same pinned buffers, redo
checks, multiplet-aware cuts); when the cuts equal the recorded ones it
replays B_m (which depends on the cuts only: the s^-1/2 weights are gathered
from the device copies of the singular values). Otherwise (the structure moves again, a decomposition would be
redone), the graphs are dropped, the environment is reset to the state
recorded at the start of move m, and the moves from m on run eagerly. The
same launches on the same data give the same results bit for bit.

State carried from one iteration to the next (the environment's C and T
tensors, the cached enlarged corners) is read by the first move from static
buffers: the recording iteration starts by copying that state into them, and
a closing graph, replayed after the last move's, copies the final state back
(the block layouts must then be identical, else no graph is kept). Every native plan a graph executes
is kept alive with the graphs (their device tables are read by address).
"""

import time
import warnings

import numpy as np
import torch

from .. import _ops
from ..config import config

# pinned host memory per recorded iteration for the read-back values and the
# uploaded weights (a SU(2) chi=256 iteration uses about 0.2 MB)
PINNED_BYTES = 16 << 20


def _key(r):
    if isinstance(r, np.ndarray):
        return (r.dtype.str, r.shape, r.tobytes())
    if isinstance(r, (tuple, list)):
        return tuple(_key(x) for x in r)
    return r


def _same_layout(a, b):
    """Same type, representations, blocks, offsets and memory order."""
    return (
        type(a) is type(b)
        and a._fortran == b._fortran
        and a._data.numel() == b._data.numel()
        and a._data.dtype == b._data.dtype
        and np.array_equal(a._bshapes, b._bshapes)
        and np.array_equal(a._boffs, b._boffs)
        and np.array_equal(a._block_irreps, b._block_irreps)
        and np.array_equal(a._signature, b._signature)
        and _key(a._row_reps + a._col_reps) == _key(b._row_reps + b._col_reps)
    )


class GraphRunner:
    """Runs the iterations of one SequentialCTMRG, recording and replaying
    settled ones (see the module docstring). `stats` counts recorded, replayed
    and dropped graph sets."""

    def __init__(self, ctm, settle=2):
        self.ctm = ctm
        self.settle = int(settle)
        self.stream = torch.cuda.Stream()
        self.moves = (ctm.up_move, ctm.right_move, ctm.down_move, ctm.left_move)
        self.capturing = False
        self.recording = False
        self.graphs = None  # per move: [A graph, B graph]
        self.closer = None  # copies of the final carried state into the static buffers
        self.decisions = None
        self.snapshots = None
        self.end_state = None
        self.kept = []
        self.static = None
        self.pool = None
        self._pin = None
        self._graph = None
        self._pin = None
        self._pin_used = 0
        self._builds = 0
        self.quiet = 0
        self.move = 0
        self.prev_cuts = [None] * len(self.moves)
        self.settings = None
        self.failed = None  # the error that stopped recording, if any
        self.stats = {"recorded": 0, "replayed": 0, "dropped": 0}

    # hooks called by _ops -------------------------------------------------
    def keep(self, obj):
        if self.recording:
            self.kept.append(obj)

    def pinned(self, n, dtype):
        """n elements of pinned host memory for a recorded copy, carved from a
        buffer allocated before the capture (the framework's host allocator
        must not run inside one) and held with the graphs."""
        nbytes = int(n) * torch.empty((), dtype=dtype).element_size()
        off = self._pin_used
        end = off + (nbytes + 255) // 256 * 256
        if end > self._pin.numel():
            msg = "graph recorder: pinned staging exhausted"
            raise RuntimeError(msg)
        self._pin_used = end
        return self._pin[off : off + nbytes].view(dtype)

    def end_segment(self, decision):
        if not self.capturing:
            _ops.wait_current()
            return
        self._end_capture()
        self.decisions[self.move] = decision
        self.stream.synchronize()

    def begin_segment(self, decision):
        m = self.move
        prev = self.prev_cuts[m]
        self.prev_cuts[m] = None if decision.cuts is None else decision.cuts.copy()
        if not self.recording:
            return
        if (decision.redone or prev is None or decision.cuts is None
                or not np.array_equal(prev, decision.cuts) or self.decisions[m] is not decision):
            self.recording = False  # the recorded graphs ran; the rest of the iteration runs eagerly
            return
        self._begin_capture()

    # iteration --------------------------------------------------------------
    def iterate(self):
        caller = torch.cuda.current_stream()
        self.stream.wait_stream(caller)
        prev_hooks = _ops.GraphHooks.active
        _ops.GraphHooks.active = self
        try:
            with torch.cuda.stream(self.stream):
                self._iterate()
        finally:
            _ops.GraphHooks.active = prev_hooks
            self.capturing = False
            caller.wait_stream(self.stream)

    def _settings(self):
        c = self.ctm
        return (id(c._env), c.chi_target, c.cutoff, c.degen_ratio, c.block_chi_ratio, c.ncv_ratio,
                tuple(id(a) for a in c._env._site["A"]))

    def _state(self):
        env = self.ctm._env
        return ({k: list(v) for k, v in env._site.items() if k != "A"},
                {k: list(v) for k, v in env._enlarged.items()})

    def _set_state(self, state):
        env = self.ctm._env
        site, enlarged = state
        for k, v in site.items():
            env._site[k] = list(v)
        env._enlarged = {k: list(v) for k, v in enlarged.items()}
        env._clear_move()

    def _state_is(self, state):
        """The environment holds exactly these tensor objects."""
        env = self.ctm._env
        site, enlarged = state
        for k, v in site.items():
            if len(env._site[k]) != len(v) or any(a is not b for a, b in zip(env._site[k], v, strict=True)):
                return False
        if set(env._enlarged) != set(enlarged):
            return False
        return all(len(env._enlarged[k]) == len(v) and all(a is b for a, b in zip(env._enlarged[k], v, strict=True))
                   for k, v in enlarged.items())

    def _iterate(self):
        if self.graphs is not None:
            if self._settings() != self.settings or not self._state_is(self.end_state):
                self._drop()
                self.quiet = 0
            else:
                m = self._replay()
                if m is None:
                    self.stats["replayed"] += 1
                    return
                self._set_state(self.snapshots[m])
                self._drop()
                self.quiet = 0
                self._eager(m)
                return
        builds = _ops.PLAN_BUILDS[0]
        if self.quiet >= self.settle and config["ctmrg_graphs"] and self.failed is None:
            self._record()
        else:
            self._eager(0)
        self.quiet = self.quiet + 1 if _ops.PLAN_BUILDS[0] == builds else 0

    def _eager(self, m0):
        for m in range(m0, len(self.moves)):
            self.move = m
            self.moves[m]()

    def _replay(self):
        """Replays the recorded iteration move by move: A_m, one wait, the
        move's decision re-evaluated, B_m; then the closing graph that
        refreshes the static buffers. Returns the index of the first move whose
        decision differs from the recorded one (the static buffers are then
        untouched), or None. Accumulates the host time spent outside the waits
        (graph launches, decisions) in stats. (Replaying all moves before one
        wait, possible since the launches after a wait depend on the cuts only,
        measured slower: p50 23.2 vs 22.8 ms, the launches of eight graphs in a
        row block the host.)"""
        t0 = time.perf_counter()
        waited = 0.0
        for m, (a, b) in enumerate(self.graphs):
            a.replay()
            tw = time.perf_counter()
            self.stream.synchronize()
            waited += time.perf_counter() - tw
            if not self.decisions[m].replay():
                return m
            b.replay()
        self.closer.replay()
        self.stats["replay_host_s"] = self.stats.get("replay_host_s", 0.0) + time.perf_counter() - t0 - waited
        return None

    # recording --------------------------------------------------------------
    def _begin_capture(self):
        self._graph = torch.cuda.CUDAGraph()
        self._builds = _ops.PLAN_BUILDS[0]
        # plans whose table copies completed need no event query inside
        self.stream.synchronize()
        _ops.settle_plan_tables()
        _ops.capture_started()
        try:
            self._graph.capture_begin(pool=self.pool, capture_error_mode="relaxed")
        except BaseException:
            _ops.capture_ended()
            raise
        self.capturing = True

    def _end_capture(self):
        g = self._graph
        self._graph = None
        self.capturing = False
        try:
            g.capture_end()
        finally:
            _ops.capture_ended()
        g.replay()  # the recorded launches run now, as they would have eagerly
        self.graphs[self.move].append(g)
        if _ops.PLAN_BUILDS[0] != self._builds:
            self.recording = False  # a plan built inside a recorded segment: keep nothing

    def _record(self):
        env = self.ctm._env
        # static copies of the carried state, read by the first move
        static = []
        for k, v in env._site.items():
            if k != "A":
                env._site[k] = [t._with(data=t._data.clone()) for t in v]
                static.extend(env._site[k])
        for k, v in env._enlarged.items():
            env._enlarged[k] = [None if t is None else t._with(data=t._data.clone()) for t in v]
            static.extend(t for t in env._enlarged[k] if t is not None)
        start = self._state()
        nm = len(self.moves)
        snapshots = []
        self.static = static
        self.graphs = [[] for _ in range(nm)]
        self.decisions = [None] * nm
        self.snapshots = snapshots
        self.kept = []
        self._pin = torch.empty((PINNED_BYTES,), dtype=torch.uint8, pin_memory=True)
        self._pin_used = 0
        self.pool = torch.cuda.graph_pool_handle()
        self.settings = self._settings()
        self.recording = True
        m = 0
        try:
            for m in range(nm):
                self.move = m
                snapshots.append(self._state())
                if self.recording:
                    self._begin_capture()
                self.moves[m]()
                if self.recording:
                    if len(self.graphs[m]) != 1:
                        # not exactly one host wait in this move: nothing to replay
                        self._end_capture()
                        self.recording = False
                        continue
                    self._end_capture()
            if self.recording:
                # the closing copies as a graph of their own: a speculative
                # replay runs it only once every decision has agreed
                self.move = nm
                self.graphs.append([])
                self._begin_capture()
                closed = self._close_loop(start)
                self._end_capture()
                self.closer = self.graphs.pop()[0]
                if not closed:
                    self.recording = False
        except BaseException as exc:
            if not self.capturing:
                self._drop()
                raise
            # an error inside a capture: the iteration goes on eagerly and this
            # runner records no more (a failing capture would otherwise be
            # retried, and paid for, every iteration)
            self.failed = repr(exc)
            warnings.warn(f"CTMRG graph recording disabled: {exc!r}", RuntimeWarning, stacklevel=2)
            self.capturing = False
            try:
                self._graph.capture_end()
            except Exception:  # noqa: BLE001 -- the capture is discarded either way
                pass
            finally:
                _ops.capture_ended()
            self._graph = None
            # the move's launches since its last wait did not run: run it again
            self._set_state(snapshots[m])
            self._drop()
            self._eager(m)
            return
        if not self.recording:
            self._drop()
            return
        self.recording = False
        self.end_state = self._state()
        self.stats["recorded"] += 1

    def _close_loop(self, start):
        """Inside the last move's capture: copy the final carried state into the
        static buffers the first move reads. False when a layout differs."""
        env = self.ctm._env
        site, enlarged = start
        pairs = []
        for k, v in site.items():
            pairs.extend(zip(v, env._site[k], strict=True))
        if set(env._enlarged) != set(enlarged):
            return False
        for k, v in enlarged.items():
            for s, f in zip(v, env._enlarged[k], strict=True):
                if (s is None) != (f is None):
                    return False
                if s is not None:
                    pairs.append((s, f))
        if not all(_same_layout(s, f) for s, f in pairs):
            return False
        for s, f in pairs:
            if s is not f:
                s._data.copy_(f._data)
        return True

    def _drop(self):
        if self.graphs is not None or self.recording:
            self.stats["dropped"] += 1
        self.recording = False
        self.graphs = None
        self.closer = None
        self.decisions = None
        self.snapshots = None
        self.end_state = None
        self.kept = []
        self.static = None
        self.pool = None
