"""Device operations: the only module that launches work through the C-ABI.

Every function takes torch tensors that live on the HIP device (float64 unless
stated) and enqueues the native kernel on torch's current stream. Tensors on
any other device are rejected: there is no host fallback.
"""

import ctypes
import threading
from collections import OrderedDict

import numpy as np
import torch

from . import _lib
from ._lib import check, lib

DTYPE = torch.float64


_HIP_AVAILABLE = None
_DEVICES = {}


def device():
    """Device that holds all block buffers (the current HIP device)."""
    global _HIP_AVAILABLE
    if _HIP_AVAILABLE is None:
        _HIP_AVAILABLE = torch.cuda.is_available()
    if not _HIP_AVAILABLE:
        msg = "frostspin_amd device operations need a HIP GPU (torch.cuda is unavailable)"
        raise RuntimeError(msg)
    idx = torch._C._cuda_getDevice()
    dev = _DEVICES.get(idx)
    if dev is None:
        dev = _DEVICES[idx] = torch.device("cuda", idx)
    return dev


_SIDE = threading.local()


def side_stream(k=0):
    """Extra stream number k of the calling thread on the current device
    (default priority, created once), for work independent of the current
    stream; None off-GPU."""
    if device().type != "cuda":
        return None
    idx = torch._C._cuda_getDevice()
    streams = getattr(_SIDE, "streams", None)
    if streams is None:
        streams = _SIDE.streams = {}
    st = streams.get((idx, k))
    if st is None:
        st = streams[(idx, k)] = torch.cuda.Stream(device=idx)
    return st


# Independent chains of small launches (the enlarged corners and the
# renormalizations of different sites) can run on separate streams; results
# are identical either way (same kernels on the same inputs).
# Measured on the SU(2) chi=256 workload with every launch from the host:
# 2-3 % slower than one stream (the chains are already issued faster than they
# run), so off by default there; while a CTMRG iteration is recorded as graphs
# the chains become parallel branches of the graph, at no host cost on replay.
CONCURRENCY = {"enabled": False, "recorded": True}


def _device_buffers(obj):
    if isinstance(obj, torch.Tensor):
        yield obj
    elif isinstance(obj, (tuple, list)):
        for o in obj:
            yield from _device_buffers(o)
    else:
        data = getattr(obj, "_data", None)
        if isinstance(data, torch.Tensor):
            yield data


def run_concurrently(thunks):
    """Call each thunk (a function enqueueing device work and returning its
    results), the first on the current stream and the others on side streams,
    then make the current stream wait for all of them. Buffers produced on a
    side stream are recorded on the current stream, so the allocator does not
    hand them to other work while the current stream may still read them.
    Inputs must be ready on the current stream when this is called (the side
    streams wait for it first)."""
    thunks = list(thunks)
    concurrent = CONCURRENCY["enabled"] or (CONCURRENCY["recorded"] and recording())
    if len(thunks) < 2 or not concurrent or KernelTimer.active is not None or device().type != "cuda":
        return [t() for t in thunks]
    main = torch.cuda.current_stream()
    sides = [side_stream(1 + i) for i in range(len(thunks) - 1)]
    for st in sides:
        st.wait_stream(main)
    try:
        results = [thunks[0]()]
        for st, t in zip(sides, thunks[1:], strict=True):
            with torch.cuda.stream(st):
                results.append(t())
    finally:
        # joined back whatever happened (a capture must not end with a side
        # stream's work left unjoined)
        for st in sides:
            main.wait_stream(st)
    for r in results[1:]:
        for buf in _device_buffers(r):
            buf.record_stream(main)
    return results


def _stream():
    """The calling thread's current stream on the current device (raw handle)."""
    return ctypes.c_void_p(torch._C._cuda_getCurrentRawStream(torch._C._cuda_getDevice()))


class KernelTimer:
    """Optional per-launch timing: HIP events recorded on the launch stream around
    every native call, with the algorithmic work of that call (flop for GEMMs,
    bytes moved for permutes). Used by the benchmark to price kernels against the
    roofline; off by default (no events, no synchronization)."""

    active = None

    def __init__(self, *, pad_cycles=0):
        self.records = []  # (kind, work, start_event, end_event)
        # pad_cycles > 0: a short spin kernel ahead of each timed launch keeps the
        # GPU busy while the host enqueues it, so the event pair brackets the
        # kernel itself rather than the host's launch latency
        self.pad_cycles = int(pad_cycles)

    def __enter__(self):
        KernelTimer.active = self
        return self

    def __exit__(self, *exc):
        KernelTimer.active = None

    def summary(self):
        torch.cuda.synchronize()
        out = {}
        for kind, work, e0, e1, disp in self.records:
            ms = e0.elapsed_time(e1)
            d = out.setdefault(kind, {"launches": 0, "dispatches": 0, "ms": 0.0, "work": 0.0})
            d["launches"] += 1
            d["dispatches"] += disp
            d["ms"] += ms
            d["work"] += work
        return out


class _Untimed:
    __slots__ = ()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        return False


_UNTIMED = _Untimed()


def _timed(kind, work, dispatches=1):
    """Event pair around one native call while a KernelTimer is active; a shared
    no-op otherwise. work may be a callable, evaluated only under a timer."""
    if KernelTimer.active is None:
        return _UNTIMED
    return _Timed(kind, work() if callable(work) else work, dispatches)


class _Timed:
    __slots__ = ("dispatches", "e0", "kind", "work")

    def __init__(self, kind, work, dispatches=1):
        self.kind = kind
        self.work = work
        self.dispatches = dispatches
        self.e0 = None

    def __enter__(self):
        kt = KernelTimer.active
        if kt is not None:
            if kt.pad_cycles:
                torch.cuda._sleep(kt.pad_cycles)
            self.e0 = torch.cuda.Event(enable_timing=True)
            self.e0.record()
        return self

    def __exit__(self, *exc):
        if self.e0 is not None and KernelTimer.active is not None:
            e1 = torch.cuda.Event(enable_timing=True)
            e1.record()
            disp = self.dispatches() if callable(self.dispatches) else self.dispatches
            KernelTimer.active.records.append((self.kind, self.work, self.e0, e1, disp))


def _ptr(t):
    if t is None:
        return None
    if not t.is_cuda:
        msg = "frostspin_amd kernels only accept device tensors"
        raise RuntimeError(msg)
    return t.data_ptr()  # an address for the library's void-pointer parameters


def empty(n, dtype=DTYPE):
    return torch.empty((int(n),), dtype=dtype, device=device())


def zeros(n, dtype=DTYPE):
    return torch.zeros((int(n),), dtype=dtype, device=device())


class NotRecordable(RuntimeError):
    """A host round trip (synchronous copy) inside a graph recording: the
    recorder stops recording and the iteration continues eagerly."""


def _refuse_recording(what):
    if recording():
        msg = f"{what} inside a recorded CTMRG iteration"
        raise NotRecordable(msg)


def upload(arr, dtype=DTYPE):
    """Host array -> new 1-D device tensor (flattened in C order)."""
    _refuse_recording("synchronous host-to-device copy")
    a = np.ascontiguousarray(arr).reshape(-1)
    return torch.from_numpy(a).to(device=device(), dtype=dtype)


def download(t):
    """Device tensor -> host numpy array (synchronizes the stream)."""
    _refuse_recording("synchronous device-to-host copy")
    return t.detach().to("cpu").numpy()


def download_start(t):
    """Start a device -> host copy on the current stream without waiting: the
    returned host tensor (pinned) holds the values once the stream has been
    synchronized (wait_current / stream.synchronize). Several readbacks then
    share one host wait. Host tensors are returned as a copy."""
    t = t.detach()
    if not t.is_cuda:
        return t.clone()
    if recording():
        # a graph node writing a pinned buffer of the recorder's on every replay
        t = t.contiguous()
        h = GraphHooks.active.pinned(t.numel(), t.dtype).view(t.shape)
        memcpy_async(h, t, t.numel() * t.element_size())
        return h
    h = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
    h.copy_(t, non_blocking=True)
    return h


def wait_current():
    """Host waits for the work queued so far on the current stream."""
    if device().type == "cuda":
        torch.cuda.current_stream().synchronize()


# ---------------------------------------------------------------------------
# hooks of the CTMRG graph recorder (frostspin_amd.ctmrg.graphs)
# ---------------------------------------------------------------------------
class _GraphHooks(threading.local):
    """`active` is the graph recorder of the CTMRG iteration running in this
    thread, if any: it is told where the host waits for read-back values and
    where it resumes launching, and keeps the native plans a recorded segment
    executes (their device tables must outlive the graph)."""

    active = None


GraphHooks = _GraphHooks()


PLAN_BUILDS = [0]  # native permute / structural plans built in this process


def host_wait(decision=None):
    """The host needs the values read back so far (singular values, status
    words): wait for the current stream. Under a recorder, the segment
    recorded since the previous wait ends here and runs; `decision` is the
    host step the recorder re-evaluates when it replays that segment."""
    rec = GraphHooks.active
    if rec is None:
        wait_current()
    else:
        rec.end_segment(decision)


def host_resume(decision=None):
    """The host decision after a host_wait is taken; launches resume."""
    rec = GraphHooks.active
    if rec is not None:
        rec.begin_segment(decision)


def recording():
    """True while a recorder captures launches into a graph."""
    rec = GraphHooks.active
    return rec is not None and rec.capturing


# Native plans dropped while some thread captures a graph are destroyed after
# the capture: a release records events on the streams that executed the plan,
# which must not happen on a stream being captured.
_CAPTURES = [0]
_DEFERRED = []
_DEFERRED_LOCK = threading.RLock()  # a collection inside it may run another __del__


def _destroy(fn, handle):
    with _DEFERRED_LOCK:
        if _CAPTURES[0]:
            _DEFERRED.append((fn, handle))
            return
    fn(handle)


def capture_started():
    with _DEFERRED_LOCK:
        _CAPTURES[0] += 1


def capture_ended():
    """Destroys the plans dropped during the captures, once none is left."""
    with _DEFERRED_LOCK:
        _CAPTURES[0] -= 1
        if _CAPTURES[0]:
            return
        pending = list(_DEFERRED)
        _DEFERRED.clear()
    for fn, h in pending:
        fn(h)


def settle_plan_tables():
    """Before a capture (after a device synchronisation): every completed plan
    table copy is marked complete, so plans used inside the capture need no
    event query (fs_plan_tables_settle)."""
    check(lib.fs_plan_tables_settle())


def memcpy_async(dst, src, nbytes):
    """Stream-ordered copy between a device tensor and a pinned host tensor
    (fs_memcpy_async): a node of the graph when recorded."""
    check(lib.fs_memcpy_async(dst.data_ptr(), src.data_ptr(), int(nbytes), _stream()))


def upload_async(arr, dtype=DTYPE, *, structural=False):
    """Host array -> new 1-D device tensor without a host wait: staged through
    pinned memory, copied on the current stream (ordered before the kernels
    launched after it). Recorded into a graph, the copy reads a pinned buffer
    of the recorder's and every replay copies the same values, so inside a
    recorded iteration only data fixed by the recorded structure (indices)
    may be uploaded: structural=True says so, anything else is refused."""
    a = torch.from_numpy(np.ascontiguousarray(arr).reshape(-1))
    dev = device()
    if dev.type != "cuda":
        return a.to(dtype=dtype, copy=True)
    if recording():
        if not structural:
            msg = "upload of values inside a recorded CTMRG iteration"
            raise NotRecordable(msg)
        h = GraphHooks.active.pinned(a.numel(), dtype)
        h.copy_(a)
        d = torch.empty(a.shape, dtype=dtype, device=dev)
        memcpy_async(d, h, h.numel() * h.element_size())
        return d
    h = torch.empty(a.shape, dtype=dtype, pin_memory=True)
    h.copy_(a)
    return h.to(dev, non_blocking=True)


# ---------------------------------------------------------------------------
# permute plans
# ---------------------------------------------------------------------------
class PermutePlan:
    """Owner of a native permute plan (host box tiles + lazily uploaded device copy)."""

    __slots__ = ("_handle", "irreps", "new_ncoeff", "old_ncoeff", "shapes")

    def __init__(self, handle):
        self._handle = handle
        nb = ctypes.c_int64()
        nc = ctypes.c_int64()
        oc = ctypes.c_int64()
        check(lib.fs_permute_plan_info(handle, ctypes.byref(nb), ctypes.byref(nc), ctypes.byref(oc)))
        self.irreps = np.empty((nb.value,), dtype=np.int8)
        shapes = np.empty((nb.value, 2), dtype=np.int64)
        check(lib.fs_permute_plan_blocks(handle, _lib.i8p(self.irreps), _lib.i64p(shapes)))
        self.shapes = shapes
        self.new_ncoeff = nc.value
        self.old_ncoeff = oc.value

    def tiles(self):
        """Host copy of the plan's box tiles (inspection and tests): an
        (ntiles, FS_PERMUTE_TILE_RECORD) int64 array laid out as documented for
        fs_permute_plan_tiles, and the number of boxes."""
        nt = ctypes.c_int64()
        nb = ctypes.c_int64()
        check(lib.fs_permute_plan_tiles(self._handle, ctypes.byref(nt), ctypes.byref(nb), None))
        out = np.zeros((nt.value, _lib.FS_PERMUTE_TILE_RECORD), dtype=np.int64)
        check(lib.fs_permute_plan_tiles(self._handle, ctypes.byref(nt), ctypes.byref(nb), _lib.i64p(out)))
        return out, nb.value

    def execute(self, src, dst):
        if GraphHooks.active is not None:
            GraphHooks.active.keep(self)
        with _timed("permute", 2.0 * dst.element_size() * self.new_ncoeff):
            if dst.dtype == torch.float64:
                check(lib.fs_permute_plan_execute_f64(self._handle, _ptr(src), _ptr(dst), _stream()))
            elif dst.dtype == torch.float32:
                check(lib.fs_permute_plan_execute_f32(self._handle, _ptr(src), _ptr(dst), _stream()))
            else:
                msg = f"unsupported dtype {dst.dtype}"
                raise TypeError(msg)

    def __del__(self):
        h = self._handle
        if h is not None and lib is not None:
            if _destroy is None:  # interpreter shutdown: module globals already cleared
                lib.fs_permute_plan_destroy(h)
            else:
                _destroy(lib.fs_permute_plan_destroy, h)
            self._handle = None


_PLAN_CACHE = OrderedDict()
_PLAN_LOCK = threading.Lock()  # host threads of concurrent requests share the cache
_PLAN_CACHE_SIZE = 2048


def abelian_permute_plan(
    shape,
    leg_charges,
    signature,
    old_nrr,
    old_block_irreps,
    old_offsets,
    old_layout_f,
    axes,
    new_nrr,
    *,
    old_dense=False,
    new_dense=False,
):
    """Cached native permute plan (keyed on structure only, never on data)."""
    shape = np.ascontiguousarray(shape, dtype=np.int64)
    sig = np.ascontiguousarray(signature, dtype=np.uint8)
    irr = np.ascontiguousarray(old_block_irreps, dtype=np.int8)
    offs = None if old_offsets is None else np.ascontiguousarray(old_offsets, dtype=np.int64)
    ax = np.ascontiguousarray(axes, dtype=np.int32)
    charges = [np.ascontiguousarray(c, dtype=np.int8) for c in leg_charges]
    key = (
        shape.tobytes(),
        b"|".join(c.tobytes() for c in charges),
        sig.tobytes(),
        int(old_nrr),
        irr.tobytes(),
        None if offs is None else offs.tobytes(),
        bool(old_layout_f),
        ax.tobytes(),
        int(new_nrr),
        bool(old_dense),
        bool(new_dense),
    )
    with _PLAN_LOCK:
        plan = _PLAN_CACHE.get(key)
        if plan is not None:
            _PLAN_CACHE.move_to_end(key)
            return plan
    ptrs = (_lib.c_vp * len(charges))(*[_lib.i8p(c) for c in charges])
    handle = ctypes.c_void_p()
    check(
        lib.fs_abelian_permute_plan_create(
            len(charges),
            _lib.i64p(shape),
            ptrs,
            _lib.u8p(sig),
            int(old_nrr),
            int(irr.size),
            _lib.i8p(irr),
            None if offs is None else _lib.i64p(offs),
            int(bool(old_layout_f)),
            int(bool(old_dense)),
            _lib.i32p(ax),
            int(new_nrr),
            int(bool(new_dense)),
            ctypes.byref(handle),
        )
    )
    PLAN_BUILDS[0] += 1
    plan = PermutePlan(handle.value)
    with _PLAN_LOCK:
        _PLAN_CACHE[key] = plan
        if len(_PLAN_CACHE) > _PLAN_CACHE_SIZE:
            _PLAN_CACHE.popitem(last=False)
    return plan


def permute(plan, src, dtype=DTYPE):
    out = empty(plan.new_ncoeff, dtype=dtype)
    if plan.new_ncoeff:
        plan.execute(src, out)
    return out


def lie_plan_tables(pattern, degen, in_block, block_off, block_rs, block_cs, *, nrows):
    """Host tables (tasks, rows, digits, new block shapes) of a structural
    transform plan for one permute pattern and one set of leg degeneracies
    (fs_lie_plan_tables)."""
    c = np.ascontiguousarray
    degen, in_block = c(degen, dtype=np.int64), c(in_block, dtype=np.int64)
    block_off, block_rs, block_cs = (c(x, dtype=np.int64) for x in (block_off, block_rs, block_cs))
    assert in_block.size == pattern.nt_in
    tasks = np.empty((pattern.ntasks, 8), dtype=np.int64)
    rows = np.empty((int(nrows), 3), dtype=np.int64)
    digits = np.empty((pattern.ntasks * pattern.ndim, 5), dtype=np.int64)
    shapes = np.empty((pattern.nt_out, 2), dtype=np.int64)
    check(lib.fs_lie_plan_tables(
        ctypes.byref(pattern), degen.size, _lib.i64p(degen), _lib.i64p(in_block), block_off.size,
        _lib.i64p(block_off), _lib.i64p(block_rs), _lib.i64p(block_cs),
        _lib.i64p(tasks), rows.shape[0], _lib.i64p(rows), _lib.i64p(digits), _lib.i64p(shapes),
    ))
    return tasks, rows, digits, shapes


class LieWeights:
    """The structural matrices W of one permute pattern on the current device
    (fs_lie_weights), shared by every plan built from the pattern. One object
    per device; plans keep a reference to the object they use."""

    __slots__ = ("_handles", "pattern", "w")

    def __init__(self, pattern, w):
        self.pattern = pattern
        self.w = np.ascontiguousarray(w, dtype=np.float64).reshape(-1)
        self._handles = {}

    def handle(self):
        dev = torch._C._cuda_getDevice()
        h = self._handles.get(dev)
        if h is None:
            out = ctypes.c_void_p()
            check(lib.fs_lie_weights_create(ctypes.byref(self.pattern), self.w.size, _lib.dblp(self.w),
                                            ctypes.byref(out)))
            h = self._handles[dev] = out.value
        return h

    def __del__(self):
        if lib is not None:
            for h in self._handles.values():
                if _destroy is None:  # interpreter shutdown: module globals already cleared
                    lib.fs_lie_weights_destroy(h)
                else:
                    _destroy(lib.fs_lie_weights_destroy, h)
        self._handles = {}


def lie_plan_build(pattern, degen, in_block, block_off, block_rs, block_cs, weights, src_size):
    """(plan, new block shapes) of one permute straight from its pattern and
    leg degeneracies (fs_lie_plan_build: tables and device plan in one call,
    W from the pattern's LieWeights)."""
    c = np.ascontiguousarray
    degen, in_block = c(degen, dtype=np.int64), c(in_block, dtype=np.int64)
    block_off, block_rs, block_cs = (c(x, dtype=np.int64) for x in (block_off, block_rs, block_cs))
    assert in_block.size == pattern.nt_in and weights.pattern is pattern
    shapes = np.empty((pattern.nt_out, 2), dtype=np.int64)
    handle = ctypes.c_void_p()
    check(lib.fs_lie_plan_build(
        ctypes.byref(pattern), degen.size, _lib.i64p(degen), _lib.i64p(in_block), block_off.size,
        _lib.i64p(block_off), _lib.i64p(block_rs), _lib.i64p(block_cs), weights.handle(),
        int(src_size), _lib.i64p(shapes), ctypes.byref(handle),
    ))
    PLAN_BUILDS[0] += 1
    dst = int((shapes[:, 0] * shapes[:, 1]).sum())
    plan = LiePlan._from_handle(handle.value, src_size, dst)
    plan.w = weights  # the plan reads the weights' device copy: keep it alive
    return plan, shapes


class LiePlan:
    """Owner of a native structural-transform plan (fs_lie_plan_*).

    tasks (n, 8), rows (n, 3), digits (n, 5): int64 tables laid out as
    fs_lie_task / fs_lie_row / fs_lie_digit; w: transposed structural matrices.
    Plans built from tables keep them for inspection (None for plans built by
    lie_plan_build)."""

    __slots__ = ("_handle", "coeffs", "digits", "dst_size", "flops", "rows", "src_size", "tasks", "w")

    @classmethod
    def _from_handle(cls, handle, src_size, dst_size):
        self = cls.__new__(cls)
        self.tasks = self.rows = self.digits = self.w = None
        self.src_size = int(src_size)
        self.dst_size = int(dst_size)
        self._handle = handle
        self._info()
        return self

    def __init__(self, tasks, rows, digits, w, src_size, dst_size):
        self._handle = None
        self.tasks = np.ascontiguousarray(tasks, dtype=np.int64).reshape(-1, 8)
        self.rows = np.ascontiguousarray(rows, dtype=np.int64).reshape(-1, 3)
        self.digits = np.ascontiguousarray(digits, dtype=np.int64).reshape(-1, 5)
        self.w = np.ascontiguousarray(w, dtype=np.float64).reshape(-1)
        self.src_size = int(src_size)
        self.dst_size = int(dst_size)
        handle = ctypes.c_void_p()
        check(
            lib.fs_lie_plan_create(
                self.tasks.shape[0], _lib.i64p(self.tasks),
                self.rows.shape[0], _lib.i64p(self.rows),
                self.digits.shape[0], _lib.i64p(self.digits),
                self.w.size, _lib.dblp(self.w),
                self.src_size, self.dst_size, ctypes.byref(handle),
            )
        )
        self._handle = handle.value
        self._info()

    def _info(self):
        nt = ctypes.c_int64()
        fl = ctypes.c_double()
        co = ctypes.c_double()
        check(lib.fs_lie_plan_info(self._handle, ctypes.byref(nt), ctypes.byref(fl), ctypes.byref(co)))
        self.flops = fl.value
        self.coeffs = co.value

    def execute(self, src, dst):
        assert dst.dtype == torch.float64 and dst.numel() >= self.dst_size
        assert src is None or src.numel() >= self.src_size
        if GraphHooks.active is not None:
            GraphHooks.active.keep(self)
        with _timed("lie", 8.0 * self.coeffs):
            check(lib.fs_lie_plan_execute_f64(
                self._handle, None if src is None else _ptr(src), _ptr(dst), _stream()))

    def __del__(self):
        h = self._handle
        if h is not None and lib is not None:
            if _destroy is None:  # interpreter shutdown: module globals already cleared
                lib.fs_lie_plan_destroy(h)
            else:
                _destroy(lib.fs_lie_plan_destroy, h)
            self._handle = None


# ---------------------------------------------------------------------------
# grouped GEMM
# ---------------------------------------------------------------------------


def _gemm_dispatches():
    return lib.fs_gemm_last_dispatches()


def grouped_gemm(groups, A, B, C, *, alpha=1.0, beta=0.0, split=False):
    """groups: (n, 11) int64 array, or a sequence of 11-tuples, (m, n, k, a_off,
    a_rs, a_cs, b_off, b_rs, b_cs, c_off, c_ld). A, B, C: one device buffer
    each, or a list with one buffer per group (the groups then read and write
    blocks of several tensors in one launch)."""
    n = len(groups)
    if n == 0:
        return
    multi = any(isinstance(x, (list, tuple)) for x in (A, B, C))
    if multi:
        As, Bs, Cs = (list(x) if isinstance(x, (list, tuple)) else [x] * n for x in (A, B, C))
        if not (len(As) == len(Bs) == len(Cs) == n):
            raise ValueError("one operand buffer per group is required")
        if split:  # the compensated form has a single-base entry point only
            by = {}
            for g, a, b, c in zip(groups, As, Bs, Cs, strict=True):
                by.setdefault((a.data_ptr(), b.data_ptr(), c.data_ptr()), (a, b, c, []))[3].append(g)
            for a, b, c, gs in by.values():
                grouped_gemm(gs, a, b, c, alpha=alpha, beta=beta, split=True)
            return
    # fs_gemm_group layout: 11 int64 per group
    garr = np.ascontiguousarray(groups, dtype=np.int64).reshape(n, 11)
    arr = _lib._addr(garr)
    def flop():
        return 2.0 * float((garr[:, 0].astype(np.float64) * garr[:, 1] * garr[:, 2]).sum())

    # the library launches one kernel per 32 groups that hold a non-empty group
    if multi:
        # one device address per group and operand, as uint64 arrays read
        # during the call (the address of each distinct tensor looked up once)
        ptrs = {}
        for t in (*As, *Bs, *Cs):
            if id(t) not in ptrs:
                ptrs[id(t)] = _ptr(t)
        bases = [np.fromiter((ptrs[id(t)] for t in lst), dtype=np.uint64, count=n) for lst in (As, Bs, Cs)]
        with _timed("gemm", flop, _gemm_dispatches):
            check(lib.fs_grouped_gemm_f64_multi(n, arr, *(_lib._addr(b) for b in bases), float(alpha), float(beta),
                                                _stream()))
        return
    fn = lib.fs_grouped_gemm_f64_split if split else lib.fs_grouped_gemm_f64
    with _timed("gemm_split" if split else "gemm", flop, _gemm_dispatches):
        check(fn(n, arr, _ptr(A), _ptr(B), _ptr(C), float(alpha), float(beta), _stream()))


# ---------------------------------------------------------------------------
# descriptor tables
# ---------------------------------------------------------------------------
_DESC_DTYPES = {}


def _desc_table(ctype, descs):
    """Descriptors as one numpy structured array laid out like `ctype` (all
    fields 8 bytes: c_int64 or c_double): one conversion per field instead of
    one ctypes object per descriptor. Returns (table, ctypes pointer)."""
    dt = _DESC_DTYPES.get(ctype)
    if dt is None:
        dt = np.dtype([(f, "<i8" if t is ctypes.c_int64 else "<f8") for f, t in ctype._fields_])
        assert dt.itemsize == ctypes.sizeof(ctype)
        _DESC_DTYPES[ctype] = dt
    if isinstance(descs, np.ndarray) and descs.dtype == dt:
        table = descs
    elif isinstance(descs, dict):  # one array (or scalar) per field
        n = max(np.size(v) for v in descs.values())
        table = np.empty((n,), dtype=dt)
        for name in dt.names:
            table[name] = descs[name]
    else:
        cols = descs.T if isinstance(descs, np.ndarray) else list(zip(*descs))
        table = np.empty((len(descs),), dtype=dt)
        for name, col in zip(dt.names, cols, strict=True):
            table[name] = col
    return table, _lib._addr(table)  # the table stays alive with the caller's reference


def _ndesc(descs):
    return max(np.size(v) for v in descs.values()) if isinstance(descs, dict) else len(descs)


# ---------------------------------------------------------------------------
# batched SVD
# ---------------------------------------------------------------------------
SVD_MAX_MIN_DIM = 1024


def batched_svd(descs, A, U, S, VT, work, *, max_sweeps=60):
    """descs: sequence of 9-tuples (m, n, a_off, a_rs, a_cs, u_off, s_off, vt_off, work_off)."""
    n = _ndesc(descs)
    if n == 0:
        return
    tab, arr = _desc_table(_lib.SvdDesc, descs)
    # work model: one Jacobi sweep ~ 3 q^2 p flop (dot products + rotations)
    def flop():
        mn = np.minimum(tab["m"], tab["n"]).astype(np.float64)
        return float((3.0 * mn * mn * np.maximum(tab["m"], tab["n"])).sum())

    with _timed("svd", flop):
        check(
            lib.fs_batched_svd_f64(
                n, arr, _ptr(A), _ptr(U), _ptr(S), _ptr(VT), _ptr(work), int(max_sweeps), _stream()
            )
        )


def batched_qr(descs, A, Q, R, work):
    """descs: sequence of 8-tuples (m, n, a_off, a_rs, a_cs, q_off, r_off, work_off);
    R may be None (only Q is formed)."""
    n = _ndesc(descs)
    if n == 0:
        return
    tab, arr = _desc_table(_lib.QrDesc, descs)
    mn = np.minimum(tab["m"], tab["n"]).astype(np.float64)
    flop = float((4.0 * tab["m"] * mn * mn).sum())  # geqrf + orgqr, m >= n
    with _timed("qr", flop):
        check(lib.fs_batched_qr_f64(n, arr, _ptr(A), _ptr(Q), _ptr(R), _ptr(work), _stream()))


def batched_cholesky(descs, G, R, status):
    """descs: sequence of (l, g_off, r_off, shift_rel); status: device int32 (len(descs))."""
    n = _ndesc(descs)
    if n == 0:
        return
    tab, arr = _desc_table(_lib.CholDesc, descs)
    flop = float((tab["l"].astype(np.float64) ** 3 / 3.0).sum())
    with _timed("cholesky", flop):
        check(lib.fs_batched_cholesky_f64(n, arr, _ptr(G), _ptr(R), _ptr(status), _stream()))


CHOLINV_MAX_ORDER = 128


def batched_cholinv(descs, G, R, Rinv, status):
    """descs: (l, g_off, g_stride, n_part, r_off, rinv_off, shift_rel): G is the sum
    of n_part l x l partials at g_off + p*g_stride; R may be None. Writes the upper
    Cholesky factor (if R) and its inverse; status: device int32 (len(descs))."""
    n = _ndesc(descs)
    if n == 0:
        return
    tab, arr = _desc_table(_lib.CholInvDesc, descs)
    with _timed("cholinv", lambda: float((2.0 * tab["l"].astype(np.float64) ** 3 / 3.0).sum())):
        check(lib.fs_batched_cholinv_f64(
            n, arr, _ptr(G), None if R is None else _ptr(R), None if Rinv is None else _ptr(Rinv),
            _ptr(status), _stream()))


def batched_triinv(descs, R, X, status=None):
    """descs: (l, r_off, x_off) per matrix (or a dict of columns): X = R^-1 for
    upper triangular R; matrices whose status entry is nonzero get NaN."""
    n = _ndesc(descs)
    if n == 0:
        return
    tab, arr = _desc_table(_lib.TriInvDesc, descs)
    with _timed("triinv", lambda: float((tab["l"].astype(np.float64) ** 3 / 3.0).sum())):
        check(lib.fs_batched_triinv_f64(n, arr, _ptr(R), _ptr(X), _ptr(status), _stream()))


def cholqr(descs, Y, Q, Rt, status, *, passes=3, shifted=True):
    """descs: (m, l, y_off, y_rs, y_cs, q_off, rt_off) per matrix (or a dict of
    columns): multi-pass shifted CholeskyQR of every Y_i (m x l, l <= 128) in
    one native call, Q packed row-major at q_off, the total triangular factor
    at rt_off when Rt is given (passes >= 2); status: device int32 of
    passes * len(descs), pass-major."""
    n = _ndesc(descs)
    if n == 0:
        return
    tab, arr = _desc_table(_lib.CholQrDesc, descs)
    size = ctypes.c_int64()
    check(lib.fs_cholqr_work_size(n, arr, int(passes), int(Rt is not None), ctypes.byref(size)))
    work = empty(int(size.value))
    def flop():
        m = tab["m"].astype(np.float64)
        l = tab["l"].astype(np.float64)
        return float(passes * (4.0 * m * l * l + 2.0 * l**3 / 3.0).sum())

    with _timed("cholqr", flop):
        check(lib.fs_cholqr_f64(
            n, arr, _ptr(Y), _ptr(Q), None if Rt is None else _ptr(Rt), _ptr(work), int(size.value),
            int(passes), int(bool(shifted)), _ptr(status), _stream()))


def batched_trsm_right_upper(descs, Y, R, O):
    """descs: sequence of (m, l, y_off, y_rs, y_cs, r_off, o_off): O = Y R^-1."""
    n = _ndesc(descs)
    if n == 0:
        return
    tab, arr = _desc_table(_lib.TrsmDesc, descs)
    f0 = tab[tab.dtype.names[0]].astype(np.float64)
    f1 = tab[tab.dtype.names[1]].astype(np.float64)
    flop = float((f0 * f1 * f1).sum())
    with _timed("trsm", flop):
        check(lib.fs_batched_trsm_right_upper_f64(n, arr, _ptr(Y), _ptr(R), _ptr(O), _stream()))


# ---------------------------------------------------------------------------
# elementwise helpers
# ---------------------------------------------------------------------------
def weighted_norm2(x, seg_off, weights):
    """Device scalar tensor sum_b w_b ||x[seg_b]||^2 (segments contiguous, ascending)."""
    seg = np.ascontiguousarray(seg_off, dtype=np.int64)
    w = np.ascontiguousarray(weights, dtype=np.float64)
    out = empty(1)
    partial = empty(_lib.FS_NORM_PARTIALS)
    check(
        lib.fs_weighted_norm2_f64(
            int(w.size), _lib.i64p(seg), _lib.dblp(w), _ptr(x) if x.numel() else None,
            _ptr(partial), _ptr(out), _stream(),
        )
    )
    return out


def scale_(x, alpha):
    if x.numel():
        check(lib.fs_scale_f64(int(x.numel()), float(alpha), _ptr(x), _stream()))


def scale_rsqrt_dev_(x, d):
    if x.numel():
        check(lib.fs_scale_rsqrt_dev_f64(int(x.numel()), _ptr(d), _ptr(x), _stream()))


def axpby_(alpha, x, beta, y):
    if y.numel():
        check(lib.fs_axpby_f64(int(y.numel()), float(alpha), _ptr(x), float(beta), _ptr(y), _stream()))


def copy2d(descs, src, dst, *, row_w=None, col_w=None, col_idx=None):
    """descs: sequence of 13-tuples
    (rows, cols, s_off, s_rs, s_cs, d_off, d_rs, d_cs, rw_off, cw_off, ci_off, alpha, beta)."""
    n = _ndesc(descs)
    if n == 0:
        return
    tab, arr = _desc_table(_lib.Copy2dDesc, descs)
    def nbytes():
        names = tab.dtype.names
        per = np.where(tab[names[12]] == 0.0, 16.0, 24.0)
        return float((per * tab[names[0]] * tab[names[1]]).sum())

    with _timed("copy2d", nbytes):
        check(
            lib.fs_batched_copy2d_f64(
                n, arr, _ptr(src), _ptr(dst), _ptr(row_w), _ptr(col_w), _ptr(col_idx), _stream()
            )
        )


def gather2d(descs, idx, w, src, dst):
    """Two-term weighted gather (fs_batched_gather2d_f64). descs: sequences of
    (rows, cols, src_rows, src_cols, s_off, s_rs, s_cs, d_off, d_rs, d_cs, ri_off,
    rw_off, ci_off, cw_off, beta); idx (int64) and w (float64) are host arrays,
    validated against the buffer sizes before upload."""
    n = _ndesc(descs)
    if n == 0:
        return
    tab, arr = _desc_table(_lib.Gather2dDesc, descs)
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    check(lib.fs_gather2d_check(n, arr, _lib.i64p(idx), idx.size, w.size, src.numel(), dst.numel()))
    idx_dev = upload(idx, dtype=torch.int64)
    w_dev = upload(w)
    names = tab.dtype.names
    moved = float((tab[names[0]].astype(np.float64) * tab[names[1]]).sum())
    with _timed("gather2d", 16.0 * moved):
        check(lib.fs_batched_gather2d_f64(n, arr, _ptr(idx_dev), _ptr(w_dev), _ptr(src), _ptr(dst), _stream()))


def weighted_trace(x, offs, ns, lds, weights):
    """Device scalar sum_b w_b trace(block_b)."""
    offs = np.ascontiguousarray(offs, dtype=np.int64)
    ns = np.ascontiguousarray(ns, dtype=np.int64)
    lds = np.ascontiguousarray(lds, dtype=np.int64)
    w = np.ascontiguousarray(weights, dtype=np.float64)
    out = empty(1)
    total = zeros(1)
    for s in range(0, max(len(w), 1), 64):
        sl = slice(s, s + 64)
        # the chunks stay bound to names until the call returns: the library
        # receives bare addresses
        o_s, n_s, l_s, w_s = (np.ascontiguousarray(a[sl]) for a in (offs, ns, lds, w))
        check(
            lib.fs_weighted_trace_f64(
                int(w_s.size), _lib.i64p(o_s), _lib.i64p(n_s), _lib.i64p(l_s), _lib.dblp(w_s),
                _ptr(x) if x.numel() else None, _ptr(out), _stream(),
            )
        )
        axpby_(1.0, out, 1.0, total)
    return total


def fill_normal_(x, seed):
    if x.numel():
        check(lib.fs_fill_normal_f64(int(x.numel()), int(seed) & 0xFFFFFFFFFFFFFFFF, _ptr(x), _stream()))


def to_f32(x):
    y = empty(x.numel(), dtype=torch.float32)
    if x.numel():
        check(lib.fs_convert_f64_to_f32(int(x.numel()), _ptr(x), _ptr(y), _stream()))
    return y


def to_f64(x):
    y = empty(x.numel())
    if x.numel():
        check(lib.fs_convert_f32_to_f64(int(x.numel()), _ptr(x), _ptr(y), _stream()))
    return y
