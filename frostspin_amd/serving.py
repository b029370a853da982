"""Continuous-batching request loop for finite-temperature CTMRG solves.

A request asks for the thermal state of a spin-1/2 Heisenberg (SU(2)) or XXZ
(U(1)) square-lattice model at inverse temperature beta: simple update from
beta = 0 at bond dimension D, then CTMRG at corner dimension chi for a fixed
number of iterations, then the bond energy. Requests of differing beta, D,
couplings and chi are served together:

* the engine owns S slots. In "process" mode (the serving mode) each slot is a
  worker process with its own HIP context: the host-side work of a request
  (Python planning, launches) runs on its own CPU core, and the kernels of
  different requests overlap on the GPU (most CTMRG kernels leave most of the
  256 CUs idle when run alone). In "thread" mode each slot is a host thread
  with its own HIP stream in the calling process (no process start-up, but
  the host work of all slots shares one interpreter lock);
* environments and all intermediate tensors stay in HBM for the life of the
  request; nothing is staged through the host except the truncation data the
  algorithm itself reads back;
* admission is by HBM footprint: a waiting request is started when a slot is
  free and its estimated peak footprint fits next to those already running
  (first come, first served among the requests that fit); requests that
  arrive while others are on the GPU join as soon as that holds;
* every request records its arrival, start and finish times, so latency
  percentiles and throughput come out of the same run;
* in process mode the workers share the irreps-only permute structures they
  build through one directory (plan_store): a pattern computed by one slot is
  loaded by the others.

Each request's launches are ordered on its own stream and share no mutable
device data with other requests, so its result does not depend on what else
is in flight (checked by the tests: alone and in the mix agree).
"""

import dataclasses
import multiprocessing as mp
import os
import queue
import shutil
import tempfile
import threading
import time

import numpy as np
import torch

from . import _ops
from .ctmrg import models
from .ctmrg.sequential_ctmrg import SequentialCTMRG


@dataclasses.dataclass
class Request:
    """One thermal CTMRG solve."""

    rid: int
    beta: float
    chi: int
    D: int
    n_iter: int = 10
    symmetry: str = "SU2"
    jz: float = 1.0  # XXZ couplings (U(1) only); SU(2) is the Heisenberg point
    jperp: float = 1.0
    tau: float = 0.01
    arrival: float = 0.0  # seconds after the start of the run
    record_iters: bool = False  # keep the completion time of every CTMRG iteration


@dataclasses.dataclass
class Result:
    rid: int
    energy: float
    chi_values: list
    arrival: float
    start: float
    finish: float
    slot: int
    iter_times: list = dataclasses.field(default_factory=list)  # monotonic seconds

    @property
    def latency(self):
        return self.finish - self.arrival

    @property
    def queue_time(self):
        return self.start - self.arrival


def footprint_bytes(req):
    """Estimated peak HBM use of a request: the four enlarged corners of a site
    pair at (chi D^2)^2 dense-equivalent entries each, times the block fraction
    that survives the symmetry (about 1/4 for U(1), 1/16 for SU(2) at these
    sizes), plus a factor 3 for the transients of a contraction."""
    frac = {"SU2": 1 / 16, "U1": 1 / 4}.get(req.symmetry, 1.0)
    side = req.chi * req.D**2
    return int(3 * 4 * side * side * 8 * frac)


def solve(req, iter_times=None):
    """Simple update to beta, n_iter CTMRG iterations, bond energy (on the
    current device stream). With a list in iter_times, the completion time
    (time.monotonic, after a stream synchronization) of every iteration is
    appended to it."""
    if req.symmetry == "SU2":
        h = models.heisenberg_bond("SU2")
    elif req.symmetry == "U1":
        h = models.xxz_bond(req.jz, req.jperp)
    else:
        raise ValueError(f"unsupported symmetry {req.symmetry}")
    _, tensors = models.thermal_tensors(h, req.D, req.beta, tau=req.tau)
    ctm = SequentialCTMRG.from_elementary_tensors("AB\nBA", tensors, req.chi, cutoff=1e-10, degen_ratio=0.99)
    # every iteration eager: a serving worker process died once while
    # recording graphs in the 8-slot serving benchmark, cause not found
    ctm.use_graphs = False
    for _ in range(req.n_iter):
        ctm.iterate()
        if iter_times is not None:
            torch.cuda.current_stream().synchronize()
            iter_times.append(time.monotonic())
    return models.bond_energy(ctm, h), list(ctm.chi_values)


def _summary(requests, results, wall, n_slots):
    lat = np.array([r.latency for r in results])
    iters = sum(r.n_iter for r in requests)
    return {
        "requests": len(results),
        "slots": n_slots,
        "wall_s": wall,
        "requests_per_s": len(results) / wall,
        "ctmrg_iters_per_s": iters / wall,
        "latency_p50_s": float(np.percentile(lat, 50)),
        "latency_p99_s": float(np.percentile(lat, 99)),
        "queue_p50_s": float(np.percentile([r.queue_time for r in results], 50)),
    }


def host_thread_share(n_slots):
    """Host threads (BLAS, OpenMP, torch CPU ops) for each of n_slots worker
    processes: the process's CPU share (OMP_NUM_THREADS when set, else the
    CPUs it may run on) split evenly, at least one. Unbounded, every worker
    would start a full-width BLAS pool for its small host-side linear algebra
    and the slots would oversubscribe the cores."""
    share = os.environ.get("OMP_NUM_THREADS")
    total = int(share) if share and share.isdigit() else len(os.sched_getaffinity(0))
    return max(1, total // max(1, n_slots))


def hw_queue_share(n_slots):
    """Hardware queues (GPU_MAX_HW_QUEUES) for each of n_slots worker
    processes: 16 split evenly, at most HIP's default of 4, at least 1. Every
    process maps its streams onto that many hardware queues; once all slots
    together ask for more than the device schedules at once, the driver
    time-slices the queues and every slot slows down (8 slots x 4 queues:
    35.5 aggregate iters/s at the SU(2) chi=256 headline, against 72-76 with 1-2
    queues each)."""
    return max(1, min(4, 16 // max(1, n_slots)))


def _worker_main(slot, q_in, q_out, debug_checks, plan_dir=None, host_threads=None, hw_queues=None):
    """Slot process: serve requests from q_in until None arrives. With a
    plan_dir, the irreps-only permute structures are shared with the other
    slots through that directory; host_threads bounds its host thread pools,
    hw_queues its hardware queues (set before the process's first HIP call)."""
    if hw_queues:
        os.environ["GPU_MAX_HW_QUEUES"] = str(int(hw_queues))
    import torch  # noqa: PLC0415 - imported in the worker process

    from .symmetric_tensor import plan_store
    from .symmetric_tensor.tools import config

    if host_threads:
        from threadpoolctl import threadpool_limits  # noqa: PLC0415

        threadpool_limits(host_threads)
        torch.set_num_threads(host_threads)
    config["debug_checks"] = debug_checks
    plan_store.set_plan_dir(plan_dir)
    try:
        total = int(torch.cuda.get_device_properties(0).total_memory)
        q_out.put(("ready", slot, total, None, 0.0, 0.0))
    except Exception as err:  # noqa: BLE001 - reported to the engine
        q_out.put(("error", slot, repr(err), None, 0.0, 0.0))
        return
    while True:
        req = q_in.get()
        if req is None:
            return
        start = time.monotonic()
        try:
            its = [] if req.record_iters else None
            energy, chis = solve(req, its)
            q_out.put(("done", slot, req.rid, (energy, chis, its or [], plan_store.stats()), start,
                       time.monotonic()))
        except Exception as err:  # noqa: BLE001 - reported to the engine
            q_out.put(("failed", slot, req.rid, repr(err), start, time.monotonic()))


class ServingEngine:
    """S slots over one GPU; see the module docstring. Use as a context manager
    (or call close()) in process mode so the workers exit."""

    def __init__(self, n_slots=4, *, hbm_budget=None, mode="thread", start_timeout=600.0, plan_dir=None,
                 share_plans=True):
        self.n_slots = int(n_slots)
        if self.n_slots < 1:
            raise ValueError("need at least one slot")
        if mode not in ("thread", "process"):
            raise ValueError(f"unknown mode {mode}")
        self.mode = mode
        self._lock = threading.Condition()
        self._waiting = []  # admitted in arrival order when they fit
        self._running_bytes = 0
        self._free_slots = list(range(self.n_slots))
        self._results = {}
        self._errors = []
        self._t0 = None
        self._workers = []
        self._plan_tmp = None
        self.worker_plan_stats = {}  # slot -> latest plan_store.stats() of that worker
        if mode == "process" and share_plans and plan_dir is None:
            # one directory for this engine's workers (and removed with it),
            # unless the caller or FROSTSPIN_AMD_PLAN_DIR names a lasting one
            from .symmetric_tensor import plan_store

            plan_dir = plan_store.plan_dir()
            if plan_dir is None:
                self._plan_tmp = tempfile.mkdtemp(prefix="fs_plans_")
                plan_dir = self._plan_tmp
        self.plan_dir = plan_dir if share_plans else None
        if mode == "process":
            # workers are started before this process touches the GPU: they
            # report the device memory once their HIP context exists
            from .symmetric_tensor.tools import config

            ctx = mp.get_context("spawn")
            self._q_out = ctx.Queue()
            self._q_in = [ctx.Queue() for _ in range(self.n_slots)]
            self._workers = [
                ctx.Process(target=_worker_main,
                            args=(i, self._q_in[i], self._q_out, config["debug_checks"], self.plan_dir,
                                  host_thread_share(self.n_slots), hw_queue_share(self.n_slots)),
                            daemon=True)
                for i in range(self.n_slots)
            ]
            for w in self._workers:
                w.start()
            total = None
            deadline = time.monotonic() + start_timeout
            for _ in range(self.n_slots):
                try:
                    kind, slot, payload, *_ = self._q_out.get(timeout=max(1.0, deadline - time.monotonic()))
                except queue.Empty:
                    self.close()
                    raise RuntimeError("serving workers did not start") from None
                if kind != "ready":
                    self.close()
                    raise RuntimeError(f"serving worker {slot} failed to start: {payload}")
                total = payload
            self._streams = None
            if hbm_budget is None:
                hbm_budget = int(0.8 * total)
        else:
            dev = _ops.device()
            on_gpu = dev.type == "cuda"
            if hbm_budget is None:
                hbm_budget = int(0.8 * torch.cuda.get_device_properties(dev).total_memory) if on_gpu else 8 << 30
            self._streams = [torch.cuda.Stream() if on_gpu else None for _ in range(self.n_slots)]
        self.hbm_budget = int(hbm_budget)

    def close(self):
        for q, w in zip(getattr(self, "_q_in", []), self._workers, strict=False):
            if w.is_alive():
                q.put(None)
        for w in self._workers:
            w.join(timeout=30)
            if w.is_alive():
                w.terminate()
        self._workers = []
        if self._plan_tmp is not None:
            shutil.rmtree(self._plan_tmp, ignore_errors=True)
            self._plan_tmp = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _now(self):
        return time.perf_counter() - self._t0

    def _fits(self, need):
        return self._running_bytes + need <= self.hbm_budget or self._running_bytes == 0

    # thread mode -----------------------------------------------------------
    def _admit(self):
        """Start every waiting request that fits (called with the lock held)."""
        i = 0
        while i < len(self._waiting) and self._free_slots:
            req = self._waiting[i]
            need = footprint_bytes(req)
            if self._fits(need):
                self._waiting.pop(i)
                slot = self._free_slots.pop(0)
                self._running_bytes += need
                threading.Thread(target=self._work, args=(req, slot, need), daemon=True).start()
            else:
                i += 1

    def _work(self, req, slot, need):
        start = self._now()
        try:
            stream = self._streams[slot]
            its = [] if req.record_iters else None
            if stream is not None:
                with torch.cuda.stream(stream):
                    energy, chis = solve(req, its)
                    stream.synchronize()
            else:
                energy, chis = solve(req, its)
            offset = time.monotonic() - time.perf_counter() + self._t0
            res = Result(req.rid, energy, chis, req.arrival, start, self._now(), slot, [t - offset for t in its or []])
        except Exception as err:  # noqa: BLE001 - reported by run()
            res = None
            with self._lock:
                self._errors.append((req.rid, err))
        with self._lock:
            if res is not None:
                self._results[req.rid] = res
            self._running_bytes -= need
            self._free_slots.append(slot)
            self._admit()
            self._lock.notify_all()

    def run(self, requests):
        """Serve requests (each arriving `arrival` seconds after the call) and
        return their results in request order plus a summary dict."""
        requests = sorted(requests, key=lambda r: r.arrival)
        self._results, self._errors = {}, []
        if self.mode == "process":
            return self._run_process(requests)
        self._t0 = time.perf_counter()
        for req in requests:
            delay = req.arrival - self._now()
            if delay > 0:
                time.sleep(delay)
            with self._lock:
                self._waiting.append(req)
                self._admit()
        with self._lock:
            while len(self._results) + len(self._errors) < len(requests):
                self._lock.wait()
        if self._errors:
            rid, err = self._errors[0]
            raise RuntimeError(f"request {rid} failed: {err}") from err
        wall = self._now()
        results = [self._results[r.rid] for r in requests]
        return results, _summary(requests, results, wall, self.n_slots)

    # process mode ----------------------------------------------------------
    def _run_process(self, requests):
        if not self._workers:
            raise RuntimeError("engine is closed")
        t0 = time.monotonic()
        arrivals = list(requests)
        waiting, running = [], {}  # running: slot -> (req, need)
        by_rid = {r.rid: r for r in requests}
        while arrivals or waiting or running:
            now = time.monotonic() - t0
            while arrivals and arrivals[0].arrival <= now:
                waiting.append(arrivals.pop(0))
            i = 0
            while i < len(waiting) and self._free_slots:
                need = footprint_bytes(waiting[i])
                if self._fits(need):
                    req = waiting.pop(i)
                    slot = self._free_slots.pop(0)
                    self._running_bytes += need
                    running[slot] = (req, need)
                    self._q_in[slot].put(req)
                else:
                    i += 1
            timeout = (arrivals[0].arrival - now) if arrivals else 1.0
            try:
                kind, slot, rid, payload, start, finish = self._q_out.get(timeout=max(1e-3, min(timeout, 1.0)))
            except queue.Empty:
                if not all(w.is_alive() for w in self._workers):
                    raise RuntimeError("a serving worker died") from None
                continue
            req, need = running.pop(slot)
            self._running_bytes -= need
            self._free_slots.append(slot)
            if kind != "done":
                raise RuntimeError(f"request {rid} failed: {payload}")
            energy, chis, its, pstats = payload
            self.worker_plan_stats[slot] = pstats
            self._results[rid] = Result(rid, energy, chis, by_rid[rid].arrival, start - t0, finish - t0, slot,
                                        [t - t0 for t in its])
        wall = time.monotonic() - t0
        results = [self._results[r.rid] for r in requests]
        summary = _summary(requests, results, wall, self.n_slots)
        stats = list(self.worker_plan_stats.values())
        if stats:
            summary["plan_structures"] = {
                "shared_dir": self.plan_dir is not None,
                "built": sum(x["built"] for x in stats),
                "loaded": sum(x["loaded"] for x in stats),
                "build_s_per_worker": float(np.mean([x["build_s"] for x in stats])),
                "load_s_per_worker": float(np.mean([x["load_s"] for x in stats])),
            }
        return results, summary


def overlapped_rate(results, skip):
    """Aggregate iterations/s over the interval in which every request was past
    its first `skip` iterations and not yet finished: iterations completed by
    all requests inside [max of their start times, min of their end times]
    divided by its length (steady-state throughput of concurrent solves)."""
    t_a = max(r.iter_times[skip - 1] if skip else r.start for r in results)
    t_b = min(r.iter_times[-1] for r in results)
    if t_b <= t_a:
        raise ValueError("the requests' timed phases do not overlap")
    n = sum(sum(1 for t in r.iter_times[skip:] if t_a < t <= t_b) for r in results)
    return n / (t_b - t_a), t_b - t_a


def poisson_requests(n, *, mean_interarrival=0.05, betas=(0.5, 4.0), chis=(32, 64, 128, 192, 256), Ds=(6, 7, 8),
                     n_iter=10, symmetry="SU2", seed=0):
    """n requests with exponential inter-arrival times, beta uniform in betas,
    chi and D drawn from the given sets."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(mean_interarrival, size=n))
    return [
        Request(
            rid=i, beta=float(rng.uniform(*betas)), chi=int(rng.choice(chis)), D=int(rng.choice(Ds)),
            n_iter=n_iter, symmetry=symmetry, arrival=float(t[i]),
        )
        for i in range(n)
    ]
