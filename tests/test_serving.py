"""Continuous-batching request loop: each request's result is the same alone
and in the mix (energy to 1e-12), admission respects the HBM budget, and the
summary carries the latency percentiles and throughput."""

import numpy as np
import pytest

from frostspin_amd import serving

BACKENDS = [
    pytest.param("emulated", id="emulated"),
    pytest.param("gpu", id="gpu", marks=pytest.mark.gpu),
]

# (beta, chi, D) per request; emulated runs use small sizes
SIZES = {
    "emulated": [(0.2, 6, 3), (0.4, 10, 4), (0.3, 8, 3)],
    "gpu": [(0.5, 32, 6), (2.0, 64, 6), (1.0, 48, 7), (3.0, 64, 8)],
}


@pytest.fixture(params=BACKENDS)
def backend(request):
    if request.param == "emulated":
        request.getfixturevalue("emulated_device")
    else:
        request.getfixturevalue("gpu")
    return request.param


def _requests(backend, *, spread):
    return [
        serving.Request(rid=i, beta=b, chi=c, D=d, n_iter=3, arrival=spread * i)
        for i, (b, c, d) in enumerate(SIZES[backend])
    ]


def test_results_alone_equal_results_in_mix(backend):
    alone = {}
    for req in _requests(backend, spread=0.0):
        res, summary = serving.ServingEngine(1).run([req])
        assert summary["requests"] == 1
        alone[req.rid] = res[0].energy
    mixed, summary = serving.ServingEngine(len(SIZES[backend])).run(_requests(backend, spread=0.01))
    for r in mixed:
        assert abs(r.energy - alone[r.rid]) <= 1e-12 * max(1.0, abs(alone[r.rid]))
        assert -0.75 < r.energy < 0.25
        assert r.finish >= r.start >= r.arrival - 1e-9
    # several slots were busy at the same time
    spans = sorted((r.start, r.finish) for r in mixed)
    assert any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))
    for key in ("latency_p50_s", "latency_p99_s", "requests_per_s", "ctmrg_iters_per_s"):
        assert np.isfinite(summary[key]) and summary[key] > 0


def test_admission_by_hbm_footprint(backend):
    reqs = _requests(backend, spread=0.0)
    assert all(serving.footprint_bytes(r) > 1 for r in reqs)
    # a budget below every footprint: a request only starts on an idle GPU
    res, summary = serving.ServingEngine(len(reqs), hbm_budget=1).run(reqs)
    spans = sorted((r.start, r.finish) for r in res)
    for (s0, f0), (s1, _) in zip(spans, spans[1:]):
        assert s1 >= f0 - 1e-6
    assert summary["requests"] == len(reqs)


def test_poisson_request_stream():
    reqs = serving.poisson_requests(64, seed=3)
    assert len(reqs) == 64
    arr = np.array([r.arrival for r in reqs])
    assert (np.diff(arr) > 0).all()
    assert 0.02 < np.mean(np.diff(arr)) < 0.1  # mean 50 ms
    assert all(0.5 <= r.beta <= 4.0 and r.chi in (32, 64, 128, 192, 256) and 6 <= r.D <= 8 for r in reqs)


@pytest.mark.gpu
def test_process_slots_match_in_process_results(gpu):
    """Worker-process slots (own HIP context each) return the energies of the
    same requests solved in this process, to 1e-12."""
    reqs = _requests("gpu", spread=0.005)
    alone = {}
    for req in reqs:
        res, _ = serving.ServingEngine(1).run([req])
        alone[req.rid] = res[0].energy
    with serving.ServingEngine(3, mode="process", start_timeout=300) as engine:
        mixed, summary = engine.run(reqs)
    assert summary["requests"] == len(reqs)
    for r in mixed:
        assert abs(r.energy - alone[r.rid]) <= 1e-12 * max(1.0, abs(alone[r.rid]))


def test_overlapped_rate_counts_common_window():
    """Aggregate rate over the interval where all requests are past warm-up."""
    mk = lambda rid, ts: serving.Result(rid, 0.0, [], 0.0, 0.0, ts[-1], 0, list(ts))  # noqa: E731
    a = mk(0, [1.0, 2.0, 3.0, 4.0, 5.0])
    b = mk(1, [1.5, 2.5, 3.5, 4.5])
    rate, window = serving.overlapped_rate([a, b], 1)
    # both past their first iteration from 1.5 on, b stops at 4.5: window (1.5, 4.5]
    # holds a's 2.0, 3.0, 4.0 and b's 2.5, 3.5, 4.5
    assert window == pytest.approx(3.0)
    assert rate == pytest.approx(6 / 3.0)
    with pytest.raises(ValueError):
        serving.overlapped_rate([mk(0, [1.0, 2.0]), mk(1, [3.0, 4.0])], 1)
