"""Continuous-batching serving: requests admitted by HBM budget overlap on the device, and each result is
the same alone and in the mix (bit-identical: every kernel of the operator is deterministic)."""
import numpy as np
import pytest

from purify_amd import serving

pytestmark = pytest.mark.gpu


def small_mix(seed, n_requests=10):
    return serving.poisson_requests(seed, n_requests, rate_per_s=1e5, sizes=(64, 128), M_range=(1e3, 3e4),
                                    w_fraction=0.3)


def run(reqs, budget, inflight):
    srv = serving.Server(budget, max_inflight=inflight)
    futs = [srv.submit(r) for r in reqs]
    res = [f.result(timeout=120) for f in futs]
    srv.close()
    return res, srv


def test_results_alone_and_in_the_mix_are_identical():
    reqs = small_mix(7)
    assert any(r.w_term for r in reqs) and any(not r.w_term for r in reqs)
    alone = []
    for r in small_mix(7):
        r.arrival = 0.0
        (res,), _ = run([r], 1 << 34, 1)
        alone.append(res)
    mixed, srv = run(reqs, 1 << 34, 4)
    assert srv.peak_inflight > 1  # the requests did overlap
    for a, b in zip(alone, mixed):
        assert a.rid == b.rid
        assert np.array_equal(a.y, b.y) and np.array_equal(a.x, b.x)
        assert b.finished >= b.admitted >= b.arrival - 1e-3
    stats = serving.latency_stats(mixed)
    assert stats["requests"] == len(reqs) and stats["p99_s"] >= stats["p50_s"] > 0


def test_admission_respects_the_hbm_budget():
    reqs = small_mix(11, 8)
    need = [serving.estimate_bytes(r) for r in reqs]
    budget = 2 * max(need)
    res, srv = run(reqs, budget, 4)
    assert len(res) == len(reqs)
    assert srv.peak_in_use <= budget
    # the estimate bounds what the operator actually holds
    for r, n in zip(res, need):
        assert r.device_bytes <= n
    with pytest.raises(ValueError):
        srv2 = serving.Server(min(need) // 2)
        try:
            srv2.submit(reqs[0])
        finally:
            srv2.close()


def test_padmm_requests_alone_and_in_the_mix_are_identical():
    def mix():
        return serving.poisson_requests(13, 6, rate_per_s=1e5, sizes=(64, 128), M_range=(2e3, 2e4), w_fraction=0.3,
                                        padmm_fraction=0.5, padmm_iters=20)
    reqs = mix()
    kinds = [r.kind for r in reqs]
    assert "padmm" in kinds and "pair" in kinds
    alone = []
    for r in mix():
        r.arrival = 0.0
        (res,), _ = run([r], 1 << 34, 1)
        alone.append(res)
    mixed, srv = run(reqs, 1 << 34, 3)
    for a, b, r in zip(alone, mixed, reqs):
        assert np.array_equal(a.x, b.x)
        if r.kind == "padmm":
            assert a.y is None and np.isfinite(b.x).all() and np.linalg.norm(b.x) > 0
