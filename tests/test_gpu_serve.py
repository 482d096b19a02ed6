"""Serving loop: requests from host buffers, continuous batching, admission by
HBM footprint.

A request's result must not depend on what it was batched with: each request
is run alone, then all of them together, and the two results must be equal
bit for bit and equal to the oracle (or, at the largest point, to the plain
per-call path, itself checked against the oracle elsewhere).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MIX_POINTS = [(8192, 438), (32768, 868), (65536, 1747), (1024, 256)]


def _catalog(O, points):
    return [(N, W, O.spec_q(N, W)) for N, W in points]


def test_serve_solo_equals_mix(O, pm):
    cat = _catalog(O, MIX_POINTS)
    kinds = ["mixed", "uniform", "max", "sparse", "small16"]
    reqs = []
    for i in range(12):
        p = i % len(cat)
        N, W, q = cat[p]
        reqs.append((p, *O.operands(kinds[i % len(kinds)], N, W, q, 300 + i)))
    with pm.Server(cat) as S:
        solo = []
        for p, a, b in reqs:
            print("solo point", p, flush=True)
            c, ms = S.wait(S.submit(p, a, b))
            assert ms > 0
            solo.append(c)
        st = S.stats(reset=True)
        assert st["completed"] == len(reqs) and st["batches"] == len(reqs)
        print("mix of", len(reqs), flush=True)
        tickets = [S.submit(p, a, b) for p, a, b in reqs]
        mix = [S.wait(t)[0] for t in tickets]
        st = S.stats()
        print("stats", st, flush=True)
        assert st["completed"] == len(reqs) and st["failed"] == 0
        assert st["batches"] < len(reqs), "requests submitted together should share batches"
    for i, (p, a, b) in enumerate(reqs):
        assert np.array_equal(solo[i], mix[i]), "request %d differs alone and in the mix" % i
    for i, (p, a, b) in enumerate(reqs):
        N, W, q = cat[p]
        if N <= 32768 or i == 2:
            assert np.array_equal(mix[i], O.OraclePlan(N, W, q).multiply(a, b)), "request %d vs oracle" % i
        else:
            with pm.Multiplier(N, W, O.q_limbs(q, O.limbs_of(W))) as m:
                assert np.array_equal(mix[i], m.run(a, b)), "request %d vs plain path" % i


def test_serve_admission_under_tiny_budget(O, pm):
    """A device budget that holds one N=32768 request at a time: every request
    still completes, correctly, and the queue head waits for memory."""
    cat = _catalog(O, [(32768, 868), (2048, 300), (65536, 1747)])
    N, W, q = cat[0]
    L, K = O.limbs_of(W), 57
    one = 2 * N * L * 4 + 2 * K * N * 4
    budget = int(one * 1.5)
    reqs = [(0, *O.operands("uniform", N, W, q, 400 + i)) for i in range(5)]
    N1, W1, q1 = cat[1]
    reqs += [(1, *O.operands("mixed", N1, W1, q1, 410 + i)) for i in range(3)]
    with pm.Server(cat, device_budget=budget, max_inflight=2) as S:
        N3, W3, q3 = cat[2]
        a3, b3 = O.operands("uniform", N3, W3, q3, 1)
        with pytest.raises(pm.PolymultError, match="above the server's budget"):
            S.submit(2, a3, b3)
        tickets = [S.submit(p, a, b) for p, a, b in reqs]
        outs = [S.wait(t)[0] for t in tickets]
        st = S.stats()
        print("stats", st, flush=True)
        assert st["completed"] == len(reqs) and st["failed"] == 0
        assert st["device_peak"] <= budget
        assert st["deferred"] > 0
    for (p, a, b), c in zip(reqs, outs):
        N, W, q = cat[p]
        assert np.array_equal(c, O.OraclePlan(N, W, q).multiply(a, b))


def test_serve_poisson_drive_short(O, pm):
    """Two seconds of open-loop Poisson load on the serving mix: every request
    completes, none fails, every result equals the plain path's product of
    its operands (checked by the load driver, pm_server_drive), percentiles are ordered."""
    cat = _catalog(O, MIX_POINTS[:3])
    with pm.Server(cat) as S:
        rep = S.drive(500.0, 2.0, [0.70, 0.25, 0.05], seed=7)
        print("drive", rep, flush=True)
        assert rep["requests"] > 700
        assert rep["completed"] == rep["requests"] and rep["failed"] == 0
        assert rep["verified"] == rep["completed"] and rep["mismatches"] == 0
        assert 0 < rep["p50_ms"] <= rep["p99_ms"] <= rep["max_ms"]
        # the server stays usable and exact after the load
        N, W, q = cat[0]
        a, b = O.operands("mixed", N, W, q, 9)
        c, _ = S.wait(S.submit(0, a, b))
        assert np.array_equal(c, O.OraclePlan(N, W, q).multiply(a, b))


def test_serve_saturate_closed_loop(O, pm):
    """Closed-loop load keeps `depth` requests in flight and completes them all."""
    cat = _catalog(O, MIX_POINTS[:3])
    with pm.Server(cat) as S:
        rep = S.saturate(16, 0.5, [0.70, 0.25, 0.05], seed=3)
        print("saturate", {k: rep[k] for k in ("requests", "completed", "achieved_rate", "mean_batch")}, flush=True)
        assert rep["completed"] == rep["requests"] > 16 and rep["failed"] == 0
        assert rep["verified"] == rep["completed"] and rep["mismatches"] == 0
        assert rep["mean_batch"] > 1.0


def test_serve_drive_expected_products_match_oracle(O, pm):
    """The load driver's reference for each result is the plain path (pm_run) at the
    serving points; pin that path against the oracle at those points, on
    operands uniform in [0, q) as the load driver draws them."""
    for N, W in MIX_POINTS[:3]:
        q = O.spec_q(N, W)
        L = O.limbs_of(W)
        a, b = O.operands("uniform", N, W, q, 4242 + N)
        with pm.Multiplier(N, W, O.q_limbs(q, L)) as m:
            got = m.run(a, b)
        assert np.array_equal(got, O.OraclePlan(N, W, q).multiply(a, b)), (N, W)


def test_serve_concurrent_submitters(O, pm):
    """Four threads submit and wait at once (the binding releases the GIL in
    every call); every result is exact."""
    import threading
    cat = _catalog(O, [(4096, 256), (8192, 438)])
    with pm.Server(cat) as S:
        errors = []

        def worker(t):
            try:
                for i in range(6):
                    p = (t + i) % 2
                    N, W, q = cat[p]
                    a, b = O.operands("uniform", N, W, q, 1000 + 10 * t + i)
                    c, _ = S.wait(S.submit(p, a, b))
                    if not np.array_equal(c, O.OraclePlan(N, W, q).multiply(a, b)):
                        errors.append((t, i))
            except Exception as e:  # surfaced below
                errors.append(repr(e))

        threads = [threading.Thread(target=worker, args=(t,)) for t in range(4)]
        for th in threads:
            th.start()
        for th in threads:
            th.join(120)
        assert not any(th.is_alive() for th in threads)
        assert not errors, errors
        assert S.stats()["completed"] == 24


@pytest.mark.parametrize("config", [
    dict(max_inflight=1, max_batch=2, host_threads=1, max_batch_bytes=1 << 20),
    dict(max_inflight=6, max_batch=64, host_threads=4, col_tile=8),
])
def test_serve_other_configs(O, pm, config):
    """One slot with tiny batches and one host thread, or six slots and
    narrow column tiles: every request submitted at once still completes,
    each equal to the plain per-call path."""
    cat = _catalog(O, [(8192, 438), (1024, 256), (32768, 868)])
    reqs = []
    for i in range(10):
        p = i % len(cat)
        N, W, q = cat[p]
        reqs.append((p, *O.operands(["mixed", "max", "uniform"][i % 3], N, W, q, 500 + i)))
    print("server config %s" % config, flush=True)
    with pm.Server(cat, **config) as S:
        tickets = [S.submit(p, a, b) for p, a, b in reqs]
        got = [S.wait(t)[0] for t in tickets]
        st = S.stats()
        assert st["completed"] == len(reqs) and st["failed"] == 0
        if config["max_batch"] == 2:
            assert st["batches"] >= len(reqs) // 2
    for i, (p, a, b) in enumerate(reqs):
        N, W, q = cat[p]
        assert np.array_equal(got[i], O.OraclePlan(N, W, q).multiply(a, b)), "request %d" % i


def test_serve_zero_copy_threshold(O, pm):
    """Batches with at most 2 MiB of operands run their kernels on the pinned
    staging (no copies); larger ones go through the copy engines. 32768/256
    (L = 8) has exactly 2 MiB of operands, 32768/288 (L = 9) 2.25 MiB, and
    1024/256 is far below. Each is served alone, then all together (which
    puts the small and the threshold request in one batch above 2 MiB), and
    every result must equal the oracle's."""
    cat = _catalog(O, [(32768, 256), (32768, 288), (1024, 256)])
    plans = [O.OraclePlan(N, W, q) for N, W, q in cat]
    reqs = []
    for i in range(6):
        p = i % len(cat)
        N, W, q = cat[p]
        reqs.append((p, *O.operands("mixed" if i < 3 else "max", N, W, q, 900 + i)))
    with pm.Server(cat) as S:
        for p, a, b in reqs:
            print("zero-copy threshold, alone: point", cat[p][:2], flush=True)
            c, _ = S.wait(S.submit(p, a, b))
            assert np.array_equal(c, plans[p].multiply(a, b)), ("alone", cat[p][:2])
        print("zero-copy threshold, together", flush=True)
        tickets = [S.submit(p, a, b) for p, a, b in reqs]
        for (p, a, b), t in zip(reqs, tickets):
            assert np.array_equal(S.wait(t)[0], plans[p].multiply(a, b)), ("together", cat[p][:2])
