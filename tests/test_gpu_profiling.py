"""Profiling parity: the diagnostics that stand in for the reference's PM_DIAG
setup benchmark and timeline (solve.cu:1238-1358), through the C ABI.

They report rates and timelines, so the checks are on shape and order: every
rate positive and below what its link can carry, the host pipeline's GPU
events in pipeline order, the server's trace files complete and ordered, and
every product computed with the timeline or the trace on still exact.
"""
import csv
import re

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_probe_copy(pm):
    # 512 MiB read and written: larger than the 256 MB last-level cache, so HBM
    # (8 TB/s read + write at most) sets the rate
    g = pm.probe_copy(512 << 20, 5)
    print("device copy %.0f GB/s" % g, flush=True)
    assert 500 < g < 8000
    for nbytes, iters in ((0, 5), (1000, 5), (1 << 20, 0)):
        with pytest.raises(pm.PolymultError, match="bad copy probe arguments"):
            pm.probe_copy(nbytes, iters)


def test_probe_transfers(pm):
    r = pm.probe_transfers(16 << 20, 4, 2)
    print(r, flush=True)
    for k, v in r.items():
        assert v > 0, k
    # PCIe Gen5 x16 carries 64 GB/s per direction; both directions at once at most twice that
    for k in ("h2d_pinned", "d2h_pinned", "h2d_pageable", "d2h_pageable"):
        assert r[k] < 70, k
    assert r["duplex_pinned"] < 140


def test_pm_profile_timeline(O, pm, monkeypatch, capfd):
    monkeypatch.setenv("PM_PROFILE", "1")
    N, W = 8192, 438
    q = O.spec_q(N, W)
    L = O.limbs_of(W)
    a, b = O.operands("uniform", N, W, q, 31)
    with pm.Multiplier(N, W, O.q_limbs(q, L)) as m:
        got = [m.run(a, b) for _ in range(2)]
    err = capfd.readouterr().err
    print(err, flush=True)
    want = O.OraclePlan(N, W, q).multiply(a, b)
    assert all(np.array_equal(g, want) for g in got)
    assert re.search(r"pm: N=8192 W=438 L=14 K=\d+, w1=6 w2=7, chunks in a \d+ / b \d+ / out \d+, "
                     r"column windows in 0 / out 0, \d+ host threads", err)
    host = re.findall(r"pm: inputs staged\+submitted ([\d.]+) us \| first out chunk ([\d.]+) \| done ([\d.]+)", err)
    gpu = re.findall(r"pm: gpu: last H2D ([\d.]+) us \| last RNS ([\d.]+) \| NTT passes ([\d.]+) \| first CRT ([\d.]+) "
                     r"\| first D2H ([\d.]+) \| last D2H ([\d.]+)", err)
    assert len(host) == 2 and len(gpu) == 2
    for h, g in zip(host, gpu):
        h, g = [float(x) for x in h], [float(x) for x in g]
        assert h[0] <= h[2] and h[1] <= h[2]
        # From the last RNS on, each event is recorded on the stream of the one
        # before, or on a stream that waited for it. The last-H2D event is
        # recorded on the input stream after it waits for the second copy
        # stream: an upper bound on the last copy, unordered with the last RNS.
        assert all(x >= 0 for x in g), g
        assert all(x <= y for x, y in zip(g[1:], g[2:])), g
        assert g[0] <= g[5], g


def test_serve_trace(O, pm, monkeypatch, tmp_path):
    base = tmp_path / "serve"
    monkeypatch.setenv("PM_SERVE_TRACE", str(base))
    cat = [(256, 300, O.spec_q(256, 300)), (1024, 100, O.spec_q(1024, 100))]
    reqs = []
    with pm.Server(cat) as S:
        for i in range(12):
            pt = i % 2
            N, W, q = cat[pt]
            a, b = O.operands("uniform", N, W, q, 40 + i)
            reqs.append((pt, a, b, S.submit(pt, a, b)))
        for pt, a, b, t in reqs:
            c, _ = S.wait(t)
            N, W, q = cat[pt]
            assert np.array_equal(c, O.OraclePlan(N, W, q).multiply(a, b))
    # the server writes its trace when it shuts down
    req = list(csv.DictReader(open(str(base) + ".requests.csv")))
    bat = list(csv.DictReader(open(str(base) + ".batches.csv")))
    print("%d requests in %d batches" % (len(req), len(bat)), flush=True)
    assert len(req) == 12
    assert sorted(int(r["point"]) for r in req) == [0] * 6 + [1] * 6
    assert sum(int(x["nreq"]) for x in bat) == 12
    io = {0: 256 * 10 * 4, 1: 1024 * 4 * 4}  # N L 4 bytes of one operand
    assert sum(int(x["bytes"]) for x in bat) == 3 * sum(io[int(r["point"])] for r in req)
    for r in req:
        assert float(r["arrive"]) <= float(r["take"]) <= float(r["done"])
    for x in bat:
        t = [float(x[k]) for k in ("take", "staged", "issued", "seen", "done")]
        assert all(u <= v for u, v in zip(t, t[1:])), t
