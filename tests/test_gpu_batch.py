"""Batched form: many products of uneven N/W in one persistent launch.

Every product of a batch must come out bit-exact against the oracle (which
picks its own primes; the result c is canonical, so the shared prime set of
the batch must not change a single bit), and equal to the plain per-call path.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _batch(O, pm, points, seed, kind="uniform"):
    prods, ops = [], []
    for i, (N, W) in enumerate(points):
        q = O.spec_q(N, W)
        prods.append((N, W, q))
        ops.append(O.operands(kind, N, W, q, seed + i))
    B = pm.Batch(prods)
    da = [pm.DeviceBuffer.from_array(a) for a, _ in ops]
    db = [pm.DeviceBuffer.from_array(b) for _, b in ops]
    dc = [pm.DeviceBuffer(N * O.limbs_of(W) * 4) for N, W in points]
    return B, prods, ops, da, db, dc


def _check(O, prods, ops, dc, idx=None):
    for i in (range(len(prods)) if idx is None else idx):
        N, W, q = prods[i]
        got = dc[i].download((N, O.limbs_of(W)))
        want = O.OraclePlan(N, W, q).multiply(*ops[i])
        assert np.array_equal(got, want), "product %d (N=%d W=%d)" % (i, N, W)


def test_batch_uneven_small(O, pm):
    points = [(64, 32), (1024, 512), (4096, 256), (2048, 868), (256, 1747), (512, 4000), (1, 8), (16, 33),
              (8192, 438), (2, 40)]
    print("batch of %d uneven products" % len(points), flush=True)
    B, prods, ops, da, db, dc = _batch(O, pm, points, 100, "mixed")
    B.run(da, db, dc)
    B.sync()
    _check(O, prods, ops, dc)
    # same result on a second launch (graph replay) and against the plain path
    for d in dc:
        d.upload(np.zeros(d.nbytes // 4, np.uint32))
    B.run(da, db, dc)
    B.sync()
    _check(O, prods, ops, dc)
    for i in (1, 3, 8):
        N, W, q = prods[i]
        with pm.Multiplier(N, W, O.q_limbs(q, O.limbs_of(W))) as m:
            assert np.array_equal(m.run(*ops[i]), dc[i].download((N, O.limbs_of(W))))


def test_batch_log_uniform_mix(O, pm):
    """The measured batch shape: N log-uniform in 2^11..2^16, W uniform in [128, 2048]."""
    rng = np.random.default_rng(4)
    points = [(int(2 ** rng.integers(11, 17)), int(rng.integers(128, 2049))) for _ in range(20)]
    print("batch", points, flush=True)
    B, prods, ops, da, db, dc = _batch(O, pm, points, 200)
    B.run(da, db, dc)
    B.sync()
    _check(O, prods, ops, dc)
    best = B.autotune(da, db, dc, 2)
    print("autotune", best, flush=True)
    assert best["us_per_launch"] > 0
    for d in dc:
        d.upload(np.zeros(d.nbytes // 4, np.uint32))
    B.run(da, db, dc)
    B.sync()
    _check(O, prods, ops, dc)


def test_batch_64_product_workload(O, pm):
    """The batched workload as bench.py times it: 64 products, N log-uniform in
    2^11..2^16, W uniform in [128, 2048], operands uniform in [0, q), one graph
    launch after autotuning. Every one of the 64 products is compared with the
    oracle (computed on host threads: the oracle's ctypes calls release the GIL)."""
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(20260)
    points = [(int(2 ** rng.integers(11, 17)), int(rng.integers(128, 2049))) for _ in range(64)]
    print("batch of 64:", points, flush=True)
    B, prods, ops, da, db, dc = _batch(O, pm, points, 20260 + 200)
    best = B.autotune(da, db, dc, 2)
    print("autotune", best, flush=True)
    for d in dc:
        d.upload(np.zeros(d.nbytes // 4, np.uint32))
    B.run(da, db, dc)
    B.sync()

    def want(i):
        N, W, q = prods[i]
        return O.OraclePlan(N, W, q).multiply(*ops[i])

    with ThreadPoolExecutor(max_workers=8) as ex:
        wants = list(ex.map(want, range(64)))
    bad = [i for i in range(64) if not np.array_equal(dc[i].download((prods[i][0], O.limbs_of(prods[i][1]))),
                                                       wants[i])]
    print("64 products checked, %d differ" % len(bad), flush=True)
    assert not bad, "products differing from the oracle: %s" % bad


def test_batch_closed_form_full_size(O, pm):
    """Eight N=65536/W=1747 products of all-(q-1) operands: c_k = (2k + 2 - N) mod q each."""
    N, W = 65536, 1747
    q = O.spec_q(N, W)
    L = O.limbs_of(W)
    am = np.tile(O.q_limbs(q - 1, L), (N, 1))
    B = pm.Batch([(N, W, q)] * 8)
    da = [pm.DeviceBuffer.from_array(am) for _ in range(8)]
    dc = [pm.DeviceBuffer(N * L * 4) for _ in range(8)]
    B.run(da, da, dc)
    B.sync()
    want = O.all_qm1(N, q)
    for i in range(8):
        assert O.unpack(dc[i].download((N, L))) == want, i


def test_batch_far_moduli(O, pm):
    """Products whose q = 2^W - delta with delta at the largest the point check
    takes (below 2^63 and below 2^(W-1)) beside a spec-prime product, in one
    batch: each equals the oracle's."""
    far = (1 << 63) - 1
    prods = [(256, 64, (1 << 64) - far), (1024, 100, (1 << 100) - far), (16, 33, (1 << 33) - ((1 << 32) - 1)),
             (4096, 256, O.spec_q(4096, 256)), (64, 4064, (1 << 4064) - far), (65536, 1747, (1 << 1747) - far)]
    ops = [O.operands(k, N, W, q, 6400 + i)
           for i, (k, (N, W, q)) in enumerate(zip(["max", "uniform", "mixed", "uniform", "mixed", "mixed"], prods))]
    B = pm.Batch(prods)
    da = [pm.DeviceBuffer.from_array(a) for a, _ in ops]
    db = [pm.DeviceBuffer.from_array(b) for _, b in ops]
    dc = [pm.DeviceBuffer(N * O.limbs_of(W) * 4) for N, W, _ in prods]
    B.run(da, db, dc)
    B.sync()
    _check(O, prods, ops, dc)


def test_batch_widest_point(O, pm):
    """W = 4064 (L = 127, K = 263 of the shared primes) beside narrow
    products, in one group and as single-product groups: every product
    equals the oracle's, which picks its own primes."""
    qw = (1 << 4064) - 1000001  # any odd modulus with a small delta
    prods = [(256, 4064, qw), (4096, 256, O.spec_q(4096, 256)), (2048, 4064, qw), (64, 32, O.spec_q(64, 32))]
    ops = [O.operands(k, N, W, q, 4100 + i) for i, (k, (N, W, q)) in enumerate(zip(["max", "mixed", "uniform", "max"], prods))]
    B = pm.Batch(prods)
    da = [pm.DeviceBuffer.from_array(a) for a, _ in ops]
    db = [pm.DeviceBuffer.from_array(b) for _, b in ops]
    dc = [pm.DeviceBuffer(N * O.limbs_of(W) * 4) for N, W, _ in prods]
    for cfg in (dict(col_tile=32, ceil_split=0, blocks_per_cu=0, group_mb=0, branches=1),
                dict(col_tile=16, ceil_split=1, blocks_per_cu=2, group_mb=1, branches=2)):
        print("widest point, %s" % cfg, flush=True)
        B.configure(**cfg)
        B.run(da, db, dc)
        B.sync()
        _check(O, prods, ops, dc)


@pytest.mark.parametrize("cfg", [
    dict(col_tile=8, ceil_split=1, blocks_per_cu=2, group_mb=0, branches=1),
    dict(col_tile=16, ceil_split=0, blocks_per_cu=0, group_mb=1, branches=2),
    dict(col_tile=32, ceil_split=1, blocks_per_cu=2, group_mb=4, branches=2),
    dict(col_tile=32, ceil_split=0, blocks_per_cu=0, group_mb=0, branches=2),
    dict(col_tile=32, ceil_split=0, blocks_per_cu=2, group_mb=1, branches=3),
])
def test_batch_configurations(O, pm, cfg):
    """Every hand-out mode, grouping (including groups of one product, which
    run the single-product kernels) and graph branching gives the same bits."""
    points = [(4096, 256), (2048, 868), (8192, 438), (1024, 1747), (32768, 300), (16, 64), (2048, 2048)]
    print("config", cfg, flush=True)
    B, prods, ops, da, db, dc = _batch(O, pm, points, 500, "mixed")
    B.configure(**cfg)
    for _ in range(2):
        for d in dc:
            d.upload(np.zeros(d.nbytes // 4, np.uint32))
        B.run(da, db, dc)
        B.sync()
        _check(O, prods, ops, dc)


def test_batch_new_buffers_each_launch(O, pm):
    """A launch with other operand and result buffers computes on those (the
    graph records single-product groups with their buffers as arguments)."""
    points = [(65536, 1747), (4096, 256), (2048, 868)]
    B, prods, ops, da, db, dc = _batch(O, pm, points, 600, "uniform")
    B.configure(col_tile=32, ceil_split=0, blocks_per_cu=0, group_mb=96, branches=2)
    B.run(da, db, dc)
    B.sync()
    _check(O, prods, ops, dc)
    ops2 = [O.operands("mixed", N, W, q, 700 + i) for i, (N, W, q) in enumerate(prods)]
    da2 = [pm.DeviceBuffer.from_array(a) for a, _ in ops2]
    db2 = [pm.DeviceBuffer.from_array(b) for _, b in ops2]
    dc2 = [pm.DeviceBuffer(N * O.limbs_of(W) * 4) for N, W in points]
    B.run(da2, db2, dc2)
    B.sync()
    _check(O, prods, ops2, dc2, idx=[1, 2])
    N, W, q = prods[0]
    with pm.Multiplier(N, W, O.q_limbs(q, O.limbs_of(W))) as m:
        assert np.array_equal(dc2[0].download((N, O.limbs_of(W))), m.run(*ops2[0]))
    # and the first buffers were left alone by the second launch
    _check(O, prods, ops, dc, idx=[1, 2])


@pytest.mark.parametrize("group_mb", [96, 0])
def test_batch_back_to_back_on_caller_stream(O, pm, group_mb):
    """Launches on a caller's stream, alternating two buffer sets with no host
    wait between them: each launch computes on its own buffers (the product
    table is rewritten only after the launch before has finished with it).
    group_mb 96: single-product groups (recaptured); 0: one group whose phase
    kernels read the table."""
    points = [(4096, 256), (2048, 868), (8192, 438), (512, 1747)]
    print("back-to-back launches, group_mb %d" % group_mb, flush=True)
    B, prods, ops, da, db, dc = _batch(O, pm, points, 800, "uniform")
    B.configure(col_tile=32, ceil_split=0, blocks_per_cu=0, group_mb=group_mb, branches=2)
    ops2 = [O.operands("mixed", N, W, q, 900 + i) for i, (N, W, q) in enumerate(prods)]
    da2 = [pm.DeviceBuffer.from_array(a) for a, _ in ops2]
    db2 = [pm.DeviceBuffer.from_array(b) for _, b in ops2]
    dc2 = [pm.DeviceBuffer(N * O.limbs_of(W) * 4) for N, W in points]
    s = pm.Stream()
    for _ in range(3):
        B.run(da, db, dc, stream=s.handle)
        B.run(da2, db2, dc2, stream=s.handle)
    s.sync()
    _check(O, prods, ops, dc)
    _check(O, prods, ops2, dc2)
    s.close()
