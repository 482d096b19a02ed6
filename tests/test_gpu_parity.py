"""Parity of the HIP path with the oracle, stage by stage and end to end.

Every comparison is bit-exact (integer work). Sizes where the C oracle takes
seconds are compared directly; the full configurations are compared directly
too (the C oracle takes ~2 s at N=65536, W=1747) and through the closed form
for all-(q-1) operands and sparse brute force.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# The reference's emulator points (tests/run_emu.sh:16-29) plus small edge cases.
EMU_POINTS = [(64, 32), (16, 33), (2, 40), (1024, 512), (256, 1747), (4096, 100), (8192, 256), (2048, 868),
              (512, 4000)]
EDGE_POINTS = [(1, 8), (2, 16), (4, 20), (8, 33), (16, 64), (32, 100), (128, 256), (2, 3), (32, 31)]
KINDS = ["uniform", "max", "sparse", "small16", "mixed", "zero"]


def _setup(O, pm, N, W):
    q = O.spec_q(N, W)
    L = O.limbs_of(W)
    return q, L, pm.Multiplier(N, W, O.q_limbs(q, L)), O.OraclePlan(N, W, q)


def _diff(c, want):
    return int(np.count_nonzero((np.asarray(c) != np.asarray(want)).reshape(len(want), -1).any(axis=1)))


@pytest.mark.parametrize("N,W", [(64, 32), (1024, 512), (4096, 256), (2048, 868), (512, 4000), (65536, 1747)])
def test_stage_rns(O, pm, N, W):
    q, L, m, P = _setup(O, pm, N, W)
    print("stage rns N=%d W=%d K=%d" % (N, W, m.K), flush=True)
    assert m.K == P.K
    for kind in ("uniform", "max", "mixed"):
        a, _ = O.operands(kind, N, W, q, 11)
        got = m.stage_rns(a)
        want = P.rns(a)
        assert np.array_equal(got, want), "%s: %d residues differ" % (kind, int(np.count_nonzero(got != want)))


@pytest.mark.parametrize("N,W", [(1, 8), (2, 16), (4, 20), (16, 33), (64, 32), (256, 1747), (1024, 512),
                                 (4096, 256), (8192, 256), (65536, 1747)])
def test_stage_ntt(O, pm, N, W):
    q, L, m, P = _setup(O, pm, N, W)
    print("stage ntt N=%d W=%d K=%d" % (N, W, m.K), flush=True)
    rng = np.random.default_rng(N + W)
    p = P.primes.astype(np.uint64)[:, None]
    R = (rng.integers(0, 1 << 62, size=(m.K, N), dtype=np.uint64) % p).astype(np.uint32)
    fwd = m.stage_ntt_fwd(R)
    assert np.array_equal(fwd, P.ntt_fwd(R))
    inv = m.stage_ntt_inv(R)
    want = (P.ntt_inv(R).astype(np.uint64) * np.uint64(N) % p).astype(np.uint32)  # unscaled inverse
    assert np.array_equal(inv, want)


@pytest.mark.parametrize("N,W", [(8192, 256), (65536, 1747)])
def test_stage_ntt_extremes(O, pm, N, W):
    """Residues at the ends of [0, p): the forward butterflies that leave
    u + v in [0, 2p) and u - v + p in [1, 2p) unreduced, and the inverse
    difference v - u + p, meet their largest and smallest values here."""
    q, L, m, P = _setup(O, pm, N, W)
    print("stage ntt extremes N=%d W=%d K=%d" % (N, W, m.K), flush=True)
    p = P.primes.astype(np.uint64)[:, None]
    rng = np.random.default_rng(7 * N + W)
    spikes = np.zeros((m.K, N), np.uint64)
    spikes[:, rng.choice(N, 64, replace=False)] = 1
    cases = {
        "all p-1": np.broadcast_to(p - 1, (m.K, N)),
        "p-1 spikes on zeros": spikes * (p - 1),
        "0, 1, p-2, p-1": np.choose(rng.integers(0, 4, size=(m.K, N)), [np.zeros_like(p), np.ones_like(p), p - 2, p - 1]),
    }
    for name, R64 in cases.items():
        R = np.ascontiguousarray(R64, dtype=np.uint64).astype(np.uint32)
        assert np.array_equal(m.stage_ntt_fwd(R), P.ntt_fwd(R)), "forward, " + name
        want = (P.ntt_inv(R).astype(np.uint64) * np.uint64(N) % p).astype(np.uint32)
        assert np.array_equal(m.stage_ntt_inv(R), want), "inverse, " + name


@pytest.mark.parametrize("N,W", [(1, 8), (64, 32), (1024, 512), (2048, 868), (512, 4000), (65536, 1747)])
def test_stage_crt(O, pm, N, W):
    q, L, m, P = _setup(O, pm, N, W)
    print("stage crt N=%d W=%d" % (N, W), flush=True)
    for kind in ("uniform", "max", "small16"):
        a, b = O.operands(kind, N, W, q, 5)
        X = P.ntt_inv(P.pointwise(P.ntt_fwd(P.rns(a)), P.ntt_fwd(P.rns(b))))
        got = m.stage_crt(X)
        want = P.crt(X)
        assert np.array_equal(got, want), "%s: %d coefficients differ" % (kind, _diff(got, want))


@pytest.mark.parametrize("N,W", EMU_POINTS + EDGE_POINTS)
def test_multiply_points(O, pm, N, W):
    q, L, m, P = _setup(O, pm, N, W)
    print("multiply N=%d W=%d K=%d" % (N, W, m.K), flush=True)
    for s, kind in enumerate(KINDS):
        a, b = O.operands(kind, N, W, q, 1000 + 7919 * s + N + W)
        got = m.run(a, b)
        want = P.multiply(a, b)
        assert np.array_equal(got, want), "%s: %d of %d coefficients differ" % (kind, _diff(got, want), N)


def test_multiply_widest_point(O, pm):
    """W = 4064, L = 127: the widest point. The RNS epilogue's 32-bit digit
    sums are tightest here (4L 128^2 257 is just under 2^31), and W is a
    multiple of 32 (the fold's whole-limb case). q is any odd 4064-bit
    modulus with a small delta; the product mod q needs no prime."""
    N, W = 256, 4064
    q = (1 << W) - 1000001
    L = O.limbs_of(W)
    P = O.OraclePlan(N, W, q)
    with pm.Multiplier(N, W, O.q_limbs(q, L)) as m:
        print("multiply N=%d W=%d K=%d" % (N, W, m.K), flush=True)
        for s, kind in enumerate(["max", "zero", "uniform", "mixed"]):
            a, b = O.operands(kind, N, W, q, 4064 + s)
            got = m.run(a, b)
            want = P.multiply(a, b)
            assert np.array_equal(got, want), "%s: %d of %d coefficients differ" % (kind, _diff(got, want), N)


# q = 2^W - delta with delta at the largest the point check takes (below 2^63
# and below 2^(W-1), solve.cu:883-896): the CRT's 2^W = delta fold then
# carries the most bits per round and runs the most rounds.
FAR_MODULI = [(64, 2, 1), (16, 33, (1 << 32) - 1), (256, 64, (1 << 63) - 1), (1024, 100, (1 << 63) - 1),
              (4096, 1000, (1 << 63) - 1), (64, 4064, (1 << 63) - 1), (8, 5000, (1 << 63) - 1),
              (65536, 1747, (1 << 63) - 1)]


@pytest.mark.parametrize("N,W,delta", FAR_MODULI)
def test_multiply_far_modulus(O, pm, N, W, delta):
    """Moduli far below 2^W, through the host and the device path: all-(q-1)
    operands against the closed form c_k = (2k + 2 - N) mod q (any odd q:
    (q-1)^2 = 1), the rest against the oracle."""
    q = (1 << W) - delta
    L = O.limbs_of(W)
    P = O.OraclePlan(N, W, q)
    with pm.Multiplier(N, W, O.q_limbs(q, L)) as m:
        print("multiply N=%d W=%d K=%d delta bits %d" % (N, W, m.K, delta.bit_length()), flush=True)
        assert m.K == P.K
        am = np.tile(O.q_limbs(q - 1, L), (N, 1))
        want = np.stack([O.q_limbs((2 * k + 2 - N) % q, L) for k in range(N)])
        assert np.array_equal(m.run(am, am), want), "all q-1"
        for s, kind in enumerate(["uniform", "mixed", "sparse"]):
            a, b = O.operands(kind, N, W, q, 6300 + s + W)
            want = P.multiply(a, b)
            got = m.run(a, b)
            assert np.array_equal(got, want), "host path, %s: %d of %d differ" % (kind, _diff(got, want), N)
            da, db = pm.DeviceBuffer.from_array(a), pm.DeviceBuffer.from_array(b)
            dc = pm.DeviceBuffer(N * L * 4)
            m.run_device(da, db, dc)
            m.sync()
            assert np.array_equal(dc.download((N, L)), want), "device path, " + kind


def test_multiply_schoolbook_small(O, pm):
    """Independent of the oracle's NTT: schoolbook negacyclic products."""
    for N, W in [(1, 8), (8, 33), (64, 32), (128, 256)]:
        q, L, m, _ = _setup(O, pm, N, W)
        for seed in range(3):
            a, b = O.operands("uniform", N, W, q, seed)
            assert O.unpack(m.run(a, b)) == O.schoolbook(O.unpack(a), O.unpack(b), q)


def test_multiply_kronecker(O, pm):
    """Independent of the oracle's NTT: Kronecker-substitution bigint product."""
    for N, W in [(4096, 256), (1024, 868)]:
        q, L, m, _ = _setup(O, pm, N, W)
        a, b = O.operands("mixed", N, W, q, 77)
        assert O.unpack(m.run(a, b)) == O.kronecker(O.unpack(a), O.unpack(b), q)


# ---- the measured configurations ----

def test_config_4096_256(O, pm):
    q, L, m, P = _setup(O, pm, 4096, 256)
    assert (m.K, (1 << 256) - q) == (18, 466943)
    for kind in ("uniform", "max", "sparse", "small16", "mixed"):
        a, b = O.operands(kind, 4096, 256, q, 4096)
        assert np.array_equal(m.run(a, b), P.multiply(a, b)), kind


def test_config_32768_868(O, pm):
    N, W = 32768, 868
    q, L, m, P = _setup(O, pm, N, W)
    assert (m.K, (1 << W) - q) == (57, 11993087)
    for kind in ("uniform", "max", "sparse", "small16"):
        a, b = O.operands(kind, N, W, q, 32768)
        got = m.run(a, b)
        assert np.array_equal(got, P.multiply(a, b)), kind
        if kind == "max":
            assert O.unpack(got) == O.all_qm1(N, q)


def test_config_65536_1747(O, pm):
    N, W = 65536, 1747
    q, L, m, P = _setup(O, pm, N, W)
    assert (m.K, (1 << W) - q) == (114, 9699327)
    a, b = O.operands("mixed", N, W, q, 65536)
    got = m.run(a, b)
    assert np.array_equal(got, P.multiply(a, b))
    # closed form, independent of every NTT detail
    am = np.tile(O.q_limbs(q - 1, L), (N, 1))
    assert O.unpack(m.run(am, am)) == O.all_qm1(N, q)
    # commutativity and device-resident path
    assert np.array_equal(m.run(b, a), got)
    da, db = pm.DeviceBuffer.from_array(a), pm.DeviceBuffer.from_array(b)
    dc = pm.DeviceBuffer(N * L * 4)
    m.run_device(da, db, dc)
    m.sync()
    assert np.array_equal(dc.download((N, L)), got)


def test_sparse_bruteforce_65536(O, pm):
    """Sparse operands against O(nnz^2) monomial brute force at full size."""
    N, W = 65536, 1747
    q, L, m, _ = _setup(O, pm, N, W)
    rng = np.random.default_rng(9)
    a = np.zeros((N, L), dtype=np.uint32)
    b = np.zeros((N, L), dtype=np.uint32)
    ia = rng.choice(N, 24, replace=False)
    ib = rng.choice(N, 24, replace=False)
    a[ia] = O.uniform_below_q(rng, 24, W, q)
    b[ib] = O.uniform_below_q(rng, 24, W, q)
    da = {int(i): v for i, v in zip(ia, O.unpack(a[ia]))}
    db = {int(i): v for i, v in zip(ib, O.unpack(b[ib]))}
    assert O.unpack(m.run(a, b)) == O.sparse_product(da, db, N, q)


def test_errors_like_reference(O, pm):
    q = O.spec_q(64, 32)
    with pytest.raises(pm.PolymultError, match="power of two"):
        pm.Multiplier(48, 32, O.q_limbs(q, 1))
    m = pm.Multiplier(64, 32, O.q_limbs(q, 1))
    with pytest.raises(pm.PolymultError, match="operand size"):
        m.run(np.zeros(10, np.uint32), np.zeros(10, np.uint32))


def test_device_calls_in_flight(O, pm):
    """Default-stream device calls overlap on the lanes: each result is exact,
    and calls into one output buffer leave the last call's result."""
    N, W = 32768, 868
    q = O.spec_q(N, W)
    L = O.limbs_of(W)
    ops = [O.operands(k, N, W, q, 900 + i) for i, k in enumerate(["uniform", "mixed", "max", "sparse", "small16"])]
    with pm.Multiplier(N, W, O.q_limbs(q, L)) as m:
        da = [pm.DeviceBuffer.from_array(a) for a, _ in ops]
        db = [pm.DeviceBuffer.from_array(b) for _, b in ops]
        dc = [pm.DeviceBuffer(N * L * 4) for _ in ops]
        for rep in range(3):
            for i in range(len(ops)):
                m.run_device(da[i], db[i], dc[i])
        shared = pm.DeviceBuffer(N * L * 4)
        for i in range(len(ops)):
            m.run_device(da[i], db[i], shared)
        m.sync()
        for i, (a, b) in enumerate(ops):
            assert np.array_equal(dc[i].download((N, L)), O.OraclePlan(N, W, q).multiply(a, b)), i
        assert np.array_equal(shared.download((N, L)), dc[-1].download((N, L)))
        # the host-buffer path on the same state afterwards
        assert np.array_equal(m.run(*ops[1]), dc[1].download((N, L)))


def test_device_calls_chained(O, pm):
    """Device calls that read earlier calls' outputs with no host wait: on the
    default stream (the lanes) x1 = a b, x2 = x1 b, x3 = x2 x1, x4 = x3 a, three
    times over; then a caller-stream result read by a default-stream call and
    the other way round. Each read waits for the write it depends on."""
    N, W = 32768, 868
    q = O.spec_q(N, W)
    L = O.limbs_of(W)
    a, b = O.operands("uniform", N, W, q, 1300)
    P = O.OraclePlan(N, W, q)
    x1 = P.multiply(a, b)
    x2 = P.multiply(x1, b)
    x3 = P.multiply(x2, x1)
    x4 = P.multiply(x3, a)
    s = pm.Stream()
    with pm.Multiplier(N, W, O.q_limbs(q, L)) as m:
        da, db = pm.DeviceBuffer.from_array(a), pm.DeviceBuffer.from_array(b)
        dx = [pm.DeviceBuffer(N * L * 4) for _ in range(6)]
        for rep in range(3):
            for d in dx:
                d.upload(np.zeros((N, L), np.uint32))
            print("chain round", rep, flush=True)
            m.run_device(da, db, dx[0])
            m.run_device(dx[0], db, dx[1])
            m.run_device(dx[1], dx[0], dx[2])
            m.run_device(dx[2], da, dx[3])
            m.run_device(dx[3], dx[3], dx[4], s.handle)  # caller stream after the lanes
            m.run_device(dx[4], da, dx[5])  # lanes after the caller stream
            m.sync()
            s.sync()
            assert np.array_equal(dx[0].download((N, L)), x1)
            assert np.array_equal(dx[1].download((N, L)), x2)
            assert np.array_equal(dx[2].download((N, L)), x3)
            assert np.array_equal(dx[3].download((N, L)), x4)
            x5 = P.multiply(x4, x4)
            assert np.array_equal(dx[4].download((N, L)), x5)
            assert np.array_equal(dx[5].download((N, L)), P.multiply(x5, a))
    s.close()


def test_time_device_leaves_product(O, pm):
    """pm_time_device repeats each kernel on the chain's residues, then runs
    whole multiplies: its times are positive and d_c ends holding a * b."""
    N, W = 4096, 256
    q, L, m, P = _setup(O, pm, N, W)
    a, b = O.operands("mixed", N, W, q, 77)
    with m:
        da, db = pm.DeviceBuffer.from_array(a), pm.DeviceBuffer.from_array(b)
        dc = pm.DeviceBuffer(N * L * 4)
        ms = m.time_device(da, db, dc, 5)
        assert len(ms) == 6 and all(t > 0 for t in ms)
        assert ms[0] >= max(ms[1:])
        assert np.array_equal(dc.download((N, L)), P.multiply(a, b))


def test_device_calls_across_streams(O, pm):
    """One state driven from two caller streams, the default stream (lanes),
    the sharded path and the host path, interleaved with no host wait: every
    result is exact (the state's shared residues are ordered across streams)."""
    N, W = 8192, 438
    q = O.spec_q(N, W)
    L = O.limbs_of(W)
    ops = [O.operands(k, N, W, q, 1200 + i) for i, k in enumerate(["uniform", "mixed", "max", "small16", "sparse"])]
    want = [O.OraclePlan(N, W, q).multiply(a, b) for a, b in ops]
    s1, s2 = pm.Stream(), pm.Stream()
    with pm.Multiplier(N, W, O.q_limbs(q, L)) as m:
        da = [pm.DeviceBuffer.from_array(a) for a, _ in ops]
        db = [pm.DeviceBuffer.from_array(b) for _, b in ops]
        dc = [pm.DeviceBuffer(N * L * 4) for _ in range(8)]
        m.run_device(da[0], db[0], dc[0], s1.handle)
        m.run_device(da[1], db[1], dc[1], s2.handle)
        m.run_device(da[2], db[2], dc[2])
        m.run_device(da[3], db[3], dc[3], s1.handle)
        m.run_device_sharded(da[4], db[4], dc[4], 2, s2.handle)
        c_host = m.run(*ops[1])
        m.run_device(da[2], db[2], dc[5], s2.handle)
        m.run_device(da[0], db[0], dc[6])
        m.run_device(da[3], db[3], dc[7], s1.handle)
        s1.sync()
        s2.sync()
        m.sync()
        assert np.array_equal(c_host, want[1])
        for k, i in enumerate([0, 1, 2, 3, 4, 2, 0, 3]):
            assert np.array_equal(dc[k].download((N, L)), want[i]), k
    s1.close()
    s2.close()


def test_distinct_states_on_threads(O, pm):
    """Four states, each created, used and freed on its own host thread at
    once (the ctypes calls release the GIL): host-buffer and device calls on
    every state are exact, as the header promises for distinct states."""
    import threading

    points = [(4096, 256), (2048, 868), (8192, 100), (1024, 1747)]
    jobs = []
    for i, (N, W) in enumerate(points):
        q = O.spec_q(N, W)
        ops = [O.operands(k, N, W, q, 1300 + 10 * i + j) for j, k in enumerate(["uniform", "mixed"])]
        jobs.append((N, W, q, ops, [O.OraclePlan(N, W, q).multiply(a, b) for a, b in ops]))
    errors = []
    start = threading.Barrier(len(jobs))

    def work(N, W, q, ops, want):
        try:
            L = O.limbs_of(W)
            start.wait(60)
            with pm.Multiplier(N, W, O.q_limbs(q, L)) as m:
                da = [pm.DeviceBuffer.from_array(a) for a, _ in ops]
                db = [pm.DeviceBuffer.from_array(b) for _, b in ops]
                dc = [pm.DeviceBuffer(N * L * 4) for _ in ops]
                for rep in range(4):
                    for j in range(len(ops)):
                        if not np.array_equal(m.run(*ops[j]), want[j]):
                            errors.append("host N=%d W=%d rep %d op %d" % (N, W, rep, j))
                        m.run_device(da[j], db[j], dc[j])
                    m.sync()
                    for j in range(len(ops)):
                        if not np.array_equal(dc[j].download((N, L)), want[j]):
                            errors.append("device N=%d W=%d rep %d op %d" % (N, W, rep, j))
        except Exception as e:  # reported below, on the test's thread
            errors.append("N=%d W=%d: %r" % (N, W, e))

    threads = [threading.Thread(target=work, args=job) for job in jobs]
    for t in threads:
        t.start()
    for t in threads:
        t.join(100)
    assert not any(t.is_alive() for t in threads), "a worker did not finish"
    assert not errors, errors


@pytest.mark.parametrize("N,W", [(16384, 256), (16384, 288), (8192, 438), (1, 8), (2048, 868)])
def test_host_small_point_path(O, pm, N, W):
    """pm_run at points whose chunk plans are single chunks (the whole output
    within its first chunk: N L 4 up to about 0.5 MiB), where the kernels read
    and write the pinned staging buffers directly on one stream. 16384/256 is
    the largest such point at W = 256 (exactly 0.5 MiB); 16384/288 (L = 9) is
    just above it and takes the chunked pipeline. Consecutive calls with new
    operands, and a device-resident call between them on the same state, must
    not see each other's data."""
    q, L, m, P = _setup(O, pm, N, W)
    print("small-point path N=%d W=%d K=%d" % (N, W, m.K), flush=True)
    with m:
        for s, kind in enumerate(["mixed", "max", "uniform"]):
            a, b = O.operands(kind, N, W, q, 1700 + s + N)
            got = m.run(a, b)
            assert np.array_equal(got, P.multiply(a, b)), (N, W, kind)
            if s == 0:
                da, db = pm.DeviceBuffer.from_array(b), pm.DeviceBuffer.from_array(a)
                dc = pm.DeviceBuffer(N * L * 4)
                m.run_device(da, db, dc)
                m.sync()
                assert np.array_equal(dc.download((N, L)), P.multiply(b, a)), (N, W, "device")


@pytest.mark.parametrize("env", [
    {"PM_IN_CHUNKS": "0.05,0.2", "PM_IN_TAIL": "0.03,0.01", "PM_IN_CHUNKS_B": "0.07", "PM_IN_TAIL_B": "0.02",
     "PM_OUT_CHUNKS": "0.1", "PM_OUT_TAIL": "0.02"},
    {"PM_IN_CHUNKS": "500", "PM_IN_TAIL": "400", "PM_IN_TAIL_B": "400", "PM_OUT_CHUNKS": "500", "PM_OUT_TAIL": "400"},
    {"PM_HOST_THREADS": "1", "PM_H2D_STREAMS": "1", "PM_D2H_STREAMS": "1"},
    {"PM_HOST_THREADS": "3", "PM_IN_CHUNKS": "1,0.3", "PM_IN_CHUNKS_B": "0.5,2", "PM_OUT_CHUNKS": "0.7,2"},
])
def test_host_pipeline_plans(O, pm, env, monkeypatch):
    """pm_run (the fherma_run path) under other chunk plans (a's and b's
    apart), host-thread counts and one copy stream each way: dozens of tiny
    chunks, one chunk larger than the operand, tails larger than N. The plans
    are read when a state is created."""
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    print("pipeline plan %s" % env, flush=True)
    for N, W in [(32768, 868), (4096, 256), (1, 8), (64, 32)]:
        q, L, m, P = _setup(O, pm, N, W)
        with m:
            for s, kind in enumerate(["mixed", "max"]):
                a, b = O.operands(kind, N, W, q, 1300 + s + N)
                got = m.run(a, b)
                assert np.array_equal(got, P.multiply(a, b)), (N, W, kind)


@pytest.mark.parametrize("env", [
    {"PM_IN_COLS": "32", "PM_OUT_COLS": "32"},
    {"PM_IN_COLS": "64", "PM_OUT_COLS": "64"},
    {"PM_IN_COLS": "128", "PM_OUT_COLS": "128", "PM_OUT_CHUNKS": "0.3,2"},
    {"PM_IN_COLS": "64"},
    {"PM_OUT_COLS": "128", "PM_HOST_THREADS": "3"},
    {"PM_OUT_COLS": "0"},
])
def test_host_column_windows(O, pm, env, monkeypatch):
    """pm_run with b staged in column windows (each window's RNS and forward
    column pass run as it lands) and the result leaving window by window (the
    first window's inverse column pass and CRT first), against the oracle.
    Windows apply to chunked two-pass plans: 65536/1747 (w1 = w2 = 8),
    32768/868 (w1 = 7, w2 = 8), 16384/288 (128 columns: a 128-column window
    is the whole row and stays off), 16384/5000 (64-bit RNS epilogue; its
    CRT is not the aliased one, so the result stays in contiguous chunks) and
    131072/64 (three passes: no windows). PM_OUT_COLS=0 is the default,
    contiguous chunks."""
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    print("column windows %s" % env, flush=True)
    for N, W, kinds in [(65536, 1747, ["mixed"]), (32768, 868, ["mixed", "max"]), (16384, 288, ["uniform"]),
                        (16384, 5000, ["mixed"]), (131072, 64, ["max"])]:
        q, L, m, P = _setup(O, pm, N, W)
        with m:
            for s, kind in enumerate(kinds):
                a, b = O.operands(kind, N, W, q, 1900 + s + N)
                got = m.run(a, b)
                assert np.array_equal(got, P.multiply(a, b)), (N, W, kind, env)


def test_host_column_windows_engage(O, pm, monkeypatch, capfd):
    """The window settings reach the pipeline: with PM_PROFILE its set-up
    line names the window widths in use at 65536/1747, and the profiled call
    is exact too."""
    monkeypatch.setenv("PM_IN_COLS", "64")
    monkeypatch.setenv("PM_OUT_COLS", "32")
    monkeypatch.setenv("PM_PROFILE", "1")
    N, W = 65536, 1747
    q, L, m, P = _setup(O, pm, N, W)
    with m:
        a, b = O.operands("mixed", N, W, q, 1999)
        got = m.run(a, b)
        assert np.array_equal(got, P.multiply(a, b))
    err = capfd.readouterr().err
    assert "column windows in 64 / out 32" in err, err[-2000:]


def test_host_output_chunks_wait_for_their_crt(O, pm, monkeypatch):
    """Every D2H chunk waits for the CRT launch that writes it. With a
    128-coefficient first chunk its copy ends long before the second CRT
    launch (the rest of the coefficients) does, so a copy stream that waited
    only for the first launch could copy stale results of the previous call;
    each call here has new operands, so stale chunks would not match. (On
    this runtime the two D2H streams' copies went out in issue order even
    without that wait, so the check guards the dependency rather than
    reproducing its absence.)"""
    monkeypatch.setenv("PM_OUT_CHUNKS", "0.02,4")
    monkeypatch.setenv("PM_OUT_TAIL", "0")
    N, W = 65536, 1747
    q, L, m, P = _setup(O, pm, N, W)
    with m:
        for s, kind in enumerate(["mixed", "uniform", "max", "mixed"]):
            a, b = O.operands(kind, N, W, q, 2100 + s)
            got = m.run(a, b)
            bad = _diff(got, P.multiply(a, b))
            assert bad == 0, (kind, s, bad)
