"""Points past the two-pass NTT and past W = 4064, which the reference accepts
("Any power-of-two N and any W work", README.md:6-7; its only limits are the
primes p = 1 mod 2N below 2^31 and K < 1024, solve.cu:899-913):

* N above 65536 runs three NTT passes (N up to 2^24) or four (2^25, the
  largest N with two such primes);
* W above 4064 (L above 127) runs the RNS decomposition with 64-bit digit
  sums; from about 9,000 bits the CRT splits into its linear part and the
  finish, and from about 12,000 bits the finish reads its slots from global
  memory.

Every result is bit-exact: against the oracle where it finishes in seconds,
else against the all-(q-1) closed form and sparse brute force. q is the
spec's prime where one exists, else q = 2^W - delta (any odd modulus with a
small delta: the product mod q needs no prime, and the reference takes any
such q).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LONG = [(131072, 64, 0), (131072, 256, 0), (262144, 32, 0), (1 << 20, 16, 15)]
WIDE = [(64, 5000), (1024, 6000), (16, 8192), (64, 10000), (16, 15000)]


def _q(O, N, W, delta):
    return O.spec_q(N, W) if delta == 0 else (1 << W) - delta


def _diff(c, want):
    return int(np.count_nonzero((np.asarray(c) != np.asarray(want)).reshape(len(want), -1).any(axis=1)))


def _check_both_paths(O, pm, N, W, q, kinds, seed):
    L = O.limbs_of(W)
    P = O.OraclePlan(N, W, q)
    with pm.Multiplier(N, W, O.q_limbs(q, L)) as m:
        print("N=%d W=%d K=%d" % (N, W, m.K), flush=True)
        assert m.K == P.K
        for s, kind in enumerate(kinds):
            a, b = O.operands(kind, N, W, q, seed + s)
            want = P.multiply(a, b)
            got = m.run(a, b)
            assert np.array_equal(got, want), "host path, %s: %d of %d differ" % (kind, _diff(got, want), N)
            da, db = pm.DeviceBuffer.from_array(a), pm.DeviceBuffer.from_array(b)
            dc = pm.DeviceBuffer(N * L * 4)
            m.run_device(da, db, dc)
            m.sync()
            assert np.array_equal(dc.download((N, L)), want), "device path, " + kind


@pytest.mark.parametrize("N,W,delta", LONG)
def test_multiply_three_passes(O, pm, N, W, delta):
    """N = 2^17, 2^18 and 2^20: three NTT passes, against the oracle."""
    _check_both_paths(O, pm, N, W, _q(O, N, W, delta), ["uniform", "max", "mixed"], 17 * N + W)


@pytest.mark.parametrize("N,W,delta", [(131072, 64, 0), (262144, 32, 0)])
def test_stage_ntt_three_passes(O, pm, N, W, delta):
    """The three-pass NTT alone, forward and (unscaled) inverse, against the
    oracle's transform of random residues."""
    q = _q(O, N, W, delta)
    L = O.limbs_of(W)
    P = O.OraclePlan(N, W, q)
    with pm.Multiplier(N, W, O.q_limbs(q, L)) as m:
        p = P.primes.astype(np.uint64)[:, None]
        R = (np.random.default_rng(N).integers(0, 1 << 62, size=(m.K, N), dtype=np.uint64) % p).astype(np.uint32)
        assert np.array_equal(m.stage_ntt_fwd(R), P.ntt_fwd(R))
        want = (P.ntt_inv(R).astype(np.uint64) * np.uint64(N) % p).astype(np.uint32)
        assert np.array_equal(m.stage_ntt_inv(R), want)


@pytest.mark.parametrize("logN,q", [(24, 251), (25, 251)])
def test_multiply_longest_points(O, pm, logN, q):
    """N = 2^24 (three passes of 8 stages) and 2^25 (four passes), W = 8, L = 1:
    all-(q-1) operands give c_k = (2k + 2 - N) mod q, since (q-1)^2 = 1;
    sparse operands give the negacyclic sums of their monomial products."""
    N, W = 1 << logN, 8
    with pm.Multiplier(N, W, O.q_limbs(q, 1)) as m:
        print("N=2^%d K=%d" % (logN, m.K), flush=True)
        am = np.full((N, 1), q - 1, np.uint32)
        k = np.arange(N, dtype=np.int64)
        want = ((2 * k + 2 - N) % q).astype(np.uint32)
        assert np.array_equal(m.run(am, am)[:, 0], want), "all q-1"
        rng = np.random.default_rng(logN)
        ia, ib = rng.choice(N, 24, replace=False), rng.choice(N, 24, replace=False)
        va, vb = rng.integers(1, q, 24), rng.integers(1, q, 24)
        a = np.zeros((N, 1), np.uint32)
        b = np.zeros((N, 1), np.uint32)
        a[ia, 0], b[ib, 0] = va, vb
        acc = {}
        for i, x in zip(ia.tolist(), va.tolist()):
            for j, y in zip(ib.tolist(), vb.tolist()):
                e, sgn = (i + j, 1) if i + j < N else (i + j - N, -1)
                acc[e] = acc.get(e, 0) + sgn * x * y
        want = np.zeros(N, np.uint32)
        for e, v in acc.items():
            want[e] = v % q
        got = m.run(a, b)[:, 0]
        assert np.array_equal(got, want), "sparse: %d differ" % int(np.count_nonzero(got != want))


@pytest.mark.parametrize("N,W", WIDE)
def test_multiply_wide(O, pm, N, W):
    """W from 5000 to 15000 bits: the 64-bit RNS epilogue, and from 10000 bits
    the split CRT (at 15000 its finish reads the slots from global memory)."""
    _check_both_paths(O, pm, N, W, (1 << W) - 1000001, ["max", "uniform", "mixed", "zero"], W)


@pytest.mark.parametrize("N,W", [(16, 8192), (16, 15000)])
def test_stage_rns_crt_wide(O, pm, N, W):
    """The RNS and CRT stages alone at wide points, against the oracle's."""
    q = (1 << W) - 1000001
    L = O.limbs_of(W)
    P = O.OraclePlan(N, W, q)
    with pm.Multiplier(N, W, O.q_limbs(q, L)) as m:
        for kind in ("uniform", "max"):
            a, b = O.operands(kind, N, W, q, 3)
            assert np.array_equal(m.stage_rns(a), P.rns(a)), "rns " + kind
            X = P.ntt_inv(P.pointwise(P.ntt_fwd(P.rns(a)), P.ntt_fwd(P.rns(b))))
            assert np.array_equal(m.stage_crt(X), P.crt(X)), "crt " + kind


def test_wide_point_sharded_and_lanes(O, pm):
    """The split CRT through the other device paths: calls in flight on the
    lanes (each lane its own slot buffer) and the prime-sharded multiply."""
    N, W = 64, 10000
    q = (1 << W) - 1000001
    L = O.limbs_of(W)
    P = O.OraclePlan(N, W, q)
    ops = [O.operands(k, N, W, q, 70 + i) for i, k in enumerate(["uniform", "mixed", "max", "small16"])]
    with pm.Multiplier(N, W, O.q_limbs(q, L)) as m:
        da = [pm.DeviceBuffer.from_array(a) for a, _ in ops]
        db = [pm.DeviceBuffer.from_array(b) for _, b in ops]
        dc = [pm.DeviceBuffer(N * L * 4) for _ in ops]
        for rep in range(2):
            for i in range(len(ops)):
                m.run_device(da[i], db[i], dc[i])
        m.sync()
        for i, (a, b) in enumerate(ops):
            assert np.array_equal(dc[i].download((N, L)), P.multiply(a, b)), i
        ds = pm.DeviceBuffer(N * L * 4)
        for shards in (1, 3):
            m.run_device_sharded(da[0], db[0], ds, shards)
            m.sync()
            assert np.array_equal(ds.download((N, L)), dc[0].download((N, L))), shards


def test_batch_refuses_wide_and_long(O, pm):
    """The batched and served forms share primes 1 mod 2^17 and run two NTT
    passes with the 32-bit RNS epilogue: they refuse N above 65536 and W above
    4064 with a message (the plain call takes both)."""
    q = (1 << 4065) - 1000001
    with pytest.raises(pm.PolymultError, match="W above 4064"):
        pm.Batch([(256, 4065, q)])
    with pytest.raises(pm.PolymultError, match="N above 65536"):
        pm.Batch([(131072, 64, O.spec_q(131072, 64))])
    with pytest.raises(pm.PolymultError, match="W above 4064"):
        pm.Server([(4096, 256, O.spec_q(4096, 256)), (256, 4065, q)])
    with pytest.raises(pm.PolymultError, match="N above 65536"):
        pm.Server([(131072, 64, O.spec_q(131072, 64))])
