"""Pins the oracle (oracle/pm_oracle.c through oracle/oracle.py) before any
GPU result is compared with it.

Each check is independent of the oracle's own NTT/CRT code, chosen so that a
plausible slip (a dropped term, a wrong sign or twiddle index, a transposed
operand, a missing prime) fails at least one of them:
  * the spec modulus: delta = 2^W - q and K for the measured points (derived
    independently with sympy; K = 114 at 65536/1747 is README.md:12);
  * schoolbook products on the reference's selftest points (tests/oracle.py:231-239);
  * Kronecker-substitution bigint products (no NTT anywhere);
  * the all-(q-1) closed form c_k = (2k + 2 - N) mod q, at full size;
  * sparse monomial brute force;
  * stage checks: residues vs Python bigint %, the NTT vs a textbook O(N^2)
    negacyclic evaluation, the CRT vs a bigint CRT;
  * the reference binary itself (oracle/_ref/solution_emu) on point dirs;
  * the golden fixtures in tests/golden/.
"""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.mark.parametrize("N,W,delta,K", [(4096, 256, 466943, 18), (8192, 438, 311295, 29),
                                         (32768, 868, 11993087, 57), (65536, 1747, 9699327, 114)])
def test_spec_q_and_K(O, N, W, delta, K):
    q = O.spec_q(N, W)
    assert (1 << W) - q == delta
    assert q % (2 * N) == 1
    P = O.OraclePlan(N, W, q)
    assert P.K == K and P.delta == delta
    p = P.primes.astype(object)
    assert all(int(x) % (2 * N) == 1 and int(x) < (1 << 31) for x in p)
    M = 1
    for x in p:
        M *= int(x)
    assert M > 8 * N * q * q


def test_sympy_agrees_on_q(O):
    sympy = pytest.importorskip("sympy")
    for N, W in [(64, 32), (16, 33), (1024, 512), (4096, 256)]:
        q = O.spec_q(N, W)
        assert sympy.isprime(q)
        assert sympy.prevprime(q + 1) == q  # q prime; the spec-derived one is the largest below 2^W with q = 1 mod 2N
        nxt = q + 2 * N
        while nxt < (1 << W):
            assert not sympy.isprime(nxt)
            nxt += 2 * N


@pytest.mark.parametrize("N,W", [(1, 8), (2, 16), (4, 20), (8, 33), (16, 64), (32, 100), (64, 32), (128, 256),
                                 (2, 40), (16, 33), (2, 3)])
def test_schoolbook(O, N, W):
    q = O.spec_q(N, W)
    P = O.OraclePlan(N, W, q)
    for s, kind in enumerate(["uniform", "max", "sparse", "small16", "mixed", "zero"]):
        a, b = O.operands(kind, N, W, q, 5 + s)
        assert O.unpack(P.multiply(a, b)) == O.schoolbook(O.unpack(a), O.unpack(b), q), kind


@pytest.mark.parametrize("N,W", [(1024, 512), (256, 1747), (4096, 100), (2048, 868), (512, 4000), (4096, 256)])
def test_kronecker(O, N, W):
    q = O.spec_q(N, W)
    P = O.OraclePlan(N, W, q)
    a, b = O.operands("mixed", N, W, q, N + W)
    assert O.unpack(P.multiply(a, b)) == O.kronecker(O.unpack(a), O.unpack(b), q)


@pytest.mark.parametrize("N,W", [(4096, 256), (32768, 868), (65536, 1747)])
def test_closed_form_all_qm1(O, N, W):
    q = O.spec_q(N, W)
    L = O.limbs_of(W)
    am = np.tile(O.q_limbs(q - 1, L), (N, 1))
    assert O.unpack(O.OraclePlan(N, W, q).multiply(am, am)) == O.all_qm1(N, q)


def test_sparse_bruteforce(O):
    N, W = 8192, 438
    q = O.spec_q(N, W)
    L = O.limbs_of(W)
    rng = np.random.default_rng(3)
    a = np.zeros((N, L), np.uint32)
    b = np.zeros((N, L), np.uint32)
    ia, ib = rng.choice(N, 40, replace=False), rng.choice(N, 40, replace=False)
    a[ia] = O.uniform_below_q(rng, 40, W, q)
    b[ib] = O.uniform_below_q(rng, 40, W, q)
    da = {int(i): v for i, v in zip(ia, O.unpack(a[ia]))}
    db = {int(i): v for i, v in zip(ib, O.unpack(b[ib]))}
    assert O.unpack(O.OraclePlan(N, W, q).multiply(a, b)) == O.sparse_product(da, db, N, q)


def test_stage_rns_bigint(O):
    N, W = 256, 1747
    q = O.spec_q(N, W)
    P = O.OraclePlan(N, W, q)
    a, _ = O.operands("uniform", N, W, q, 1)
    R = P.rns(a)
    vals = O.unpack(a)
    for j in range(0, P.K, 7):
        p = int(P.primes[j])
        assert [v % p for v in vals] == [int(x) for x in R[j]]


def test_stage_ntt_textbook(O):
    """Forward output at bit-reversed position k is sum_n x_n psi^((2 brv(k) + 1) n);
    the inverse undoes it exactly."""
    for N, W in [(16, 64), (64, 32)]:
        q = O.spec_q(N, W)
        P = O.OraclePlan(N, W, q)
        logn = N.bit_length() - 1
        rng = np.random.default_rng(N)
        R = np.array([rng.integers(0, int(p), N) for p in P.primes], dtype=np.uint32)
        F = P.ntt_fwd(R)
        for j in range(P.K):
            p, psi = int(P.primes[j]), int(P.psi[j])
            for k in range(N):
                br = int(format(k, "0%db" % logn)[::-1], 2) if logn else 0
                e = 2 * br + 1
                want = sum(int(R[j, n]) * pow(psi, e * n, p) for n in range(N)) % p
                assert int(F[j, k]) == want
        assert np.array_equal(P.ntt_inv(F), R)


def test_stage_crt_bigint(O):
    N, W = 64, 1747
    q = O.spec_q(N, W)
    P = O.OraclePlan(N, W, q)
    import random

    rng = random.Random(2)
    M = 1
    for p in P.primes:
        M *= int(p)
    bound = N * q * q  # |x| < N q^2 < M/8 for every negacyclic product
    assert 8 * bound < M
    xs = [rng.randrange(-bound + 1, bound) for _ in range(N - 3)] + [bound - 1, -(bound - 1), 0]
    X = np.array([[x % int(p) for x in xs] for p in P.primes], dtype=np.uint32)
    assert O.unpack(P.crt(X)) == [x % q for x in xs]


def test_reference_binary(O, tmp_path):
    """The reference's own solve.cu (built on its CPU emulator) agrees with the oracle."""
    exe = O.ref_binary()
    if exe is None:
        pytest.skip("oracle/_ref/solution_emu not built")
    for N, W in [(64, 32), (16, 33), (1024, 512), (256, 1747), (512, 4000)]:
        q = O.spec_q(N, W)
        d = tmp_path / ("pt_%d_%d" % (N, W))
        cases = [O.operands(k, N, W, q, N * W + i) for i, k in enumerate(["uniform", "max", "sparse", "small16"])]
        O.write_point(d, N, W, q, cases)
        r = subprocess.run([str(exe), str(d)], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        P = O.OraclePlan(N, W, q)
        for i, (a, b) in enumerate(cases):
            assert np.array_equal(O.read_output(d, i, N, O.limbs_of(W)), P.multiply(a, b)), (N, W, i)


def test_golden_fixtures(O):
    """Committed vectors (tests/golden/make_golden.py) against the oracle."""
    index = json.loads((GOLDEN / "index.json").read_text())
    assert index["cases"]
    for case in index["cases"]:
        N, W = case["N"], case["W"]
        L = O.limbs_of(W)
        q = int(case["q"], 16)
        assert q == O.spec_q(N, W)
        data = np.load(GOLDEN / case["file"], allow_pickle=False)
        a, b, c = data["a"], data["b"], data["c"]
        assert a.shape == (N, L) and c.shape == (N, L)
        assert np.array_equal(O.OraclePlan(N, W, q).multiply(a, b), c), case["file"]


def test_plan_errors_match_reference(O):
    q = O.spec_q(64, 32)
    ql = O.q_limbs(q, 1)
    with pytest.raises(RuntimeError, match="N must be a power of two"):
        O.OraclePlan(48, 32, ql)
    with pytest.raises(RuntimeError, match="inconsistent W / L"):
        O.OraclePlan(64, 32, ql, L=2)
    with pytest.raises(RuntimeError, match="q out of range"):
        O.OraclePlan(64, 32, np.array([q - 1], np.uint32))
    with pytest.raises(RuntimeError, match="too far below"):
        O.OraclePlan(64, 32, np.array([3], np.uint32))


# Planted mistakes (oracle/pm_oracle.c, PMO_MUTATION): each is a slip the checks
# above are chosen to catch. The reference's README claims its tests catch two
# such bugs (too few primes, a wrong inverse-butterfly sign) with no harness to
# show it; here each planted build must fail at least one end-to-end check.
MUTANTS = {
    1: "one prime fewer than the K rule",
    2: "inverse butterfly difference with the wrong sign",
    3: "T[v] term dropped from the CRT",
    4: "N^-1 dropped from the inverse NTT",
    5: "final conditional subtraction of q dropped",
    6: "a's transform squared instead of multiplied by b's",
    7: "fold without the factor delta",
}
ORACLE_C = Path(__file__).resolve().parent.parent / "oracle" / "pm_oracle.c"


def _oracle_build(tmp_path, m):
    out = tmp_path / ("liboracle_m%d.so" % m)
    subprocess.run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-DPMO_MUTATION=%d" % m, "-o", str(out),
                    str(ORACLE_C), "-lm"], check=True)
    return out


def _end_to_end_failures(O):
    """The end-to-end checks of this file (schoolbook, Kronecker, the
    all-(q-1) closed form) that the oracle build in use fails."""
    failed = []

    def check(name, fn):
        try:
            ok = fn()
        except ValueError:  # the CRT's margin check
            ok = False
        if not ok:
            failed.append(name)

    for N, W in [(1, 8), (2, 16), (4, 20), (8, 33), (16, 64), (64, 32), (128, 256)]:
        q = O.spec_q(N, W)

        def school(N=N, W=W, q=q):
            P = O.OraclePlan(N, W, q)
            for s, kind in enumerate(["uniform", "max", "sparse", "small16", "mixed"]):
                a, b = O.operands(kind, N, W, q, 5 + s)
                if O.unpack(P.multiply(a, b)) != O.schoolbook(O.unpack(a), O.unpack(b), q):
                    return False
            return True

        check("schoolbook %d/%d" % (N, W), school)
    N, W = 1024, 512
    q = O.spec_q(N, W)
    a, b = O.operands("mixed", N, W, q, N + W)
    check("kronecker", lambda: O.unpack(O.OraclePlan(N, W, q).multiply(a, b)) == O.kronecker(O.unpack(a), O.unpack(b), q))
    N, W = 4096, 256
    q = O.spec_q(N, W)
    am = np.tile(O.q_limbs(q - 1, O.limbs_of(W)), (N, 1))
    check("all q-1", lambda: O.unpack(O.OraclePlan(N, W, q).multiply(am, am)) == O.all_qm1(N, q))
    return failed


def test_planted_mistakes_are_caught(O, tmp_path):
    with O.library(_oracle_build(tmp_path, 0)):
        assert _end_to_end_failures(O) == []
    for m, what in MUTANTS.items():
        with O.library(_oracle_build(tmp_path, m)):
            failed = _end_to_end_failures(O)
        print("planted mistake %d (%s): caught by %s" % (m, what, ", ".join(failed) or "nothing"))
        assert failed, "planted mistake %d (%s) passed every check" % (m, what)
