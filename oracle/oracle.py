"""Bigint oracle, operand generators, point-directory I/O, and a ctypes view of
the C restatement (liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker. Only tests/, the smoke check and
bench.py's CPU-baseline leg import this module; the product never does.

What is computed here, and what pins it:
  spec_q          the spec's modulus: the largest prime q < 2^W with
                  q = 1 (mod 2N) (reference tests/oracle.py:48-57 restated).
                  Pinned by delta = 2^W - q values derived independently with
                  sympy (tests/test_oracle.py), and by q = 1 (mod 2N) /
                  primality checks.
  schoolbook      negacyclic product by definition (tests/oracle.py:131-140
                  restated); the ground truth for N <= 256.
  kronecker       an NTT-free bigint product: pack, one big multiply, fold
                  X^N = -1. Independent of every NTT detail.
  all_qm1         closed form for a = b = all (q-1): c_k = (2k + 2 - N) mod q.
  sparse          O(nnz^2) monomial brute force.
"""
from __future__ import annotations

import contextlib
import ctypes
import json
import os
import random
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent

# ----------------------------------------------------------------------------
# primality and the spec modulus
# ----------------------------------------------------------------------------

_SMALL_PRIMES = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47)


def is_probable_prime(n: int, rounds: int = 24) -> bool:
    """Strong probable-prime test: trial division, then Miller-Rabin to the
    first 15 prime bases (deterministic below 3.3e24) plus seeded random
    bases for larger n."""
    if n < 2:
        return False
    for p in _SMALL_PRIMES:
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    rng = random.Random(n & 0xFFFFFFFF)
    bases = list(_SMALL_PRIMES)
    if n.bit_length() > 80:
        bases += [rng.randrange(2, n - 1) for _ in range(rounds)]
    for a in bases:
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


_Q_CACHE: dict[tuple[int, int], int] = {}


def spec_q(N: int, W: int) -> int:
    """The largest prime q < 2^W with q = 1 (mod 2N)."""
    key = (N, W)
    if key in _Q_CACHE:
        return _Q_CACHE[key]
    m = 2 * N
    q = (1 << W) - 1
    q -= (q - 1) % m
    while q > 1:
        if is_probable_prime(q):
            _Q_CACHE[key] = q
            return q
        q -= m
    raise RuntimeError("no prime q for N=%d W=%d" % (N, W))


def limbs_of(W: int) -> int:
    return (W + 31) // 32


# ----------------------------------------------------------------------------
# limb packing (little-endian u32 limbs, coefficient-major [N][L])
# ----------------------------------------------------------------------------


def pack(values, L: int) -> np.ndarray:
    """Python ints -> uint32 array [len][L]."""
    nb = 4 * L
    raw = b"".join(int(x).to_bytes(nb, "little") for x in values)
    return np.frombuffer(raw, dtype="<u4").reshape(len(values), L).copy()


def unpack(arr: np.ndarray) -> list[int]:
    """uint32 array [N][L] -> Python ints."""
    arr = np.ascontiguousarray(arr, dtype="<u4")
    N, L = arr.shape
    raw = arr.tobytes()
    nb = 4 * L
    return [int.from_bytes(raw[i * nb:(i + 1) * nb], "little") for i in range(N)]


def q_limbs(q: int, L: int) -> np.ndarray:
    return pack([q], L).reshape(L)


# ----------------------------------------------------------------------------
# independent products
# ----------------------------------------------------------------------------


def schoolbook(a: list[int], b: list[int], q: int) -> list[int]:
    N = len(a)
    c = [0] * N
    for i in range(N):
        ai = a[i]
        if not ai:
            continue
        for j in range(N):
            t = i + j
            if t < N:
                c[t] += ai * b[j]
            else:
                c[t - N] -= ai * b[j]
    return [v % q for v in c]


def kronecker(a: list[int], b: list[int], q: int) -> list[int]:
    """Negacyclic product via Kronecker substitution: A(2^s) B(2^s) with a slot
    wide enough for any coefficient of the plain product, then X^N = -1."""
    N = len(a)
    s = 2 * q.bit_length() + max(1, N.bit_length()) + 2
    nb = (s + 7) // 8
    s = 8 * nb

    def to_int(v):
        return int.from_bytes(b"".join(int(x).to_bytes(nb, "little") for x in v), "little")

    prod = to_int(a) * to_int(b)
    raw = prod.to_bytes(2 * N * nb, "little")
    coef = [int.from_bytes(raw[i * nb:(i + 1) * nb], "little") for i in range(2 * N)]
    return [(coef[k] - coef[k + N]) % q for k in range(N)]


def all_qm1(N: int, q: int) -> list[int]:
    """a = b = all (q-1) = -(1 + X + ... + X^(N-1)): c_k = (2k + 2 - N) mod q."""
    return [(2 * k + 2 - N) % q for k in range(N)]


def sparse_product(a: dict[int, int], b: dict[int, int], N: int, q: int) -> list[int]:
    """Monomial brute force: aX^i * bX^j = +-ab X^((i+j) mod N)."""
    c = [0] * N
    for i, x in a.items():
        for j, y in b.items():
            t = i + j
            if t < N:
                c[t] += x * y
            else:
                c[t - N] -= x * y
    return [v % q for v in c]


# ----------------------------------------------------------------------------
# seeded operand generators (numpy, fast at N = 65536, L = 55)
# ----------------------------------------------------------------------------


def _less_than_q(x: np.ndarray, ql: np.ndarray) -> np.ndarray:
    """Row-wise x < q for uint32 [n][L] limb arrays."""
    n, L = x.shape
    lt = np.zeros(n, dtype=bool)
    decided = np.zeros(n, dtype=bool)
    for l in range(L - 1, -1, -1):
        xl = x[:, l]
        less = (xl < ql[l]) & ~decided
        more = (xl > ql[l]) & ~decided
        lt |= less
        decided |= less | more
    return lt  # rows equal to q are not < q


def uniform_below_q(rng: np.random.Generator, n: int, W: int, q: int) -> np.ndarray:
    L = limbs_of(W)
    ql = q_limbs(q, L)
    top_bits = W - 32 * (L - 1)
    out = np.empty((n, L), dtype=np.uint32)
    todo = np.arange(n)
    while todo.size:
        x = rng.integers(0, 1 << 32, size=(todo.size, L), dtype=np.uint64).astype(np.uint32)
        if top_bits < 32:
            x[:, L - 1] &= np.uint32((1 << top_bits) - 1)
        ok = _less_than_q(x, ql)
        out[todo[ok]] = x[ok]
        todo = todo[~ok]
    return out


def small_values(rng: np.random.Generator, n: int, L: int, bits: int, q: int) -> np.ndarray:
    """Values below min(2^bits, q), bits <= 32."""
    if bits > 32:
        raise ValueError("small operands are < 2^32")
    out = np.zeros((n, L), dtype=np.uint32)
    hi = min(1 << bits, q)
    if hi > (1 << 32):
        hi = 1 << 32
    v = rng.integers(0, hi, size=n, dtype=np.uint64)
    out[:, 0] = (v & 0xFFFFFFFF).astype(np.uint32)
    return out


def operands(kind: str, N: int, W: int, q: int, seed: int) -> tuple[np.ndarray, np.ndarray]:
    """Seeded operand pair [N][L] for one distribution:
    uniform | max | zero | sparse (1% nonzero) | small16 | mixed
    (per coefficient: 1/2 uniform, 1/4 q-1, 1/4 < 2^32)."""
    L = limbs_of(W)
    rng = np.random.default_rng(seed)
    qm1 = q_limbs(q - 1, L)

    def one():
        if kind == "uniform":
            return uniform_below_q(rng, N, W, q)
        if kind == "max":
            return np.tile(qm1, (N, 1))
        if kind == "zero":
            return np.zeros((N, L), dtype=np.uint32)
        if kind == "sparse":
            x = np.zeros((N, L), dtype=np.uint32)
            nnz = max(1, N // 100)
            idx = rng.choice(N, size=nnz, replace=False)
            x[idx] = uniform_below_q(rng, nnz, W, q)
            return x
        if kind == "small16":
            return small_values(rng, N, L, 16, q)
        if kind == "mixed":
            x = uniform_below_q(rng, N, W, q)
            sel = rng.integers(0, 4, size=N)
            x[sel == 2] = qm1
            sm = small_values(rng, N, L, 32, q)
            x[sel == 3] = sm[sel == 3]
            return x
        raise ValueError("unknown operand kind " + kind)

    return one(), one()


# ----------------------------------------------------------------------------
# point directories, in the layout the reference harness reads (main.cpp:3-15)
# ----------------------------------------------------------------------------


def write_point(root, N: int, W: int, q: int, cases: list[tuple[np.ndarray, np.ndarray]]):
    root = Path(root)
    L = limbs_of(W)
    (root / "point").mkdir(parents=True, exist_ok=True)
    q_limbs(q, L).astype("<u4").tofile(root / "point" / "q.bin")
    for i, (a, b) in enumerate(cases):
        d = root / "cases" / ("%06d" % i)
        d.mkdir(parents=True, exist_ok=True)
        np.ascontiguousarray(a, dtype="<u4").tofile(d / "a.bin")
        np.ascontiguousarray(b, dtype="<u4").tofile(d / "b.bin")
    with open(root / "manifest.json", "w") as f:
        json.dump({"point": {"N": N, "W": W, "L": L}, "N": N, "W": W, "L": L, "cases": len(cases)}, f)


def read_output(root, i: int, N: int, L: int) -> np.ndarray:
    path = Path(root) / "out" / ("%06d" % i) / "c.bin"
    return np.fromfile(path, dtype="<u4").reshape(N, L)


def read_results(root) -> dict:
    with open(Path(root) / "out" / "results.json") as f:
        return json.load(f)


# ----------------------------------------------------------------------------
# the C restatement through ctypes
# ----------------------------------------------------------------------------

_LIB = None


def _bind(path):
    L = ctypes.CDLL(str(path))
    u32p = ctypes.POINTER(ctypes.c_uint32)
    L.pmo_plan_create.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, u32p,
                                  ctypes.POINTER(ctypes.c_void_p), ctypes.c_char_p, ctypes.c_size_t]
    L.pmo_plan_create.restype = ctypes.c_int
    L.pmo_plan_free.argtypes = [ctypes.c_void_p]
    L.pmo_plan_info.argtypes = [ctypes.c_void_p, u32p, ctypes.POINTER(ctypes.c_uint64), u32p, u32p]
    for name in ("pmo_rns", "pmo_crt"):
        getattr(L, name).argtypes = [ctypes.c_void_p, u32p, u32p]
    L.pmo_crt.restype = ctypes.c_int
    for name in ("pmo_ntt_fwd", "pmo_ntt_inv"):
        getattr(L, name).argtypes = [ctypes.c_void_p, u32p]
    L.pmo_pointwise.argtypes = [ctypes.c_void_p, u32p, u32p, u32p]
    L.pmo_multiply.argtypes = [ctypes.c_void_p, u32p, u32p, u32p]
    L.pmo_multiply.restype = ctypes.c_int
    return L


def lib():
    global _LIB
    if _LIB is None:
        path = HERE / "liboracle.so"
        if not path.exists():
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        _LIB = _bind(path)
    return _LIB


@contextlib.contextmanager
def library(path):
    """Plans created inside use another build of pm_oracle.c at `path` (the
    tests' planted-mistake builds); each plan keeps the build it was made with."""
    global _LIB
    saved = lib()
    _LIB = _bind(path)
    try:
        yield
    finally:
        _LIB = saved


def _ptr(a: np.ndarray):
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


class OraclePlan:
    """The C restatement for one point (N, W, L, q)."""

    def __init__(self, N: int, W: int, q: int | np.ndarray, L: int | None = None):
        L = limbs_of(W) if L is None else L
        ql = q if isinstance(q, np.ndarray) else q_limbs(q, L)
        ql = np.ascontiguousarray(ql, dtype=np.uint32)
        self.N, self.W, self.L = N, W, L
        h = ctypes.c_void_p()
        err = ctypes.create_string_buffer(256)
        self._lib = lib()
        rc = self._lib.pmo_plan_create(N, W, L, _ptr(ql), ctypes.byref(h), err, 256)
        if rc != 0:
            raise RuntimeError(err.value.decode())
        self._h = h
        K = ctypes.c_uint32()
        delta = ctypes.c_uint64()
        self._lib.pmo_plan_info(h, ctypes.byref(K), ctypes.byref(delta), None, None)
        self.K, self.delta = K.value, delta.value
        self.primes = np.zeros(self.K, dtype=np.uint32)
        self.psi = np.zeros(self.K, dtype=np.uint32)
        self._lib.pmo_plan_info(h, None, None, _ptr(self.primes), _ptr(self.psi))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None:
            self._lib.pmo_plan_free(h)
            self._h = None

    def rns(self, a: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(a, dtype=np.uint32)
        R = np.zeros((self.K, self.N), dtype=np.uint32)
        self._lib.pmo_rns(self._h, _ptr(a), _ptr(R))
        return R

    def ntt_fwd(self, R: np.ndarray) -> np.ndarray:
        R = np.array(R, dtype=np.uint32, order="C", copy=True)
        self._lib.pmo_ntt_fwd(self._h, _ptr(R))
        return R

    def ntt_inv(self, R: np.ndarray) -> np.ndarray:
        R = np.array(R, dtype=np.uint32, order="C", copy=True)
        self._lib.pmo_ntt_inv(self._h, _ptr(R))
        return R

    def pointwise(self, A: np.ndarray, B: np.ndarray) -> np.ndarray:
        A = np.ascontiguousarray(A, dtype=np.uint32)
        B = np.ascontiguousarray(B, dtype=np.uint32)
        C = np.zeros_like(A)
        self._lib.pmo_pointwise(self._h, _ptr(A), _ptr(B), _ptr(C))
        return C

    def crt(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.uint32)
        c = np.zeros((self.N, self.L), dtype=np.uint32)
        if self._lib.pmo_crt(self._h, _ptr(X), _ptr(c)):
            raise ValueError("CRT margin violated: operands outside [0, q)")
        return c

    def multiply(self, a: np.ndarray, b: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(a, dtype=np.uint32)
        b = np.ascontiguousarray(b, dtype=np.uint32)
        c = np.zeros((self.N, self.L), dtype=np.uint32)
        if self._lib.pmo_multiply(self._h, _ptr(a), _ptr(b), _ptr(c)):
            raise ValueError("CRT margin violated: operands outside [0, q)")
        return c


def ref_binary() -> Path | None:
    """The reference's own solve.cu built on its CPU emulator, if present."""
    p = HERE / "_ref" / "solution_emu"
    return p if p.exists() and os.access(p, os.X_OK) else None
