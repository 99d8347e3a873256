"""Pins the CPU oracle (oracle/) before anything is compared against it.

Each check is independent of the formula it tests: literal records of the reference
(host/sv/runtime.py:208), brute-force index enumeration, dense Kronecker-embedded operators,
np.transpose versus an index loop, textbook gate identities, closed-form fidelities.
"""
import math

import numpy as np
import pytest

from oracle import dense, judge, model
from oracle.isa import MAGIC, OP_MUX2


def _rand_state(rng, n, dtype=np.complex128):
    s = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return (s / np.linalg.norm(s)).astype(dtype)


def _rand_u(rng):
    z = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def _op_words(t, cmask, M0, M1):
    fl = np.zeros(16, dtype=np.float32)
    for i, M in enumerate((M0, M1)):
        f = np.asarray(M, dtype=np.complex64).reshape(-1)
        fl[8 * i:8 * i + 8:2] = f.real
        fl[8 * i + 1:8 * i + 8:2] = f.imag
    return [OP_MUX2, t, cmask, 0, 0, 0, 0, 0] + [int(x) for x in fl.view(np.int32)]


def _segment(L, k, passes):
    """passes: [(m, hi, [(t, cmask, M0, M1)])] -> int32 words (QSV2)."""
    w = [MAGIC, L, len(passes), k, 0, 0, 0, 0]
    for m, hi, ops in passes:
        h = list(hi) + [0] * (4 - len(hi))
        w += [m, h[0], h[1], len(ops), h[2], h[3], 0, 0]
        for op in ops:
            w += _op_words(*op)
    return (np.asarray(w, dtype=np.int64) & 0xFFFFFFFF).astype(np.uint32).view(np.int32)


def test_zero_segment_literal():
    seg = model.zero_segment(L=6, k=4)
    L, k, passes = model.decode(seg)
    assert (L, k) == (6, 4) and len(passes) == 1
    m, hi, ops = passes[0]
    assert m == 0 and hi == [] and len(ops) == 1
    code, t, cm, M0, M1 = ops[0]
    assert (code, t, cm) == (OP_MUX2, 0, 0) and not M0.any() and not M1.any()
    s = _rand_state(np.random.default_rng(0), 6, np.complex64)
    model.run_segment(s, 1, seg)
    assert not s.any()


def test_bad_magic_asserts():
    seg = model.zero_segment(4, 3).copy()
    seg[0] = 0
    with pytest.raises(AssertionError):
        model.decode(seg)


@pytest.mark.parametrize("L", [3, 4, 5, 6, 7])
def test_tile_index_map_bruteforce(L):
    rng = np.random.default_rng(L)
    for k in range(1, L + 1):
        for m in range(0, min(4, k) + 1):
            cand = list(range(k - m, L))
            hi = sorted(rng.choice(cand, size=m, replace=False).tolist()) if m else []
            idx = model.tile_index_map(L, k, m, hi)
            assert idx.shape == (1 << (L - k), 1 << k)
            assert np.array_equal(np.sort(idx.ravel()), np.arange(1 << L))
            free = [b for b in range(L) if b not in set(range(k - m)) | set(hi)]
            for t in range(1 << (L - k)):
                base = sum(((t >> i) & 1) << b for i, b in enumerate(free))
                for j in range(1 << k):
                    off = (j & ((1 << (k - m)) - 1)) + sum(((j >> (k - m + i)) & 1) << h for i, h in enumerate(hi))
                    assert idx[t, j] == base + off


def _physical_ops(L, k, m, hi, ops):
    """Tile coordinates -> physical bits (tile qubit j < k-m is j, k-m+i is hi[i])."""
    tmap = list(range(k - m)) + list(hi)
    out = []
    for t, cm, M0, M1 in ops:
        pc = 0
        for b in range(k):
            if (cm >> b) & 1:
                pc |= 1 << tmap[b]
        out.append((tmap[t], pc, M0, M1))
    return out


@pytest.mark.parametrize("seed", range(6))
def test_run_segment_vs_dense_operator(seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(3, 7))
    P = 1 << int(rng.integers(0, min(2, n - 2) + 1))
    L = n - (P.bit_length() - 1)
    k = int(rng.integers(2, L + 1))
    passes = []
    for _ in range(3):
        m = int(rng.integers(0, min(3, k) + 1))
        hi = sorted(rng.choice(list(range(k - m, L)), size=m, replace=False).tolist()) if m else []
        ops = []
        for _ in range(int(rng.integers(1, 6))):
            t = int(rng.integers(0, k))
            others = [b for b in range(k) if b != t]
            cm = 0
            for b in others:
                if rng.random() < 0.3:
                    cm |= 1 << b
            M0 = _rand_u(rng) if rng.random() < 0.7 else np.eye(2)
            M1 = _rand_u(rng)
            ops.append((t, cm, M0.astype(np.complex64), M1.astype(np.complex64)))
        passes.append((m, hi, ops))
    seg = _segment(L, k, passes)
    s0 = _rand_state(rng, n)
    got = model.run_segment(s0.copy(), P, seg, np.complex128)
    ref = s0.copy()
    for m, hi, ops in passes:
        for t, pc, M0, M1 in _physical_ops(L, k, m, hi, ops):
            O = dense.mux2_operator(L, t, pc, M0.astype(np.complex128), M1.astype(np.complex128))
            for p in range(P):
                ref[p << L:(p + 1) << L] = O @ ref[p << L:(p + 1) << L]
    assert np.max(np.abs(got - ref)) < 1e-12


@pytest.mark.parametrize("n", [1, 2, 5, 8])
def test_remap_vs_index_loop(n):
    rng = np.random.default_rng(n)
    s = _rand_state(rng, n, np.complex64)
    for _ in range(5):
        perm = [int(x) for x in rng.permutation(n)]
        assert np.array_equal(model.remap(s, n, perm), model.remap_bruteforce(s, n, perm))
    # composition: remap by p then q == remap by (q o p) in perm[new] = old form
    p = [int(x) for x in rng.permutation(n)]
    q = [int(x) for x in rng.permutation(n)]
    comp = [p[q[b]] for b in range(n)]
    assert np.array_equal(model.remap(model.remap(s, n, p), n, q), model.remap(s, n, comp))


def test_unpermute_identity_and_inverse():
    rng = np.random.default_rng(7)
    n = 6
    s = _rand_state(rng, n)
    assert np.array_equal(model.unpermute(s, n, list(range(n))), s)
    fmap = [int(x) for x in rng.permutation(n)]
    # state stored with logical l at physical fmap[l]: build it from the logical-order state
    inv = [0] * n
    for l, p in enumerate(fmap):
        inv[p] = l
    stored = model.remap(s, n, inv)
    assert np.array_equal(model.unpermute(stored, n, fmap), s)


def test_judge_closed_forms():
    rng = np.random.default_rng(3)
    s = _rand_state(rng, 5)
    r = judge.metrics(s, np.exp(0.7j) * s)
    assert abs(r["fidelity"] - 1) < 1e-14 and r["one_minus_f"] < 1e-28 and r["max_abs_err"] < 1e-15
    for eps in (1e-3, 0.1, 0.5):
        a = np.array([1, 0], dtype=np.complex128)
        b = np.array([math.cos(eps), math.sin(eps)], dtype=np.complex128)
        m = judge.metrics(a, b)
        assert abs(m["fidelity"] - math.cos(eps) ** 2) < 1e-14
        assert abs(m["one_minus_f"] - math.sin(eps) ** 2) < 1e-12
        assert abs(judge.fidelity(a, b) - math.cos(eps) ** 2) < 1e-14


def test_dense_gate_identities():
    g = dense.gate_matrix
    H, X, Z = g("h", 1), g("x", 1), g("z", 1)
    assert np.allclose(H @ Z @ H, X, atol=1e-14)
    assert np.allclose(g("s", 1) @ g("s", 1), Z, atol=1e-14)
    assert np.allclose(g("t", 1) @ g("t", 1), g("s", 1), atol=1e-14)
    assert np.allclose(g("u", 1, (math.pi / 2, 0, math.pi)), H, atol=1e-14)
    assert np.allclose(g("sx", 1) @ g("sx", 1), X, atol=1e-14)
    th = 0.731
    assert np.allclose(H @ g("rz", 1, (th,)) @ H, g("rx", 1, (th,)), atol=1e-14)
    assert np.allclose(g("u", 1, (th, 0, 0)), g("ry", 1, (th,)), atol=1e-14)
    cx = g("cx", 2)
    assert np.allclose(cx @ cx, np.eye(4), atol=1e-14)
    # CX(0 -> 1): |q1 q0> = |0 1> (index 1) -> |1 1> (index 3)
    assert cx[3, 1] == 1 and cx[1, 3] == 1 and cx[0, 0] == 1 and cx[2, 2] == 1
    swap = g("swap", 2)
    cx10 = swap @ cx @ swap
    assert np.allclose(cx @ cx10 @ cx, swap, atol=1e-14)
    assert np.allclose(g("cp", 2, (th,)), np.diag([1, 1, 1, np.exp(1j * th)]), atol=1e-14)
    assert np.allclose(g("cz", 2), np.diag([1, 1, 1, -1]), atol=1e-14)
    assert np.allclose(g("rzz", 2, (th,)), np.diag(np.exp(-0.5j * th * np.array([1, -1, -1, 1]))), atol=1e-14)


def test_apply_gate_matches_kron():
    rng = np.random.default_rng(11)
    n = 4
    s = _rand_state(rng, n)
    U = np.kron(_rand_u(rng), _rand_u(rng))   # acts on (q_b, q_a) = kron(U_b, U_a) -> qubits (a, b)
    a, b = 1, 3
    got = dense.apply_gate(s, n, U, [a, b])
    # brute force: amplitude index bits a, b form the gate index bit_a + 2 bit_b
    ref = np.zeros_like(s)
    for i in range(1 << n):
        gi = ((i >> a) & 1) | (((i >> b) & 1) << 1)
        for go in range(4):
            j = (i & ~((1 << a) | (1 << b))) | ((go & 1) << a) | (((go >> 1) & 1) << b)
            ref[j] += U[go, gi] * s[i]
    assert np.max(np.abs(got - ref)) < 1e-14


def test_c64_oracle_rounding_is_reference_order():
    """complex64 apply evaluates M[r,0]*x0 + M[r,1]*x1 in complex64 (host/sv/model.py:53-54)."""
    rng = np.random.default_rng(5)
    T = _rand_state(rng, 4, np.complex64).reshape(1, 16)
    M = _rand_u(rng).astype(np.complex64)
    x0 = T.reshape(1, 8, 2, 1)[:, :, 0, :].copy()
    x1 = T.reshape(1, 8, 2, 1)[:, :, 1, :].copy()
    y0 = M[0, 0] * x0 + M[0, 1] * x1
    y1 = M[1, 0] * x0 + M[1, 1] * x1
    model.apply_op_tile(T, OP_MUX2, 0, 0, M, M)
    V = T.reshape(1, 8, 2, 1)
    assert np.array_equal(V[:, :, 0, :], y0) and np.array_equal(V[:, :, 1, :], y1)
