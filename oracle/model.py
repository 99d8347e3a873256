"""Golden model of the QSV2 MUX2 stream: a numpy restatement of host/sv/model.py.

Each function follows the reference function named in its docstring.  ``ctype`` selects the
arithmetic: complex64 reproduces the reference model (the same numpy expression order, so the
same rounding on the same numpy build); complex128 is the error judge (the f32 matrix payload is
widened exactly, so c64-vs-c128 differences measure arithmetic error only).

Pinning (tests/test_oracle.py):
  decode            -- literal zero segment (host/sv/runtime.py:208), encode->decode round trips.
  tile_index_map    -- brute force over every (L, k, m, hi) at small L: every tile is a
                       permutation of the slice, sub-block j*2^(k-m) offsets equal sum_i 2^h_i.
  apply_op_tile /
  run_segment       -- against oracle.dense (an independent gate-level complex128 simulator built
                       from Pauli exponentials) on compiled circuits, and against a dense
                       2^n x 2^n Kronecker-embedded MUX2 operator for random (t, cmask, M0, M1).
  remap / unpermute -- against a per-index bit-permutation loop (brute force) and composition.
The c64 rounding order is the reference's (host/sv/model.py:53-54); bit-identity with the
reference module itself is pinned by the committed vectors in tests/golden/ when present.
"""
import numpy as np

from .isa import MAGIC, HDR_WORDS, PASS_WORDS, OP_WORDS, OP_MUX2, OP_HIGHX, OP_DENSE


def decode(seg, ctype=np.complex64):
    """host/sv/model.py:7-22 -- parse a segment into (L, k, [(m, hi, [(code, t, cmask, M0, M1)])])."""
    w = np.asarray(seg, dtype=np.int32)
    assert int(w[0]) == MAGIC, "bad QSV2 magic"
    L, npass, k = int(w[1]), int(w[2]), int(w[3])
    pos = HDR_WORDS
    passes = []
    for _ in range(npass):
        hdr = [int(v) for v in w[pos:pos + PASS_WORDS]]
        pos += PASS_WORDS
        m, nops = hdr[0], hdr[3]
        highs = [hdr[1], hdr[2], hdr[4], hdr[5]][:m]
        if m > 4:
            # this project's MI355X extension (the reference never emits m > 4): an 8-word
            # continuation record [OP_HIGHX, h4, .., h10] follows the pass header
            ext = [int(v) for v in w[pos:pos + PASS_WORDS]]
            assert ext[0] == OP_HIGHX, "pass with m > 4 lacks its high-qubit record"
            highs = highs + ext[1:1 + m - 4]
            pos += PASS_WORDS
        ops = []
        for _ in range(nops):
            if int(w[pos]) == OP_DENSE:
                # this project's fused form: [OP_DENSE, f, q0..q4, 0] + 2 * 4^f f32 words of U
                f = int(w[pos + 1])
                qs = [int(v) for v in w[pos + 2:pos + 2 + f]]
                fl = np.ascontiguousarray(w[pos + 8:pos + 8 + (2 << (2 * f))]).view(np.float32)
                U = np.empty(1 << (2 * f), dtype=ctype)
                U.real = fl[0::2]
                U.imag = fl[1::2]
                ops.append((OP_DENSE, qs, 0, U.reshape(1 << f, 1 << f), None))
                pos += 8 + (2 << (2 * f))
                continue
            rec = w[pos:pos + OP_WORDS]
            pos += OP_WORDS
            fl = np.ascontiguousarray(rec[8:24]).view(np.float32)
            z = np.empty(8, dtype=ctype)
            z.real = fl[0::2]
            z.imag = fl[1::2]
            ops.append((int(rec[0]), int(rec[1]), int(rec[2]), z[:4].reshape(2, 2), z[4:].reshape(2, 2)))
        passes.append((m, highs, ops))
    return L, k, passes


def _deposit(values, positions):
    """Scatter bit i of each value to bit positions[i] (vectorised pdep)."""
    out = np.zeros_like(values)
    for i, b in enumerate(positions):
        out |= ((values >> i) & 1) << b
    return out


def tile_index_map(L, k, m, hi):
    """host/sv/model.py:25-40 -- idx[tile, j] = slice index of element j of tile `tile`.

    Tile element j: its low k-m bits are physical bits 0..k-m-1, bit k-m+i is physical hi[i].
    Tiles enumerate the remaining ("free") bits of [0, L) in ascending order.
    """
    hi = list(hi)[:m]
    low = k - m
    taken = set(range(low)) | set(hi)
    free = [b for b in range(L) if b not in taken]
    assert len(free) == L - k, "tile window does not partition the slice"
    j = np.arange(1 << k, dtype=np.int64)
    off = (j & ((1 << low) - 1)) | _deposit(j >> low, hi)
    tiles = np.arange(1 << (L - k), dtype=np.int64)
    base = _deposit(tiles, free)
    return base[:, None] | off[None, :]


def apply_op_tile(T, code, t, cmask, M0, M1):
    """host/sv/model.py:43-56 -- MUX2 on tile qubit t of every tile (rows of T), in place.

    Pairs (i0, i0 | 2^t) with bit t of i0 clear use M1 when cmask != 0 and (i0 & cmask) == cmask,
    otherwise M0.  Evaluated as M[r,0]*x0 + M[r,1]*x1 in T's dtype (the reference's order).
    """
    assert code == OP_MUX2, "unknown opcode"
    nt, size = T.shape
    V = T.reshape(nt, size >> (t + 1), 2, 1 << t)
    i0 = (np.arange(size >> (t + 1))[:, None] << (t + 1)) | np.arange(1 << t)[None, :]
    if cmask:
        use1 = (i0 & cmask) == cmask
    else:
        use1 = np.zeros(i0.shape, dtype=bool)
    x0 = V[:, :, 0, :].copy()
    x1 = V[:, :, 1, :].copy()
    for M, where in ((M0, ~use1), (M1, use1)):
        y0 = M[0, 0] * x0 + M[0, 1] * x1
        y1 = M[1, 0] * x0 + M[1, 1] * x1
        V[:, :, 0, :] = np.where(where[None], y0, V[:, :, 0, :])
        V[:, :, 1, :] = np.where(where[None], y1, V[:, :, 1, :])


def apply_dense_tile(T, qs, U):
    """This project's fused record: a dense 2^f x 2^f U on tile qubits qs (index sum_i bit(qs[i]) << i)
    applied to every tile (rows of T), in place, in T's dtype.  Pinned against the product of the
    MUX2 ops it replaces (tests/test_oracle.py)."""
    nt, size = T.shape
    k = size.bit_length() - 1
    f = len(qs)
    V = T.reshape((nt,) + (2,) * k)
    axes = [1 + (k - 1 - qs[f - 1 - i]) for i in range(f)]
    Ut = U.reshape((2,) * (2 * f))
    out = np.tensordot(V, Ut, axes=(axes, list(range(f, 2 * f))))
    # out axes: rest of V (in order), then U's output bits (f-1 .. 0)
    rest = [a for a in range(k + 1) if a not in axes]
    order = rest + axes
    T[:] = np.transpose(out, np.argsort(order)).reshape(nt, size)


def run_segment(state, P, seg, ctype=None):
    """host/sv/model.py:59-71 -- run one segment on each of P contiguous 2^L slices, in place."""
    ctype = state.dtype if ctype is None else ctype
    L, k, passes = decode(seg, ctype)
    assert state.size == P << L, "state size does not match P * 2^L"
    for m, hi, ops in passes:
        idx = tile_index_map(L, k, m, hi)
        for p in range(P):
            sl = state[p << L:(p + 1) << L]
            T = sl[idx]
            for code, t, cm, M0, M1 in ops:
                if code == OP_DENSE:
                    apply_dense_tile(T, t, M0)
                else:
                    apply_op_tile(T, code, t, cm, M0, M1)
            sl[idx] = T
    return state


def remap(state, n, perm):
    """host/sv/model.py:74-78 -- qubit permutation, perm[new_bit] = old_bit, as an axis transpose.

    A [2]*n view has axis a <-> bit n-1-a; new axis a takes old bit perm[n-1-a].
    """
    view = state.reshape((2,) * n)
    axes = [n - 1 - perm[n - 1 - a] for a in range(n)]
    return np.ascontiguousarray(np.transpose(view, axes)).reshape(-1)


def remap_bruteforce(state, n, perm):
    """Independent check of remap: out[j] = in[sum_b bit_b(j) << perm[b]] (index loop, vectorised)."""
    j = np.arange(1 << n, dtype=np.int64)
    src = np.zeros_like(j)
    for b in range(n):
        src |= ((j >> b) & 1) << perm[b]
    return state[src]


def unpermute(state, n, final_map):
    """host/sv/model.py:81-84 -- return the state in logical order (logical l at physical final_map[l])."""
    return remap(state, n, [final_map[l] for l in range(n)])


def zero_state(n, ctype=np.complex64):
    s = np.zeros(1 << n, dtype=ctype)
    s[0] = 1
    return s


def run_program(prog, n, P, final_map, state=None, ctype=np.complex64):
    """host/sv/model.py:87-95 -- execute [("segment", words) | ("remap", perm)] then unpermute."""
    if state is None:
        state = zero_state(n, ctype)
    for kind, payload in prog:
        if kind == "segment":
            run_segment(state, P, payload, ctype)
        else:
            state = remap(state, n, payload)
    return unpermute(state, n, final_map)


def zero_segment(L, k):
    """The reference runtime's literal zeroing segment (host/sv/runtime.py:205-209): one pass,
    one MUX2(t=0, cmask=0, Z, Z) with Z = 0."""
    w = [MAGIC, L, 1, k, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0] + [OP_MUX2, 0, 0, 0, 0, 0, 0, 0] + [0] * 16
    return np.asarray(w, dtype=np.int64).astype(np.int32)
