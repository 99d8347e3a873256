"""Parity of the HIP pass kernel (libsvx svx_run_segment) with the CPU golden model.

Tolerances: the complex128 kernel must match the complex128 oracle to 1e-12 (FMA contraction is
the only difference); the complex64 kernel must stay within the FP32 rounding bound per amplitude
of tests/_c64.py (C * nops * 2^-24 * (|a_i| + rms)) and 1 - F < 1e-9 against the complex128
oracle run of the same records.  Ops that only move data (exact X, controlled exact X) must be
bit-exact.
"""
import numpy as np
import pytest

from oracle import judge, model
from oracle.isa import MAGIC, OP_MUX2
from _c64 import check_c64, seg_nops

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def _lib():
    from u55c_statevector_30q_amd import libsvx
    return libsvx


def _rand_state(rng, n, dtype=np.complex128):
    s = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return (s / np.linalg.norm(s)).astype(dtype)


def _rand_u(rng):
    z = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def _mat(rng, kind):
    if kind == "u":
        return _rand_u(rng)
    if kind == "i":
        return np.eye(2)
    if kind == "x":
        return np.array([[0, 1], [1, 0]])
    if kind == "d":
        return np.diag(np.exp(1j * rng.uniform(0, 6.28, 2)))
    if kind == "p":
        return np.diag([1, np.exp(1j * rng.uniform(0, 6.28))])
    raise ValueError(kind)


OP_KINDS = [("u", "u"), ("d", "d"), ("x", "x"), ("i", "u"), ("i", "d"), ("i", "p"), ("i", "x"), ("u", "d"),
            ("p", "x"), ("i", "i")]


def _op_words(t, cmask, M0, M1):
    fl = np.zeros(16, dtype=np.float32)
    for i, M in enumerate((M0, M1)):
        f = np.asarray(M, dtype=np.complex64).reshape(-1)
        fl[8 * i:8 * i + 8:2] = f.real
        fl[8 * i + 1:8 * i + 8:2] = f.imag
    return [OP_MUX2, t, cmask, 0, 0, 0, 0, 0] + [int(x) for x in fl.view(np.int32)]


def random_segment(rng, L, k, n_passes, max_ops=12, kinds=OP_KINDS, max_m=4, ctrl_p=0.3):
    w = [MAGIC, L, n_passes, k, 0, 0, 0, 0]
    for _ in range(n_passes):
        m = int(rng.integers(0, min(max_m, k) + 1))
        hi = sorted(rng.choice(list(range(k - m, L)), size=m, replace=False).tolist()) if m else []
        h = hi + [0] * (11 - len(hi))
        nops = int(rng.integers(1, max_ops + 1))
        w += [m, h[0], h[1], nops, h[2], h[3], 0, 0]
        if m > 4:
            w += [0x48] + h[4:11]
        for _ in range(nops):
            t = int(rng.integers(0, k))
            cm = 0
            for b in range(k):
                if b != t and rng.random() < ctrl_p:
                    cm |= 1 << b
            a, b = kinds[int(rng.integers(len(kinds)))]
            w += _op_words(t, cm, _mat(rng, a), _mat(rng, b))
    return (np.asarray(w, dtype=np.int64) & 0xFFFFFFFF).astype(np.uint32).view(np.int32)


def run_hip(state_np, words, P=1):
    torch = _torch()
    t = torch.from_numpy(state_np.copy()).to("cuda")
    _lib().run_segment(t, words, n_slices=P)
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _check_c64(ref128, got64, what, seg):
    check_c64(ref128, got64, what, seg_nops(seg))


def _opts(**kw):
    return _lib().options(**kw)


CASES = [(4, 3, 1), (6, 4, 1), (8, 5, 2), (9, 6, 1), (10, 6, 1), (10, 10, 1), (11, 8, 2), (12, 12, 1),
         (13, 9, 4), (14, 12, 1), (14, 13, 1), (15, 11, 2), (16, 12, 1), (18, 13, 1), (20, 12, 1), (20, 14, 1)]


@pytest.mark.parametrize("n,k,P", CASES)
def test_random_segments_c128_exact_and_c64(gpu, n, k, P):
    rng = np.random.default_rng(1000 * n + 10 * k + P)
    L = n - (P.bit_length() - 1)
    k = min(k, L)
    max_m = min(8, k) if k >= 8 else min(4, k)
    seg = random_segment(rng, L, k, n_passes=3, max_m=max_m)
    s0 = _rand_state(rng, n)
    ref = model.run_segment(s0.copy(), P, seg, np.complex128)
    if k <= 13:
        got128 = run_hip(s0, seg, P)
        assert np.max(np.abs(got128 - ref)) < 1e-12, (n, k, P)
    got64 = run_hip(s0.astype(np.complex64), seg, P)
    _check_c64(ref, got64, (n, k, P), seg)


@pytest.mark.parametrize("k", [1, 2, 3, 5, 8])
def test_every_target_and_control_small_tiles(gpu, k):
    """Every target position and control pattern of small tiles (k <= 8), the case the reference
    found only on hardware (docs/BASELINE_AUDIT.md:61-67)."""
    rng = np.random.default_rng(k)
    for n in (k + 2, 12):
        if n < k:
            continue
        w = [MAGIC, n, 0, k, 0, 0, 0, 0]
        npass = 0
        for t in range(k):
            ops = []
            for cm in ([0] + [1 << b for b in range(k) if b != t] + [((1 << k) - 1) & ~(1 << t)]):
                for a, b in (("u", "u"), ("i", "u"), ("i", "x"), ("i", "p"), ("u", "d")):
                    ops.append(_op_words(t, cm, _mat(rng, a), _mat(rng, b)))
            for i in range(0, len(ops), 64):
                chunk = ops[i:i + 64]
                m = min(2, k)
                hi = sorted(rng.choice(list(range(k - m, n)), size=m, replace=False).tolist())
                h = hi + [0] * (4 - len(hi))
                w += [m, h[0], h[1], len(chunk), h[2], h[3], 0, 0]
                for o in chunk:
                    w += o
                npass += 1
        w[2] = npass
        seg = (np.asarray(w, dtype=np.int64) & 0xFFFFFFFF).astype(np.uint32).view(np.int32)
        s0 = _rand_state(rng, n)
        ref = model.run_segment(s0.copy(), 1, seg, np.complex128)
        got = run_hip(s0, seg)
        assert np.max(np.abs(got - ref)) < 1e-11, (k, n)


@pytest.mark.parametrize("n,k", [(8, 4), (12, 10), (16, 13), (22, 13)])
def test_permutation_ops_bit_exact(gpu, n, k):
    """X and controlled-X move data only: the c64 kernel must equal the c64 golden model bit for bit."""
    rng = np.random.default_rng(n)
    seg = random_segment(rng, n, k, n_passes=4, kinds=[("x", "x"), ("i", "x")], max_m=min(8, k) if k >= 8 else 4)
    s0 = _rand_state(rng, n, np.complex64)
    ref = model.run_segment(s0.copy(), 1, seg, np.complex64)
    got = run_hip(s0, seg)
    assert np.array_equal(got, ref)


def test_zero_segment_and_errors(gpu):
    from u55c_statevector_30q_amd import libsvx
    torch = _torch()
    rng = np.random.default_rng(0)
    s0 = _rand_state(rng, 12, np.complex64)
    got = run_hip(s0, model.zero_segment(12, 12))
    assert not got.any()
    t = torch.zeros(1 << 12, dtype=torch.complex64, device="cuda")
    bad = model.zero_segment(12, 12).copy()
    bad[0] = 7
    with pytest.raises(AssertionError):
        libsvx.run_segment(t, bad)
    with pytest.raises(ValueError):
        libsvx.run_segment(t, model.zero_segment(12, 12)[:20])
    with pytest.raises(ValueError):
        libsvx.run_segment(t, model.zero_segment(11, 11))      # L does not match the slice
    big = torch.zeros(1 << 16, dtype=torch.complex128, device="cuda")
    with pytest.raises(RuntimeError):
        libsvx.run_segment(big, model.zero_segment(16, 14))   # c128 tile of 2^14 exceeds LDS


def test_engine_reference_abi(gpu):
    """svx_engine(st, ins_off): instructions stored after the slice in the same buffer
    (kernels/sv/sv_engine.cpp:290-302, host/sv/runtime.py:218-233)."""
    torch = _torch()
    from u55c_statevector_30q_amd import libsvx
    rng = np.random.default_rng(4)
    L = 12
    seg = random_segment(rng, L, 10, n_passes=2)
    s0 = _rand_state(rng, L, np.complex64)
    ins_off = (1 << L) * 8 // 32
    buf = np.zeros((1 << L) * 8 + seg.nbytes + 64, dtype=np.uint8)
    buf[:(1 << L) * 8] = s0.view(np.uint8)
    buf[(1 << L) * 8:(1 << L) * 8 + seg.nbytes] = seg.view(np.uint8)
    d = torch.from_numpy(buf).to("cuda")
    libsvx.engine(d, ins_off)
    torch.cuda.synchronize()
    got = d.cpu().numpy()[:(1 << L) * 8].view(np.complex64)
    ref = model.run_segment(s0.astype(np.complex128), 1, seg, np.complex128)
    _check_c64(ref, got, "engine", seg)


def test_prepared_replay_is_identical(gpu):
    torch = _torch()
    from u55c_statevector_30q_amd import libsvx
    rng = np.random.default_rng(9)
    seg = random_segment(rng, 16, 13, n_passes=5, max_m=7)
    s0 = _rand_state(rng, 16, np.complex64)
    h = libsvx.prepare(seg, libsvx.C64, 16)
    outs = []
    for _ in range(2):
        t = torch.from_numpy(s0.copy()).to("cuda")
        libsvx.run_prepared(t, h)
        torch.cuda.synchronize()
        outs.append(t.cpu().numpy())
    assert np.array_equal(outs[0], outs[1])
    npass, nlaunch = libsvx.prepared_info(h)
    assert npass == 5 and nlaunch <= 5


@pytest.mark.parametrize("n,k,max_m", [(21, 12, 4), (21, 12, 7), (22, 13, 7), (22, 13, 8)])
def test_mux_kernel_vector_units_matches_oracle(gpu, n, k, max_m):
    """The FP32 MUX2 kernel with every op on the vector units (single-buffer, two register buffers,
    ops regrouped over their commutation DAG) matches the complex128 golden model to the FP32
    bound at QSV2 tiles of 2^12 and 2^13 amplitudes with up to 8 gathered high qubits."""
    rng = np.random.default_rng(n * 100 + k + max_m)
    seg = random_segment(rng, n, k, n_passes=2, max_ops=10, max_m=max_m)
    s0 = _rand_state(rng, n, np.complex64)
    ref = model.run_segment(s0.astype(np.complex128), 1, seg, np.complex128)
    with _opts(fuse="off", diag_tables=False, mux_engine="valu"):
        got = run_hip(s0, seg)
    _check_c64(ref, got, (n, k, max_m), seg)


FUSE_SETTINGS = {"off": dict(fuse="off"), "auto": dict(fuse="auto"),
                 "all5": dict(fuse="auto", fuse_cost4=-1, fuse_cost5=-1, fuse_width=5),
                 "all3": dict(fuse="auto", fuse_cost4=-1, fuse_cost5=-1, fuse_width=3),
                 "all1": dict(fuse="auto", fuse_cost4=-1, fuse_cost5=-1, fuse_width=1),
                 "dense4": dict(fuse="dense4"), "dense4_split16": dict(fuse="dense4", dense_math="split16")}


@pytest.mark.parametrize("fuse", list(FUSE_SETTINGS))
@pytest.mark.parametrize("n,k,max_m", [(12, 10, 4), (16, 12, 6), (20, 13, 7), (21, 12, 4), (22, 13, 7)])
def test_pass_fusion_matches_oracle(gpu, n, k, max_m, fuse):
    """Pass fusion (MUX2 ops of a pass clustered into dense units on the matrix cores): every
    setting, from off to fusing every block of 1, 3 or 5 tile bits, matches the complex128 golden
    model; ops with controls and several kinds of matrices, including controls wider than a block.
    FP32 units are held to the FP32 bound; the split-FP16 opt-in only to FP32-level fidelity."""
    rng = np.random.default_rng(n * 31 + k + max_m)
    seg = random_segment(rng, n, k, n_passes=3, max_ops=40, max_m=max_m, ctrl_p=0.12)
    s0 = _rand_state(rng, n)
    ref = model.run_segment(s0.copy(), 1, seg, np.complex128)
    with _opts(**FUSE_SETTINGS[fuse]):
        got = run_hip(s0.astype(np.complex64), seg)
    if fuse.endswith("split16"):
        m = judge.metrics(ref, got)
        assert m["one_minus_f"] < 1e-9, (n, k, fuse, m)
    else:
        _check_c64(ref, got, (n, k, max_m, fuse), seg)


@pytest.mark.parametrize("n,k", [(16, 12), (22, 13)])
def test_pass_fusion_of_permutations_is_exact(gpu, n, k):
    """Blocks of exact X / controlled-X ops fuse into 0/1 matrices, which run on the FP32 matrix
    cores exactly (also when split FP16 is requested): the result equals the c64 golden model bit
    for bit."""
    rng = np.random.default_rng(5)
    seg = random_segment(rng, n, k, n_passes=3, max_ops=40, kinds=[("x", "x"), ("i", "x")], max_m=6)
    s0 = _rand_state(rng, n, np.complex64)
    ref = model.run_segment(s0.copy(), 1, seg, np.complex64)
    for math in ("f32", "split16"):
        with _opts(**FUSE_SETTINGS["all5"], dense_math=math):
            assert np.array_equal(run_hip(s0, seg), ref), math


def _gain_segment(rng, L, k, n_passes, n_ops, gain, max_m=6):
    """Passes of random (controlled) 2x2 ops multiplied by `gain` (non-unitary for gain != 1)."""
    w = [MAGIC, L, n_passes, k, 0, 0, 0, 0]
    for _ in range(n_passes):
        m = int(rng.integers(0, min(max_m, k - 6) + 1))
        hi = sorted(rng.choice(list(range(k - m, L)), size=m, replace=False).tolist()) if m else []
        h = hi + [0] * (11 - len(hi))
        w += [m, h[0], h[1], n_ops, h[2], h[3], 0, 0]
        if m > 4:
            w += [0x48] + h[4:11]
        for _ in range(n_ops):
            t = int(rng.integers(0, k))
            c = int(rng.integers(0, k))
            cm = 0 if c == t or rng.random() < 0.5 else 1 << c
            w += _op_words(t, cm, gain * _rand_u(rng), gain * _rand_u(rng))
    return (np.asarray(w, dtype=np.int64) & 0xFFFFFFFF).astype(np.uint32).view(np.int32)


# unit bit sets whose lowest four free tile bits put two lanes of a 16-lane LDS access on one bank
# (the library then maps the 16 columns of a block to other free bits); listed with conflict-free
# sets in the same pass so both column mappings meet in one tile
_CONFLICTING_UNITS = [(3, 4, 5, 6), (2, 3, 4, 7), (1, 3, 5, 6), (0, 3, 4, 5), (2, 4, 6, 8), (1, 4, 5, 6),
                      (0, 1, 4, 5), (2, 3, 6, 7)]


@pytest.mark.parametrize("n,k,hi", [(20, 13, [16, 17, 18, 19]), (22, 13, [])])
def test_dense_units_with_remapped_columns(gpu, n, k, hi):
    """Fused-form dense records on tile-bit sets for which the library moves the column lane bits
    off the lowest free bits (to keep LDS reads free of bank conflicts), interleaved with units on
    ordinary sets: the complex64 result against the complex128 golden model, to the FP32 bound."""
    rng = np.random.default_rng(n + len(hi))
    units = []
    for i, u in enumerate(_CONFLICTING_UNITS):
        units.append(tuple(int(b) for b in rng.permutation(u)))   # matrix bit order need not be ascending
        if i % 2:
            units.append(tuple(int(b) for b in rng.choice(13, 4, replace=False)))
    seg = _dense_pass_words(rng, n, k, hi, units)
    s0 = _rand_state(rng, n)
    ref = model.run_segment(s0.copy(), 1, seg, np.complex128)
    got = run_hip(s0.astype(np.complex64), seg)
    _check_c64(ref, got, units, seg)


@pytest.mark.parametrize("gain,scale", [(1.0, 1.0), (1.0, 1e6), (1.0, 1e-25), (1.15, 1.0), (2.0, 1.0), (0.5, 1e3)])
@pytest.mark.parametrize("math", ["f32", "split16_tile", "split16_col"])
def test_pass_fusion_tile_scaling(gpu, gain, scale, math):
    """Dense groups on unnormalised states far from unit norm and non-unitary ops that grow or
    shrink the amplitudes within a pass: FP32 matrix cores, and the split-FP16 opt-in on a
    pre-scaled tile (the pass kernel scales each tile by a power of two from the tile maximum and
    the host's growth bound of the pass's matrices) and with per-column scaling, all match the
    complex128 golden model."""
    rng = np.random.default_rng(int(gain * 100) + int(np.log10(scale)) + 50)
    n, k = 18, 13
    seg = _gain_segment(rng, n, k, n_passes=3, n_ops=24, gain=gain)
    s0 = _rand_state(rng, n) * scale
    ref = model.run_segment(s0.copy(), 1, seg, np.complex128)
    kw = dict(fuse="auto", fuse_cost4=-1, fuse_cost5=-1, fuse_width=5)
    if math != "f32":
        kw.update(dense_math="split16", tile_scale=(math == "split16_tile"))
    with _opts(**kw):
        got = run_hip(s0.astype(np.complex64), seg)
    assert np.all(np.isfinite(got))
    if math == "f32":
        _check_c64(ref, got, (gain, scale, math), seg)
    else:
        m = judge.metrics(ref, got)
        assert m["one_minus_f"] < 1e-9, (gain, scale, math, m)
    nr = np.linalg.norm(ref)
    assert abs(np.linalg.norm(got.astype(np.complex128)) / nr - 1) < 1e-5


@pytest.mark.parametrize("math", ["f32", "split16"])
@pytest.mark.parametrize("n,k", [(18, 12), (20, 13), (22, 13)])
def test_dense_only_kernel_matches_general_kernel(gpu, n, k, math):
    """Every op of every pass in a dense 4-qubit unit (fuse=dense4).  Split FP16: the passes run on
    the 512-thread dense-only kernel, bit-identical to the general dense kernel (dense_only off: same
    units, same arithmetic, other thread count).  FP32: the MUX2 kernel runs them either way.  Both
    match the complex128 golden model."""
    from u55c_statevector_30q_amd import libsvx
    rng = np.random.default_rng(3 * n + k)
    seg = random_segment(rng, n, k, n_passes=3, max_ops=30, max_m=6, ctrl_p=0.08)
    s0 = _rand_state(rng, n)
    ref = model.run_segment(s0.copy(), 1, seg, np.complex128)
    got, names = {}, {}
    for d4 in (True, False):
        with _opts(fuse="dense4", dense_math=math, dense_only=d4):
            got[d4] = run_hip(s0.astype(np.complex64), seg)
            h = libsvx.prepare(seg, libsvx.C64, n)
            names[d4] = [libsvx.prepared_kernel(h, p) for p in range(3)]
            libsvx.release_prepared(h)
        if math == "f32":
            _check_c64(ref, got[d4], (n, k, d4), seg)
        else:
            assert judge.metrics(ref, got[d4])["one_minus_f"] < 1e-9
    if math == "f32":
        assert all(nm.startswith("k_mux<") for nm in names[True] + names[False]), names
    else:
        assert any(nm.startswith("k_pass_d4<") for nm in names[True]), names
        assert not any(nm.startswith("k_pass_d4<") for nm in names[False]), names
    assert np.array_equal(got[True], got[False]), names


def _dense_pass_words(rng, n, k, hi, unit_qubits):
    """One pass of fused-form dense records (random 16 x 16 unitaries) on the given tile qubits."""
    from u55c_statevector_30q_amd import compiler
    w = [MAGIC, n, 1, k, 0, 0, 0, 0]
    m = len(hi)
    h = list(hi) + [0] * (11 - m)
    w += [m, h[0], h[1], len(unit_qubits), h[2], h[3], 0, 0]
    if m > 4:
        w += [0x48] + h[4:11]
    for qs in unit_qubits:
        z = rng.normal(size=(16, 16)) + 1j * rng.normal(size=(16, 16))
        q, r = np.linalg.qr(z)
        w += compiler.dense_words(q * (np.diag(r) / np.abs(np.diag(r))), qs)
    return (np.asarray(w, dtype=np.int64) & 0xFFFFFFFF).astype(np.uint32).view(np.int32)


@pytest.mark.parametrize("units", [
    [(0, 1, 2, 3), (4, 5, 6, 7)],                          # tile bit 0 in the first unit: 16-byte loads
    [(5, 6, 7, 8), (9, 10, 11, 12)],                       # tile bit 0 outside every unit
    [(3, 0, 9, 6), (12, 2, 7, 10), (1, 4, 5, 8), (11, 0, 2, 3)],   # bit 0 inside, unordered bits
])
def test_dense_units_on_chosen_bits(gpu, units):
    """Fused-form dense records on chosen tile bits (20 qubits, 2^13 tiles, 4 high qubits), both
    16-byte LDS paths of the dense-only kernel, bits in any order; vs the golden model, FP32 units
    held to the FP32 bound."""
    n, k = 20, 13
    rng = np.random.default_rng(sum(sum(u) for u in units))
    seg = _dense_pass_words(rng, n, k, [16, 17, 18, 19], units)
    s0 = _rand_state(rng, n)
    ref = model.run_segment(s0.copy(), 1, seg, np.complex128)
    got = run_hip(s0.astype(np.complex64), seg)
    _check_c64(ref, got, units, seg)


@pytest.mark.parametrize("math", ["f32", "split16_tile", "split16_col"])
@pytest.mark.parametrize("n,k", [(24, 13), (23, 12)])
def test_dense_only_kernel_large_states(gpu, n, k, math):
    """Dense-only passes on states of 2^11 tiles: split FP16 (pre-scaled tiles and per-column
    scaling) on the dense-only kernel, bit-identical to the general dense kernel; FP32 on the MUX2
    kernel either way; norm kept."""
    from u55c_statevector_30q_amd import libsvx
    rng = np.random.default_rng(5 * n + k)
    seg = random_segment(rng, n, k, n_passes=3, max_ops=30, max_m=6, ctrl_p=0.08)
    s0 = _rand_state(rng, n, np.complex64)
    kw = dict(fuse="dense4")
    if math != "f32":
        kw.update(dense_math="split16", tile_scale=(math == "split16_tile"))
    got, names = {}, {}
    for d4 in (True, False):
        with _opts(**kw, dense_only=d4):
            h = libsvx.prepare(seg, libsvx.C64, n)
            names[d4] = [libsvx.prepared_kernel(h, p) for p in range(libsvx.prepared_info(h)[1])]
            libsvx.release_prepared(h)
            got[d4] = run_hip(s0, seg)
    if math == "f32":
        assert all(nm.startswith("k_mux<") for nm in names[True] + names[False]), names
    else:
        assert any(nm.startswith("k_pass_d4<") for nm in names[True]), names
        assert not any(nm.startswith("k_pass_d4<") for nm in names[False]), names
    assert np.array_equal(got[True], got[False])
    assert abs(np.linalg.norm(got[True].astype(np.complex128)) - 1) < 1e-5


@pytest.mark.parametrize("dt", ["c64", "c128"])
def test_tile_order_tuning_keeps_results(gpu, dt):
    """svx_tune_prepared picks each pass's workgroup-to-tile order by timing; the order only changes
    which workgroup takes which tile, so the tuned run, the default order and the fixed identity
    order give bit-identical states, and the stats report the order in use."""
    from u55c_statevector_30q_amd import libsvx
    torch = _torch()
    n, k = 22, 13
    rng = np.random.default_rng(77 if dt == "c64" else 78)
    seg = random_segment(rng, n, k, n_passes=4, max_ops=20, max_m=7)
    npdt = np.complex64 if dt == "c64" else np.complex128
    code = libsvx.C64 if dt == "c64" else libsvx.C128
    s0 = _rand_state(rng, n, npdt)
    ref = model.run_segment(s0.astype(np.complex128), 1, seg, np.complex128)

    def run(h):
        t = torch.from_numpy(s0.copy()).to("cuda")
        libsvx.run_prepared(t, h)
        torch.cuda.synchronize()
        return t.cpu().numpy()

    h = libsvx.prepare(seg, code, n)
    before = run(h)
    scratch = torch.from_numpy(s0.copy()).to("cuda")
    libsvx.tune_prepared(scratch, h, reps=1)
    orders = [libsvx.prepared_pass_stats(h, p)["tile_order"] for p in range(libsvx.prepared_info(h)[1])]
    assert set(orders) <= {-2, 0}, orders
    tuned = run(h)
    with _opts(tile_order="identity"):
        hi = libsvx.prepare(seg, code, n)
    assert all(libsvx.prepared_pass_stats(hi, p)["tile_order"] == 0 for p in range(libsvx.prepared_info(hi)[1]))
    ident = run(hi)
    assert np.array_equal(before, tuned) and np.array_equal(before, ident)
    if dt == "c64":
        _check_c64(ref, before, (n, k), seg)
    else:
        assert np.max(np.abs(before - ref)) < 1e-12
    with pytest.raises(ValueError):
        libsvx.tune_prepared(scratch, h, reps=0)
