"""Parity of the fused dense-unitary form (OP_DENSE records applied on matrix cores inside the pass
kernel) with the complex128 golden model of the same records.

Tolerance: the FP32 rounding bound per amplitude of tests/_c64.py against the complex128 oracle run
of the same records (dense units run on the FP32 matrix cores by default).
"""
import numpy as np
import pytest

from oracle import dense, judge, model
from oracle.isa import MAGIC
from u55c_statevector_30q_amd import circuit as C
from u55c_statevector_30q_amd import compiler as K

from test_gpu_pass import _mat, _op_words, _rand_state, run_hip
from _c64 import C_BOUND, amp_error_ratio, check_c64, prog_nops, seg_nops

pytestmark = pytest.mark.gpu


def _rand_unitary(rng, d):
    z = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def dense_segment(rng, L, k, n_passes, fs=(1, 2, 3, 4, 5), max_m=4, n_ops=6, mux_p=0.3):
    """Passes mixing dense records of random widths/qubits with MUX2 records."""
    w = [MAGIC, L, n_passes, k, 0, 0, 0, 0]
    for _ in range(n_passes):
        m = int(rng.integers(0, min(max_m, k - 5) + 1))
        hi = sorted(rng.choice(list(range(k - m, L)), size=m, replace=False).tolist()) if m else []
        h = hi + [0] * (11 - len(hi))
        w += [m, h[0], h[1], n_ops, h[2], h[3], 0, 0]
        if m > 4:
            w += [0x48] + h[4:11]
        for _ in range(n_ops):
            if rng.random() < mux_p:
                t = int(rng.integers(0, k))
                c = int(rng.integers(0, k))
                cm = 0 if c == t else 1 << c
                w += _op_words(t, cm, _mat(rng, "u"), _mat(rng, "u"))
            else:
                f = int(rng.choice(fs))
                qs = rng.choice(k, size=f, replace=False).tolist()
                w += K.dense_words(_rand_unitary(rng, 1 << f), qs)
    return (np.asarray(w, dtype=np.int64) & 0xFFFFFFFF).astype(np.uint32).view(np.int32)


def _check(ref128, got64, what, nops):
    check_c64(ref128, got64, what, nops)


@pytest.mark.parametrize("n,k", [(5, 5), (9, 8), (10, 10), (12, 11), (14, 12), (16, 13), (18, 13), (20, 14),
                                 (21, 12), (22, 13)])
def test_random_dense_records(gpu, n, k):
    """Dense records of every width 1..5 (widths below 4 are padded to 4 on the device) on random
    tile qubits, mixed with MUX2 records, with gathered high qubits."""
    rng = np.random.default_rng(77 * n + k)
    seg = dense_segment(rng, n, k, n_passes=3, max_m=min(7, k - 6) if k > 10 else 4)
    s0 = _rand_state(rng, n)
    ref = model.run_segment(s0.copy(), 1, seg, np.complex128)
    got = run_hip(s0.astype(np.complex64), seg)
    _check(ref, got, (n, k), seg_nops(seg))


@pytest.mark.parametrize("f", [4, 5])
def test_dense_permutation_matrix_is_exact(gpu, f):
    """A permutation matrix routes every amplitude through one MFMA product with 1.0 and sums with
    exact zeros: the result must equal the c64 golden model bit for bit."""
    rng = np.random.default_rng(f)
    n, k = (14, 12) if f == 4 else (22, 13)
    w = [MAGIC, n, 2, k, 0, 0, 0, 0]
    for _ in range(2):
        w += [2, 12, 13, 2, 0, 0, 0, 0]
        for _ in range(2):
            P = np.eye(1 << f)[rng.permutation(1 << f)]
            w += K.dense_words(P, rng.choice(k, size=f, replace=False).tolist())
    seg = (np.asarray(w, dtype=np.int64) & 0xFFFFFFFF).astype(np.uint32).view(np.int32)
    s0 = _rand_state(rng, n, np.complex64)
    ref = model.run_segment(s0.copy(), 1, seg, np.complex64)
    assert np.array_equal(run_hip(s0, seg), ref)


@pytest.mark.parametrize("n,width", [(9, 4), (12, 4), (12, 5), (16, 4), (16, 5), (21, 5), (22, 4)])
def test_compile_fused_vs_golden_model(gpu, n, width):
    """compile_fused streams of the u3+cx layer recipe: GPU c64 against the c128 golden model of
    the same records, and (n <= 12) against the gate-level simulator."""
    import torch
    from u55c_statevector_30q_amd import libsvx
    qc = C.u3_cx_layers(n, 10, 5)
    prog, fmap, st = K.compile_fused(qc, width=width)
    assert st.n_dense > 0
    gold = model.run_program(prog, n, 1, fmap, ctype=np.complex128)
    s = torch.empty(1 << n, dtype=torch.complex64, device="cuda")
    libsvx.zero_state(s)
    for kind, w in prog:
        assert kind == "segment"
        libsvx.run_segment(s, w)
    torch.cuda.synchronize()
    got = model.unpermute(s.cpu().numpy(), n, fmap)
    _check(gold, got, (n, width), prog_nops(prog))
    if n <= 12:
        _check(dense.simulate(qc.tuples(), n), got, (n, width, "dense"), prog_nops(prog))


def test_dense_records_rejected_for_c128(gpu):
    import torch
    from u55c_statevector_30q_amd import libsvx
    rng = np.random.default_rng(1)
    seg = dense_segment(rng, 12, 12, n_passes=1, fs=(4,))
    t = torch.zeros(1 << 12, dtype=torch.complex128, device="cuda")
    with pytest.raises(ValueError):
        libsvx.run_segment(t, seg)


@pytest.mark.parametrize("scale", [1.0, 1e5, 1e-20])
def test_split_fp16_and_fp32_paths(gpu, scale):
    """Both matrix-core paths for general dense records: FP32 MFMA (the default, held to the FP32
    bound) and the split-FP16 opt-in (per-column power-of-two scaling, so unnormalised states far
    from unit norm keep FP32-level fidelity)."""
    from u55c_statevector_30q_amd import libsvx
    rng = np.random.default_rng(11)
    n, k = 16, 13
    seg = dense_segment(rng, n, k, n_passes=3, fs=(4, 5), mux_p=0.0)
    s0 = _rand_state(rng, n) * scale
    ref = model.run_segment(s0.copy(), 1, seg, np.complex128)
    for math in ("f32", "split16"):
        with libsvx.options(dense_math=math):
            got = run_hip(s0.astype(np.complex64), seg)
        m = judge.metrics(ref, got)
        assert m["one_minus_f"] < 1e-9, (math, scale, m)
        if math == "f32":
            _check(ref, got, (math, scale), seg_nops(seg))


@pytest.mark.parametrize("n,k", [(21, 12), (22, 13)])
def test_dense_records_op_by_op_mux2_kernel(gpu, n, k):
    """Dense records with pass fusion off and every MUX2 op on the vector units (the single-buffer
    kernels: FP32 dense units on the matrix cores between the MUX2 register groups) against the
    complex128 golden model."""
    from u55c_statevector_30q_amd import libsvx
    rng = np.random.default_rng(5 * n + k)
    seg = dense_segment(rng, n, k, n_passes=2, max_m=4)
    s0 = _rand_state(rng, n)
    ref = model.run_segment(s0.copy(), 1, seg, np.complex128)
    with libsvx.options(fuse="off", diag_tables=False, mux_engine="valu"):
        got = run_hip(s0.astype(np.complex64), seg)
    _check(ref, got, (n, k), seg_nops(seg))


@pytest.mark.parametrize("n,recipe", [(16, "u3cx"), (18, "u3cx"), (16, "rqc"), (17, "qft")])
def test_cluster_scheduled_mux2_stream_vs_golden_model(gpu, n, recipe):
    """MUX2 streams whose passes were drawn cluster by cluster (compile_gpu's clustered schedule),
    run with the library's default pass fusion (whole clusters become FP32 dense units): GPU c64
    against the c128 golden model of the same records."""
    import torch
    from u55c_statevector_30q_amd import libsvx
    qc = C.u3_cx_layers(n, 6, 7) if recipe == "u3cx" else C.make(recipe, n)
    prog, fmap, st = K.compile_gpu(qc, cluster_always=True)
    assert st.clustered
    gold = model.run_program(prog, n, 1, fmap, ctype=np.complex128)
    s = torch.empty(1 << n, dtype=torch.complex64, device="cuda")
    libsvx.zero_state(s)
    hs = [libsvx.prepare(w, libsvx.C64, n) for kind, w in prog]
    units = sum(libsvx.prepared_pass_stats(h, p)["dense4"] for h in hs for p in range(libsvx.prepared_info(h)[0]))
    for h in hs:
        libsvx.run_prepared(s, h)
    torch.cuda.synchronize()
    for h in hs:
        libsvx.release_prepared(h)
    if recipe == "u3cx":
        assert units > 0
    got = model.unpermute(s.cpu().numpy(), n, fmap)
    _check(gold, got, (n, recipe), prog_nops(prog))
