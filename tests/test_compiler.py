"""Compiler restatement (host/sv/compiler.py) against the reference's recorded statistics, the
oracle's independent gate derivations and the gate-level complex128 simulator (the reference's
tests/test_compiler_model.py strategy: families and gate zoos, swept over P, k, fuse, tiling and
initial map, compiled then executed by the golden model)."""
import itertools
import json
import math
import os

import numpy as np
import pytest

from oracle import dense, judge, model
from u55c_statevector_30q_amd import circuit as C
from u55c_statevector_30q_amd import compiler as K
from u55c_statevector_30q_amd import gates

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def zoo(n, seed):
    """Random circuit over the gate zoo (same gate families as tests/circuit_zoo.py:6-22)."""
    rng = np.random.default_rng(seed)
    qc = C.Circuit(n, f"zoo_{n}_{seed}")
    one = ["h", "x", "y", "z", "s", "sdg", "t", "tdg", "sx", "sxdg"]
    par1 = ["rx", "ry", "rz", "p"]
    two = ["cx", "cy", "cz", "ch", "swap", "iswap", "ecr", "dcx", "csx"]
    par2 = ["rxx", "ryy", "rzz", "rzx", "cp", "crx", "cry", "crz"]
    for _ in range(12 * n):
        r = rng.integers(6)
        q = [int(x) for x in rng.choice(n, 3, replace=False)]
        th = float(rng.uniform(-np.pi, np.pi))
        if r == 0:
            getattr(qc, str(rng.choice(one)))(q[0])
        elif r == 1:
            getattr(qc, str(rng.choice(par1)))(th, q[0])
        elif r == 2:
            getattr(qc, str(rng.choice(two)))(q[0], q[1])
        elif r == 3:
            getattr(qc, str(rng.choice(par2)))(th, q[0], q[1])
        elif r == 4:
            qc.u(th, th / 2, -th / 3, q[0])
        else:
            qc.ccx(q[0], q[1], q[2])
    return qc


GATES_1Q = [("h", 0), ("x", 0), ("y", 0), ("z", 0), ("s", 0), ("sdg", 0), ("t", 0), ("tdg", 0), ("sx", 0),
            ("sxdg", 0), ("rx", 1), ("ry", 1), ("rz", 1), ("p", 1), ("u", 3), ("u2", 2), ("r", 2), ("id", 0)]
GATES_2Q = [("cx", 0), ("cy", 0), ("cz", 0), ("ch", 0), ("swap", 0), ("iswap", 0), ("ecr", 0), ("dcx", 0),
            ("csx", 0), ("cp", 1), ("crx", 1), ("cry", 1), ("crz", 1), ("rxx", 1), ("ryy", 1), ("rzz", 1),
            ("rzx", 1), ("cu", 4), ("cs", 0), ("csdg", 0), ("xx_plus_yy", 2)]
PAR = [0.7311, -1.2207, 2.4113, 0.3359]


@pytest.mark.parametrize("name,np_", GATES_1Q + GATES_2Q + [("ccx", 0), ("ccz", 0), ("cswap", 0)])
def test_gate_table_vs_oracle_derivation(name, np_):
    nq = 1 if (name, np_) in GATES_1Q else (2 if (name, np_) in GATES_2Q else 3)
    A = gates.matrix(name, nq, PAR[:np_])
    B = dense.gate_matrix(name, nq, PAR[:np_])
    assert np.allclose(A, B, atol=1e-12), name


def _load_stats():
    return json.load(open(os.path.join(GOLDEN, "sched_stats.json")))


def _instance(r):
    qc = C.make(r["family"], r["n"], depth=r["depth"]) if r["family"] == "rqc" else C.make(r["family"], r["n"])
    return qc.padded(r["n_run"])


def _imap(layout, n, P):
    """bench/fpga/sv_bench.py:20-26"""
    if layout == "block":
        return None
    b = P.bit_length() - 1
    return [(n - b + l) if l < b else (l - b) for l in range(n)]


def test_op_lists_match_every_recorded_run():
    """n_ops and fused_1q of all 285 recorded reference runs (every family, n = 4..30)."""
    recs = _load_stats()
    assert len(recs) == 285
    for r in recs:
        st = K.Stats()
        K.circuit_to_ops(_instance(r), st, bool(r["fuse"]), max_width=r["k"])
        assert (st.n_ops, st.fused_1q) == (r["stats"]["n_ops"], r["stats"]["fused_1q"]), r


def test_schedule_matches_recorded_ghz_runs():
    """GHZ's op list is a dependency chain, so every order-respecting scheduler agrees: passes,
    remaps and segments are pinned exactly by all 60 recorded GHZ runs (P = 1..32 slices, k = 9..12,
    MAXH 2 and 4, block/modulo layouts, tiling on/off, fusion on/off)."""
    recs = [r for r in _load_stats() if r["family"] == "ghz"]
    assert len(recs) == 60
    for r in recs:
        prog, fmap, st = K.compile_circuit(_instance(r), r["L"], r["k"], fuse_1q=bool(r["fuse"]),
                                           tiling=bool(r["tiling"]), initial_map=_imap(r["layout"], r["n_run"], r["P"]),
                                           max_high=r["maxh"])
        got = dict(n_ops=st.n_ops, n_passes=st.n_passes, n_remaps=st.n_remaps, n_segments=st.n_segments,
                   fused_1q=st.fused_1q)
        assert got == r["stats"], r


FAMS = ["ghz", "qft", "rqc", "vqe", "grover"]


@pytest.mark.parametrize("n", [5, 7])
def test_compile_then_golden_model_vs_dense(n):
    circs = [zoo(n, s) for s in range(2)] + [C.make(f, n) for f in FAMS]
    for qc in circs:
        ref = dense.simulate(qc.tuples(), n)
        for P, k, fuse, tiling in itertools.product((1, 2, 4), (3, 4, 5), (True, False), (True, False)):
            L = n - (P.bit_length() - 1)
            if k > L:
                continue
            for init in (None, list(range(n))[::-1]):
                prog, fmap, st = K.compile_circuit(qc, L, k, fuse_1q=fuse, tiling=tiling, initial_map=init)
                out = model.run_program(prog, n, P, fmap, ctype=np.complex128)
                assert 1 - judge.fidelity(ref, out) < 1e-10, (qc.name, P, k, fuse, tiling, init)


@pytest.mark.parametrize("n", [6, 9])
def test_compile_gpu_streams_vs_dense(n):
    for qc in [zoo(n, 5), C.make("qft", n), C.make("rqc", n), C.u3_cx_layers(n, 6, 4), C.cp_ladder(n)]:
        ref = dense.simulate(qc.tuples(), n)
        for k in (4, 5, 6):
            # window scheduling cluster by cluster (forced: these circuits are below the size the
            # default clusters from) and op by op, and in-order scheduling
            for sched, cluster in (("window", K.GPU_CLUSTER), ("window", 0), ("inorder", 0)):
                prog, fmap, st = K.compile_gpu(qc, k=k, scheduler=sched, cluster=cluster, cluster_always=True)
                assert st.clustered == (sched == "window" and cluster > 0)
                out = model.run_program(prog, n, 1, fmap, ctype=np.complex128)
                assert 1 - judge.fidelity(ref, out) < 1e-10, (qc.name, k, sched, cluster)


@pytest.mark.parametrize("n", [6, 9])
def test_compile_fused_vs_dense(n):
    """Fused form (OP_DENSE blocks padded to the fusion width + leftover MUX2 ops) through the
    golden model against the gate-level simulator; records carry float32 matrices."""
    for qc in [zoo(n, 7), C.make("qft", n), C.make("rqc", n), C.u3_cx_layers(n, 8, 4)]:
        ref = dense.simulate(qc.tuples(), n)
        for width in (4, 5):
            for k in (max(width, 6), n):
                prog, fmap, st = K.compile_fused(qc, k=k, width=width)
                out = model.run_program(prog, n, 1, fmap, ctype=np.complex128)
                assert 1 - judge.fidelity(ref, out) < 1e-10, (qc.name, k, width)
                if qc.name.startswith("u3cx"):
                    assert st.n_dense > 0


def test_fused_form_reduces_passes_at_30q():
    """30 qubits, depth 20: fusing into 5-qubit dense blocks needs fewer sweeps than plain MUX2."""
    qc = C.workload_gpu30()
    _, _, plain = K.compile_gpu(qc)
    _, _, fused = K.compile_fused(qc, width=5)
    assert fused.n_dense > 0 and fused.n_passes < plain.n_passes


def test_encode_decode_roundtrip():
    rng = np.random.default_rng(0)
    qc = zoo(8, 9)
    prog, fmap, st = K.compile_circuit(qc, 8, 6)
    for kind, w in prog:
        if kind != "segment":
            continue
        L, k, passes = model.decode(w, np.complex64)
        assert (L, k) == (8, 6)
        for m, hi, ops in passes:
            assert len(hi) == m and all(k - m <= h < L for h in hi)
            for code, t, cm, M0, M1 in ops:
                assert code == 1 and 0 <= t < k and not (cm >> t) & 1
    del rng


def test_gpu_stream_high_record_roundtrip():
    qc = C.u3_cx_layers(16, 4, 3)
    prog, fmap, st = K.compile_gpu(qc, k=10, max_high=7)
    seen_ext = False
    for kind, w in prog:
        L, k, passes = model.decode(w)
        for m, hi, ops in passes:
            seen_ext |= m > 4
            assert len(set(hi)) == m and all(k - m <= h < L for h in hi)
    assert seen_ext


def test_transpositions_compose_to_perm():
    rng = np.random.default_rng(1)
    for n in (3, 6, 11):
        perm = [int(x) for x in rng.permutation(n)]
        p = list(range(n))
        for a, b in K.transpositions(perm):
            assert a < b
            p[a], p[b] = p[b], p[a]
        # transpositions(perm) lists swaps that, applied to perm, sort it: undo them to rebuild perm
        q = list(range(n))
        for a, b in reversed(K.transpositions(perm)):
            q[a], q[b] = q[b], q[a]
        assert q == perm


def test_workloads_are_seeded_and_shaped():
    a, b = C.workload_gpu30(), C.workload_gpu30()
    assert [g.params for g in a.data] == [g.params for g in b.data]
    assert len(a.data) == 20 * (30 + 15)
    for layer in range(20):
        cx = [g for g in a.data[layer * 45:(layer + 1) * 45] if g.name == "cx"]
        assert len(cx) == 15
        touched = sorted(q for g in cx for q in g.qubits)
        assert touched == list(range(30))
        for h in (26, 27, 28, 29):
            pair = [g for g in cx if h in g.qubits][0]
            other = [q for q in pair.qubits if q != h][0]
            assert other < 26
    c1 = C.workload_cpu20()
    assert len(c1.data) == 20 * 30 and c1.num_qubits == 20
    lad = C.cp_ladder(30)
    assert sum(g.name == "cp" for g in lad.data) == 435 and sum(g.name == "h" for g in lad.data) == 30
    st = K.Stats()
    ops, _ = K.circuit_to_ops(lad, st)
    assert len(ops) == 465
    assert math.isclose(lad.data[1].params[0], math.pi / 2)


def test_schedule_clusters_keeps_dense_blocks_whole():
    """Every block that schedule_clusters takes as a dense unit lands in one pass, its ops in
    program order; the passes hold every op once, and per-qubit op order is preserved."""
    qc = C.u3_cx_layers(16, 10, 5)
    ops, _ = K.circuit_to_ops(qc, K.Stats(), True, max_width=10)
    k, M = 10, 4
    passes = K.schedule_clusters(ops, 16, k, K.Stats(), M, 4)
    where = {id(op): (i, j) for i, p in enumerate(passes) for j, (op, _) in enumerate(p)}
    assert sorted(where) == sorted(id(op) for op in ops)
    dense_blocks = [b for b in K.fuse_ops(ops, 4)
                    if len(b.qubits) >= 2 and sum(K.op_cost(o) for o in b.ops) > K.dense_cost(4)]
    assert dense_blocks
    for b in dense_blocks:
        pos = [where[id(o)] for o in b.ops]
        assert len({i for i, _ in pos}) == 1 and pos == sorted(pos)
    for q in range(16):   # ops touching qubit q keep their program order
        seq = [where[id(op)] for op in ops if q in op.q]
        assert seq == sorted(seq)
    for p in passes:
        assert K.pass_window({x for _, ph in p for x in ph}, k, M) is not None


def test_fuse_ops_terminates_on_ops_wider_than_the_block():
    """A multi-controlled X on 6 qubits (Grover's diffusion) is wider than a 4-qubit block: it
    forms a block of its own and clustering still covers every op once, in dependency order."""
    qc = C.make("grover", 6)
    ops, _ = K.circuit_to_ops(qc, K.Stats(), True, max_width=13)
    assert max(len(op.q) for op in ops) > 4
    blocks = K.fuse_ops(ops, 4)
    flat = [id(o) for b in blocks for o in b.ops]
    assert sorted(flat) == sorted(id(o) for o in ops)
    for b in blocks:
        assert b.ops and (len(b.qubits) <= 4 or len(b.ops) == 1)
    for n in (6, 8):
        qc = C.make("grover", n)
        ref = dense.simulate(qc.tuples(), n)
        for prog, fmap, _ in (K.compile_gpu(qc, k=6), K.compile_fused(qc, k=6, width=4)):
            out = model.run_program(prog, n, 1, fmap, ctype=np.complex128)
            assert 1 - judge.fidelity(ref, out) < 1e-10


def test_compile_gpu_takes_clustered_schedule_by_pass_count():
    """From CLUSTER_MIN_QUBITS on, the default schedule is the clustered one exactly when it adds at
    most CLUSTER_MAX_EXTRA_PASSES of passes; below, passes are drawn op by op.  Results match the
    golden model run of the op-by-op stream either way."""
    seen = set()
    for qc in [C.u3_cx_layers(22, 4, 3), C.make("vqe", 22), C.make("rqc", 22), C.make("qft", 22),
               C.u3_cx_layers(21, 4, 3)]:
        n = qc.num_qubits
        k = 13
        M = K.GPU_MAX_HIGH
        ops, _ = K.circuit_to_ops(qc, K.Stats(), True, max_width=k)
        by_op = K.schedule_window(ops, n, k, K.Stats(), M)
        by_cl = K.schedule_clusters(ops, n, k, K.Stats(), M, K.GPU_CLUSTER)
        want = n >= K.CLUSTER_MIN_QUBITS and len(by_cl) <= len(by_op) * (1 + K.CLUSTER_MAX_EXTRA_PASSES)
        prog, fmap, st = K.compile_gpu(qc, k=k)
        assert st.clustered == want and st.n_passes == len(by_cl if want else by_op), qc.name
        seen.add(want)
    assert seen == {True, False}
