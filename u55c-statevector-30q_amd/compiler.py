"""Host compiler: circuit -> MUX2 ops -> passes -> QSV2 segments.

Restates the reference compiler (host/sv/compiler.py) without Qiskit: gate matrices come from
the closed forms in ``gates``; everything from gate classification to the pass scheduler and the
encoder follows the reference's rules, so the same circuit yields the same op list, the same
passes/remaps (pinned against the reference's recorded statistics) and a stream the reference
golden model and this project's HIP library execute identically.

Program = list of ("segment", int32 QSV2 words) and ("remap", perm) with perm[new_phys] = old_phys.

MI355X retarget (``target="mi355x"``): one slice (P = 1, L = n), a 2^13-amplitude LDS tile and
up to 7 gathered high qubits per pass (sub-blocks of >= 64 contiguous amplitudes), so a pass can
hold any op of <= 7 qubits above the 6 always-local ones and remaps are only needed for wider
ops.  ``scheduler="window"`` additionally reorders commuting ops (disjoint qubits) to fill each
pass's window before closing it.
"""
from dataclasses import dataclass, field

import numpy as np
from scipy.linalg import cossin, sqrtm

from . import gates
from .isa import (MAGIC, OP_MUX2, MAX_OPS_PER_PASS, MAX_HIGH,
                  GPU_K, GPU_MAX_HIGH, GPU_MAX_OPS, GPU_CLUSTER)

SKIP = {"measure", "barrier", "delay", "save_statevector", "id"}       # host/sv/compiler.py:20
I2 = np.eye(2, dtype=complex)
_X = gates.X
_Z = gates.Z


def _mc_matrix(name, params):
    """Multi-controlled gates built directly as MUX2 (host/sv/compiler.py:25-29)."""
    if name in ("ccx", "c3x", "c4x", "mcx", "mcx_gray", "mcx_recursive"):
        return _X
    if name == "ccz":
        return _Z
    if name in ("mcphase", "mcp"):
        return np.diag([1, np.exp(1j * float(params[0]))])
    return None


@dataclass
class Op:
    """One MUX2: target t, controls ctrls (all 1 selects M1), in storage (logical) qubits."""
    t: int
    ctrls: tuple
    M0: np.ndarray
    M1: np.ndarray
    src: object = None

    @property
    def q(self):
        return (self.t,) + tuple(self.ctrls)


@dataclass
class Stats:
    """host/sv/compiler.py:45-56"""
    n_input_gates: int = 0
    native: dict = field(default_factory=dict)
    relabel: int = 0
    lowered: dict = field(default_factory=dict)
    csd_2q: int = 0
    fused_1q: int = 0
    n_ops: int = 0
    n_passes: int = 0
    n_remaps: int = 0
    n_segments: int = 0
    clustered: bool = False          # MI355X compiler: passes drawn cluster by cluster (compile_gpu)


def _count(d, key):
    d[key] = d.get(key, 0) + 1


def controlled_target(M, nq):
    """host/sv/compiler.py:63-75 -- (j, U) if M is a 2x2 U on qubit j controlled by all others."""
    dim = 1 << nq
    full = dim - 1
    for j in range(nq):
        ones = full & ~(1 << j)
        sel = [ones, ones | (1 << j)]
        rest = M.copy()
        Ub = M[np.ix_(sel, sel)].copy()
        rest[np.ix_(sel, sel)] = I2
        if np.allclose(rest, np.eye(dim), atol=1e-12):
            return j, Ub
    return None


def ops_2q(M, a, b, stats):
    """host/sv/compiler.py:78-92 -- 4x4 (index bit_a + 2 bit_b) -> 1 MUX2 (diagonal / controlled)
    or 3 MUX2 (cosine-sine decomposition over bit b)."""
    if np.allclose(M, np.diag(np.diag(M)), atol=1e-12):
        d = np.diag(M)
        return [Op(b, (a,), np.diag([d[0], d[2]]), np.diag([d[1], d[3]]))]
    ct = controlled_target(M, 2)
    if ct is not None:
        j, Ub = ct
        t, c = (a, b) if j == 0 else (b, a)
        return [Op(t, (c,), I2, Ub)]
    (u1, u2), theta, (v1h, v2h) = cossin(M, p=2, q=2, separate=True)
    cs, sn = np.cos(theta), np.sin(theta)
    R0 = np.array([[cs[0], -sn[0]], [sn[0], cs[0]]])
    R1 = np.array([[cs[1], -sn[1]], [sn[1], cs[1]]])
    stats.csd_2q += 1
    return [Op(a, (b,), v1h, v2h), Op(b, (a,), R0, R1), Op(a, (b,), u1, u2)]


def _lower(name, qs, params):
    """Decompositions of gates without a MUX2 form (the reference asks Qiskit's transpiler,
    host/sv/compiler.py:159-166).  CSWAP = CX(c,b) CCX(a,b,c) CX(c,b)."""
    if name == "cswap":
        a, b, c = qs
        return [("cx", (c, b), ()), ("ccx", (a, b, c), ()), ("cx", (c, b), ())]
    raise RuntimeError(f"gate {name} on {len(qs)} qubits has no MUX2 form and no lowering here")


def _mc_ops(Um, ctrls, t, width):
    """C^c(Um) as MUX2 ops of at most `width` qubits (Barenco et al. 1995, Lemma 7.5):
    C(V)[c_last -> t] C^(c-1)X[rest -> c_last] C(V^dag)[c_last -> t] C^(c-1)X[rest -> c_last]
    C^(c-1)(V)[rest -> t] with V^2 = Um, recursing until every op fits."""
    ctrls = list(ctrls)
    if len(ctrls) + 1 <= width or len(ctrls) <= 1:
        return [Op(t, tuple(ctrls), I2, np.asarray(Um, dtype=complex))]
    V = sqrtm(np.asarray(Um, dtype=complex))
    Vd = V.conj().T
    rest, last = ctrls[:-1], ctrls[-1]
    return (_mc_ops(V, [last], t, width) + _mc_ops(_X, rest, last, width)
            + _mc_ops(Vd, [last], t, width) + _mc_ops(_X, rest, last, width)
            + _mc_ops(V, rest, t, width))


def circuit_to_ops(qc, stats, fuse_1q=True, max_width=None):
    """host/sv/compiler.py:95-173 -- returns (ops, relab); relab[l] = storage qubit of logical l."""
    n = qc.num_qubits
    relab = list(range(n))
    ops = []
    pending = {}

    def flush(q):
        if q in pending:
            Um = pending.pop(q)
            ops.append(Op(q, (), Um, Um))

    def emit(lst):
        for op in lst:
            for q in op.q:
                flush(q)
            ops.append(op)

    def lower(name, qs, params):
        stats.native[name] -= 1
        _count(stats.lowered, name)
        mc = _mc_matrix(name, params)
        if mc is not None:
            st = [relab[q] for q in qs]
            emit(_mc_ops(mc, st[:-1], st[-1], max(2, max_width or 2)))
            return
        for sub in _lower(name, qs, params):
            handle(sub[0], list(sub[1]), sub[2], lowered=True)

    def handle(name, qs, params, lowered=False):
        if name in SKIP:
            return
        if not lowered:
            stats.n_input_gates += 1
        if name == "swap":
            a, b = qs
            relab[a], relab[b] = relab[b], relab[a]
            stats.relabel += 1
            return
        st = [relab[q] for q in qs]
        mc = _mc_matrix(name, params) if len(qs) > 2 else None
        if mc is not None and (max_width is None or len(qs) <= max_width):
            if not lowered:
                _count(stats.native, name)
            emit([Op(st[-1], tuple(st[:-1]), I2, mc, src=(name, tuple(qs)))])
            return
        if len(qs) > 2 and ((max_width is not None and len(qs) > max_width) or len(qs) > 10):
            if not lowered:
                _count(stats.native, name)
            lower(name, qs, params)
            return
        M = gates.matrix(name, len(qs), params)
        if not lowered:
            _count(stats.native, name)
        if len(qs) == 1:
            s = st[0]
            if fuse_1q and s in pending:
                pending[s] = M @ pending[s]
                stats.fused_1q += 1
            elif fuse_1q:
                pending[s] = M
            else:
                ops.append(Op(s, (), M, M))
            return
        if len(qs) == 2:
            emit(ops_2q(M, st[0], st[1], stats))
            return
        ct = controlled_target(M, len(qs))
        if ct is not None:
            j, Ub = ct
            emit([Op(st[j], tuple(x for i, x in enumerate(st) if i != j), I2, Ub, src=(name, tuple(qs)))])
            return
        lower(name, qs, params)

    for ins in qc.data:
        name, qs, params = ins
        handle(name, list(qs), params)
    for q in list(pending):
        flush(q)
    stats.n_ops = len(ops)
    return ops, relab


def pass_window(phys, k, M=None):
    """host/sv/compiler.py:176-183 -- smallest m <= M with |{q >= k-m}| <= m, and those qubits."""
    M = MAX_HIGH if M is None else M
    for m in range(M + 1):
        hi = sorted(q for q in phys if q >= k - m)
        if len(hi) <= m:
            return m, hi
    return None


def schedule(ops, n, L, k, stats, remap_policy="lru", tiling=True, l2p0=None, M=MAX_HIGH,
             max_ops=MAX_OPS_PER_PASS):
    """host/sv/compiler.py:186-260 -- greedy in-order pass packing with remaps for bank qubits
    (physical >= L) or ops that no window holds; evicts the qubit used furthest ahead."""
    l2p = list(l2p0) if l2p0 is not None else list(range(n))
    steps = []
    seg = []
    cur, cur_set = [], set()
    k = min(k, L)

    def close_pass():
        nonlocal cur, cur_set
        if cur:
            seg.append(cur)
        cur, cur_set = [], set()

    def close_segment():
        close_pass()
        if seg:
            steps.append(("segment", list(seg)))
            seg.clear()

    def next_use(p, idx):
        logical = l2p.index(p)
        for j in range(idx, len(ops)):
            if logical in ops[j].q:
                return j - idx
        return 10 ** 9

    for idx, op in enumerate(ops):
        phys = [l2p[q] for q in op.q]
        if any(p >= L for p in phys) or pass_window(set(phys), k, M) is None:
            low = k - M
            assert len(phys) <= low + M, "op wider than the tile window"
            far = lambda p: -next_use(p, idx)
            bank = [p for p in phys if p >= L]
            high_local = [p for p in phys if low <= p < L]
            hi_slots = sorted((p for p in range(max(k, L - 6), L) if p not in phys), key=far)
            if bank and len(high_local) + len(bank) <= M and len(hi_slots) >= len(bank):
                move, victims = bank, hi_slots[:len(bank)]
            else:
                keep = high_local[:M]
                move = [p for p in phys if p >= low and p not in keep]
                low_slots = sorted((p for p in range(low) if p not in phys), key=far)
                high_slots = sorted((p for p in range(low, L) if p not in phys), key=far)
                high_free = M - len(keep)
                victims = []
                for _ in move:
                    if low_slots:
                        victims.append(low_slots.pop(0))
                    elif high_free > 0 and high_slots:
                        victims.append(high_slots.pop(0))
                        high_free -= 1
                    else:
                        raise RuntimeError("cannot fit op into tile window")
            close_segment()
            perm = list(range(n))
            for a, b in zip(move, victims):
                perm[a], perm[b] = b, a
            steps.append(("remap", perm))
            stats.n_remaps += 1
            where = {old: new for new, old in enumerate(perm)}
            l2p = [where[p] for p in l2p]
            phys = [l2p[q] for q in op.q]
        merged = cur_set | set(phys)
        if (not tiling and cur) or len(cur) >= max_ops or pass_window(merged, k, M) is None:
            close_pass()
            merged = set(phys)
        cur.append((op, phys))
        cur_set = merged
    close_segment()
    return steps, l2p


def _is_diagonal(op):
    """True when the op is diagonal in the computational basis (a MUX2 whose M0 and M1 are both
    diagonal); diagonal ops commute with each other even when they share qubits."""
    if not hasattr(op, "M0"):
        return False
    M0, M1 = np.asarray(op.M0), np.asarray(op.M1)
    return M0[0, 1] == 0 and M0[1, 0] == 0 and M1[0, 1] == 0 and M1[1, 0] == 0


def _dependency_graph(ops):
    """Edges i -> j (i < j) when ops i and j share a qubit and do not both act diagonally."""
    nq = len(ops)
    last_nd = {}           # qubit -> last non-diagonal op
    diag_since = {}        # qubit -> diagonal ops after it
    preds = [0] * nq
    succs = [[] for _ in range(nq)]
    for i, op in enumerate(ops):
        deps = set()
        diag = _is_diagonal(op)
        for q in op.q:
            if q in last_nd:
                deps.add(last_nd[q])
            if not diag:
                deps.update(diag_since.get(q, ()))
        for q in op.q:
            if diag:
                diag_since.setdefault(q, []).append(i)
            else:
                last_nd[q] = i
                diag_since[q] = []
        deps.discard(i)
        preds[i] = len(deps)
        for d in deps:
            succs[d].append(i)
    return preds, succs


def schedule_window(ops, n, k, stats, M, max_ops=GPU_MAX_OPS, max_cost=None):
    """MI355X pass packer for one slice (L = n): list scheduling over the op dependency DAG.

    An op is ready when every earlier op sharing a qubit has been placed.  A pass keeps taking the
    earliest ready op whose qubits keep the pass inside one tile window (pass_window over the
    union) until none fits, max_ops is reached or the estimated VALU cost exceeds max_cost; ops on
    disjoint qubits commute, so the state is unchanged by the reordering.  Ops wider than the
    window are not handled here (the caller falls back to the reference scheduler).
    """
    nq = len(ops)
    preds, succs = _dependency_graph(ops)
    ready = sorted(i for i in range(nq) if preds[i] == 0)
    passes = []
    done = 0
    import heapq
    heapq.heapify(ready)
    while done < nq:
        cur, cur_set, cost = [], set(), 0.0
        skipped = []
        while ready:
            i = heapq.heappop(ready)
            op = ops[i]
            merged = cur_set | set(op.q)
            c = op_cost(op)
            if (len(cur) >= max_ops or pass_window(merged, k, M) is None
                    or (max_cost is not None and cur and cost + c > max_cost)):
                skipped.append(i)
                continue
            cur.append((op, list(op.q)))
            cur_set = merged
            cost += c
            done += 1
            for s in succs[i]:
                preds[s] -= 1
                if preds[s] == 0:
                    heapq.heappush(ready, s)
        for i in skipped:
            heapq.heappush(ready, i)
        if not cur:
            raise RuntimeError("cannot fit op into tile window")
        passes.append(cur)
    return passes


def _is_identity(Mx):
    return bool(np.array_equal(np.asarray(Mx, dtype=np.complex64), np.eye(2, dtype=np.complex64)))


def op_cost(op):
    """Rough VALU cost of one MUX2 per amplitude (a general 2x2 = 8 FMA per amplitude); a fused
    block carries its own cost.  An op's cost is kept on the op after the first call (ops are not
    changed once built)."""
    if hasattr(op, "cost"):
        return op.cost
    c = op.__dict__.get("_cost")
    if c is None:
        c = op.__dict__["_cost"] = _mux_cost(op)
    return c


def _mux_cost(op):
    M1 = np.asarray(op.M1)
    if op.ctrls and _is_identity(op.M0):
        if np.count_nonzero(M1 - np.diag(np.diag(M1))) == 0:
            return 1.0
        return 2.0
    if np.count_nonzero(M1 - np.diag(np.diag(M1))) == 0 and np.count_nonzero(op.M0 - np.diag(np.diag(op.M0))) == 0:
        return 4.0
    return 8.0


def encode_segment(passes, L, k, M=MAX_HIGH):
    """host/sv/compiler.py:263-285 -- QSV2 words for one segment."""
    words = [MAGIC, L, len(passes), k, 0, 0, 0, 0]
    for p in passes:
        phys_set = {q for _, ph in p for q in ph}
        m, hi = pass_window(phys_set, k, M)
        h = list(hi) + [0] * (4 - len(hi)) if len(hi) <= 4 else list(hi)
        tmap = {q: q for q in phys_set if q < k - m}
        for i, q in enumerate(hi):
            tmap[q] = k - m + i
        if m > 4:
            raise ValueError("QSV2 pass headers hold at most 4 high qubits; use encode_segment_gpu")
        words += [m, h[0], h[1], len(p), h[2], h[3], 0, 0]
        for op, ph in p:
            words += _op_words(op, [tmap[q] for q in ph])
    arr = np.asarray(words, dtype=np.int64)
    return (arr & 0xFFFFFFFF).astype(np.uint32).view(np.int32)


def _op_words(op, tq):
    cmask = 0
    for c in tq[1:]:
        cmask |= 1 << c
    fl = np.zeros(16, dtype=np.float32)
    for i, Mat in enumerate((op.M0, op.M1)):
        f = np.asarray(Mat, dtype=np.complex64).reshape(-1)
        fl[8 * i:8 * i + 8:2] = f.real
        fl[8 * i + 1:8 * i + 8:2] = f.imag
    return [OP_MUX2, tq[0], cmask, 0, 0, 0, 0, 0] + [int(x) for x in fl.view(np.int32)]


# Pass headers of the retargeted stream: m up to GPU_MAX_HIGH (> 4) high qubits do not fit the
# reference's 8-word pass header, so a pass with m > 4 carries the extra high qubits in a
# continuation record placed right after the header: [OP_HIGHX=0x48, h4, h5, h6, h7, h8, h9, h10]
# (an 8-word record; the reference format never has m > 4, so reference streams are unchanged).
OP_HIGHX = 0x48


def encode_segment_gpu(passes, L, k, M):
    words = [MAGIC, L, len(passes), k, 0, 0, 0, 0]
    for p in passes:
        phys_set = {q for _, ph in p for q in ph}
        m, hi = pass_window(phys_set, k, M)
        tmap = {q: q for q in phys_set if q < k - m}
        for i, q in enumerate(hi):
            tmap[q] = k - m + i
        h = list(hi) + [0] * (11 - len(hi))
        words += [m, h[0], h[1], len(p), h[2], h[3], 0, 0]
        if m > 4:
            words += [OP_HIGHX] + h[4:11]
        for op, ph in p:
            words += _op_words(op, [tmap[q] for q in ph])
    arr = np.asarray(words, dtype=np.int64)
    return (arr & 0xFFFFFFFF).astype(np.uint32).view(np.int32)


def compile_circuit(qc, L, k, fuse_1q=True, tiling=True, remap_policy="lru", initial_map=None,
                    max_high=MAX_HIGH):
    """host/sv/compiler.py:288-314 -- (prog, final_map, stats); final_map[l] = physical of logical l."""
    k = min(k, L)
    max_high = max(0, min(max_high, k - 3))
    stats = Stats()
    ops, relab = circuit_to_ops(qc, stats, fuse_1q, max_width=k)
    n = qc.num_qubits
    steps, l2p_end = schedule(ops, n, L, k, stats, remap_policy, tiling, initial_map, max_high)
    prog = []
    if initial_map is not None and list(initial_map) != list(range(n)):
        perm = [0] * n
        for l, p in enumerate(initial_map):
            perm[p] = l
        prog.append(("remap", perm))
    for kind, payload in steps:
        if kind == "segment":
            prog.append(("segment", encode_segment(payload, L, k, max_high)))
            stats.n_segments += 1
            stats.n_passes += len(payload)
        else:
            prog.append(("remap", payload))
    final_map = [l2p_end[relab[l]] for l in range(n)]
    return prog, final_map, stats


# The clustered schedule may add at most this share of passes, and is drawn only from this many
# qubits on.  Measured on MI355X (GHZ, QFT, RQC and VQE at 20, 26 and 30 qubits, U3+CX at 30 and 32,
# library fusion on): every case with the same or fewer passes ran as fast or faster clustered (U3+CX
# 30: 0.93x, RQC 20 / 26: 0.83x / 0.91x), every case with 15-25% more passes ran slower (VQE 20 / 26,
# RQC 30: 1.05-1.15x), U3+CX 32 with 2% more passes ran 0.94x.  Below 22 qubits a pass takes ~10 us and
# the host time of clustering (1.4 ms for the 287 ops of RQC 20, 7 ms for the 900 of U3+CX 30) exceeds
# what it saves (RQC 20: 40 us per run).
CLUSTER_MAX_EXTRA_PASSES = 0.05
CLUSTER_MIN_QUBITS = 22


def schedule_clusters(ops, n, k, stats, M, width, max_ops=GPU_MAX_OPS, max_cost=None):
    """Window scheduling over clusters: `fuse_ops` groups the ops into blocks of <= width qubits; a
    block worth one dense width-qubit unit (MUX2 cost above dense_cost(width), the rule of
    `compile_fused`) whose qubits fit one tile window is scheduled as a single unit, the other ops
    one by one; each pass then lists the member ops of its blocks in order.  Returns passes of
    (op, physical qubits) like schedule_window."""
    units = []
    for b in fuse_ops(ops, width):
        if (2 <= len(b.qubits) <= width and sum(op_cost(op) for op in b.ops) > dense_cost(width)
                and pass_window(set(b.qubits), k, M) is not None):
            b.cost = dense_cost(width)
            units.append(b)
        else:
            units.extend(b.ops)
    return [[(o, list(o.q)) for u, _ in p for o in (u.ops if isinstance(u, Block) else [u])]
            for p in schedule_window(units, n, k, stats, M, max_ops, max_cost)]


def compile_gpu(qc, k=GPU_K, max_high=None, fuse_1q=True, scheduler="window", max_ops=GPU_MAX_OPS,
                max_cost=None, cluster=GPU_CLUSTER, cluster_always=False):
    """MI355X retarget of compile_circuit: P = 1, L = n, LDS-width tile k, up to max_high gathered
    high qubits.  Returns (prog, final_map, stats) like compile_circuit.

    cluster = w > 0 (window scheduler, n >= CLUSTER_MIN_QUBITS): the passes are drawn twice, op by op
    (`schedule_window`) and cluster by cluster (`schedule_clusters`: no pass boundary cuts a cluster
    worth a dense w-qubit unit; its member ops are emitted in order as ordinary MUX2 records), and the
    clustered schedule is taken when it adds at most CLUSTER_MAX_EXTRA_PASSES of passes
    (stats.clustered says which).  The stream stays
    plain MUX2; with the clustered schedule the library's pass fusion finds whole clusters inside
    single passes (30-qubit U3+CX depth 20: 178 ms per step against 192 ms op by op).  cluster = 0
    schedules op by op, which suits a stream run with library fusion off; cluster_always takes the
    clustered schedule at any size and pass count."""
    n = qc.num_qubits
    k = min(k, n)
    if max_high is None:
        max_high = max(0, min(GPU_MAX_HIGH, k - 6)) if k >= 8 else max(0, k - 1)
    max_high = max(0, min(max_high, k - 1))
    stats = Stats()
    ops, relab = circuit_to_ops(qc, stats, fuse_1q, max_width=k)
    wide = any(pass_window(set(op.q), k, max_high) is None for op in ops)
    prog = []
    if scheduler == "window" and not wide:
        passes = schedule_window(ops, n, k, stats, max_high, max_ops, max_cost)
        if cluster and passes and (n >= CLUSTER_MIN_QUBITS or cluster_always):
            by_cluster = schedule_clusters(ops, n, k, stats, max_high, cluster, max_ops, max_cost)
            if cluster_always or len(by_cluster) <= len(passes) * (1 + CLUSTER_MAX_EXTRA_PASSES):
                passes = by_cluster
                stats.clustered = True
        if passes:
            prog.append(("segment", encode_segment_gpu(passes, n, k, max_high)))
            stats.n_segments = 1
            stats.n_passes = len(passes)
        l2p_end = list(range(n))
    else:
        steps, l2p_end = schedule(ops, n, n, k, stats, "lru", True, None, max_high, max_ops)
        for kind, payload in steps:
            if kind == "segment":
                prog.append(("segment", encode_segment_gpu(payload, n, k, max_high)))
                stats.n_segments += 1
                stats.n_passes += len(payload)
            else:
                prog.append(("remap", payload))
    final_map = [l2p_end[relab[l]] for l in range(n)]
    return prog, final_map, stats


def transpositions(perm):
    """host/sv/runtime.py:35-44 -- perm[new] = old as a list of bit transpositions (a < b)."""
    p = list(perm)
    out = []
    for i in range(len(p)):
        while p[i] != i:
            j = p[i]
            out.append((min(i, j), max(i, j)))
            p[i], p[j] = p[j], p[i]
    return out


def split_passes(words):
    """One single-pass segment per pass of a QSV2 segment (same header, n_passes = 1)."""
    w = np.asarray(words, dtype=np.int32)
    if int(w[0]) & 0xFFFFFFFF != MAGIC:
        raise AssertionError("bad QSV2 magic")
    out = []
    pos = 8
    for _ in range(int(w[2])):
        m, nops = int(w[pos]), int(w[pos + 3])
        length = 8 + (8 if m > 4 else 0)
        for _ in range(nops):
            code = int(w[pos + length])
            length += 24 if code == OP_MUX2 else 8 + (2 << (2 * int(w[pos + length + 1])))
        seg = np.concatenate([w[:8], w[pos:pos + length]]).astype(np.int32)
        seg[2] = 1
        out.append(seg)
        pos += length
    return out


# ------------------------------------------------------------------ fused dense-unitary form
# Record of this project's fused form (the reference stream has OP_MUX2 only, host/sv/isa.py:30):
#   [OP_DENSE, f, q0, q1, q2, q3, q4, 0] followed by 2 * 4^f float32 words: the 2^f x 2^f complex64
#   matrix U row-major as (re, im), basis index sum_i bit(q_i) << i over the listed tile qubits.
OP_DENSE = 2
DENSE_MAX_F = 5


def dense_words(Um, tq):
    f = len(tq)
    U = np.asarray(Um, dtype=np.complex64).reshape(1 << f, 1 << f)
    q = list(tq) + [0] * (5 - f)
    fl = np.empty(2 << (2 * f), dtype=np.float32)
    fl[0::2] = U.real.ravel()
    fl[1::2] = U.imag.ravel()
    return [OP_DENSE, f] + q + [0] + [int(x) for x in fl.view(np.int32)]


def op_dense(op, qubits):
    """Dense matrix of one MUX2 op on the ordered qubit list (index sum_i bit(qubits[i]) << i)."""
    f = len(qubits)
    pos = {q: i for i, q in enumerate(qubits)}
    bt = 1 << pos[op.t]
    cm = 0
    for c in op.ctrls:
        cm |= 1 << pos[c]
    O = np.zeros((1 << f, 1 << f), dtype=np.complex128)
    for i0 in range(1 << f):
        if i0 & bt:
            continue
        M = op.M1 if (cm and (i0 & cm) == cm) else op.M0
        i1 = i0 | bt
        O[i0, i0], O[i0, i1] = M[0, 0], M[0, 1]
        O[i1, i0], O[i1, i1] = M[1, 0], M[1, 1]
    return O


class Block:
    """A fused cluster: ordered qubits, its ops, and whether it is emitted as one dense op."""

    def __init__(self, qubits, ops):
        self.qubits = list(qubits)
        self.ops = ops
        self.dense = False

    @property
    def q(self):
        return tuple(self.qubits)

    def matrix(self):
        U = np.eye(1 << len(self.qubits), dtype=np.complex128)
        for op in self.ops:
            U = op_dense(op, self.qubits) @ U
        return U


def fuse_ops(ops, width):
    """Greedy DAG clustering: a block grows by the ready op that adds the fewest new qubits
    (earliest first on ties) while its qubit set stays <= width.  An op wider than width forms a
    block of its own."""
    nq = len(ops)
    last = {}
    preds = [0] * nq
    succs = [[] for _ in range(nq)]
    for i, op in enumerate(ops):
        deps = set()
        for q in op.q:
            if q in last:
                deps.add(last[q])
            last[q] = i
        preds[i] = len(deps)
        for d in deps:
            succs[d].append(i)
    mask = [sum(1 << q for q in set(op.q)) for op in ops]
    ready = {i for i in range(nq) if preds[i] == 0}
    blocks = []
    while ready:
        S, ns = 0, 0          # the block's qubits as a bit mask, and their count
        members = []
        while True:
            best, bcost = None, None
            for i in ready:
                add = (mask[i] & ~S).bit_count()
                if ns + add > width and ns:
                    continue
                key = (add, i)
                if bcost is None or key < bcost:
                    best, bcost = i, key
            if best is None:
                break
            ready.discard(best)
            members.append(best)
            S |= mask[best]
            ns = S.bit_count()
            for s in succs[best]:
                preds[s] -= 1
                if preds[s] == 0:
                    ready.add(s)
            if ns > width:
                break
        blocks.append(Block([q for q in range(S.bit_length()) if S >> q & 1], [ops[i] for i in members]))
    return blocks


def dense_cost(f):
    """Matrix-core cost of one dense f-qubit op in VALU-FMA units per amplitude: 8 * 2^f real flops
    per amplitude at twice the non-packed FP32 VALU rate."""
    return 8.0 * (1 << f) / 2.0 / 2.0


def compile_fused(qc, k=GPU_K, width=4, max_high=None, fuse_1q=True, max_ops=GPU_MAX_OPS, max_cost=None):
    """Fused dense-unitary form: ops are clustered into blocks of <= width qubits; a block whose
    MUX2 cost exceeds the dense cost of its padded width becomes one OP_DENSE record (padded to
    `width` qubits with idle tile qubits), the rest stay MUX2.  Blocks are packed into passes by the
    same window rule as MUX2 ops.  Returns (prog, final_map, stats) with stats.n_dense added."""
    n = qc.num_qubits
    k = min(k, n)
    if max_high is None:
        max_high = max(0, min(GPU_MAX_HIGH, k - 6)) if k >= 8 else max(0, k - 1)
    stats = Stats()
    ops, relab = circuit_to_ops(qc, stats, fuse_1q, max_width=k)
    blocks = fuse_ops(ops, width)
    units = []
    for b in blocks:
        mux_cost = sum(op_cost(op) for op in b.ops)
        if 2 <= len(b.qubits) <= width and mux_cost > dense_cost(width):
            b.dense = True
            b.cost = dense_cost(width)
            units.append(b)
        else:
            for op in b.ops:
                units.append(op)
    # pass packing over units (blocks and plain ops) with the window scheduler
    passes = schedule_window(units, n, k, stats, max_high, max_ops, max_cost)
    words = [MAGIC, n, len(passes), k, 0, 0, 0, 0]
    n_dense = 0
    for p in passes:
        phys_set = {q for _, ph in p for q in ph}
        m, hi = pass_window(phys_set, k, max_high)
        tmap = {q: q for q in phys_set if q < k - m}
        for i, q in enumerate(hi):
            tmap[q] = k - m + i
        h = list(hi) + [0] * (11 - len(hi))
        words += [m, h[0], h[1], len(p), h[2], h[3], 0, 0]
        if m > 4:
            words += [OP_HIGHX] + h[4:11]
        used_tile = set(tmap.values())
        for u, ph in p:
            if isinstance(u, Block):
                tq = [tmap[q] for q in u.qubits]
                U = u.matrix()
                pad = [t for t in range(k) if t not in tq and t not in used_tile][:max(0, width - len(tq))]
                if len(pad) < width - len(tq):
                    pad += [t for t in range(k) if t not in tq and t not in pad][:width - len(tq) - len(pad)]
                for _ in pad:
                    U = np.kron(np.eye(2), U)
                words += dense_words(U, tq + pad)
                n_dense += 1
            else:
                words += _op_words(u, [tmap[q] for q in ph])
    arr = np.asarray(words, dtype=np.int64)
    seg = (arr & 0xFFFFFFFF).astype(np.uint32).view(np.int32)
    stats.n_segments = 1
    stats.n_passes = len(passes)
    stats.n_dense = n_dense
    final_map = [relab[l] for l in range(n)]
    return [("segment", seg)], final_map, stats
