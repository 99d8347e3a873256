"""Qiskit-free circuit front-end: a gate list with the QuantumCircuit method names the reference's
generators use, the reference's benchmark families, and the seeded workloads the engine is
measured on.

Families restate circuits/generators.py (same PCG64 seeds and draw order, so instances and the
compiler's pass/remap counts are reproducible against the reference's recorded statistics).
"""
import math
from dataclasses import dataclass

import numpy as np

SEED_BASE = 20261005       # circuits/generators.py:10


@dataclass
class Instr:
    name: str
    qubits: tuple
    params: tuple = ()

    def __iter__(self):       # (name, qubits, params) tuple protocol
        return iter((self.name, self.qubits, self.params))


_SIMPLE1 = ("h", "x", "y", "z", "s", "sdg", "t", "tdg", "sx", "sxdg", "id")
_PAR1 = ("rx", "ry", "rz", "p", "u1")
_SIMPLE2 = ("cx", "cy", "cz", "ch", "swap", "iswap", "ecr", "dcx", "csx", "cs", "csdg")
_PAR2 = ("cp", "crx", "cry", "crz", "rxx", "ryy", "rzz", "rzx", "cu1")


class Circuit:
    """Minimal gate list: ``Circuit(n).h(0).cx(0, 1)``; ``data`` holds Instr records."""

    def __init__(self, n, name=""):
        self.num_qubits = int(n)
        self.name = name
        self.data = []

    def append(self, name, qubits, params=()):
        qs = tuple(int(q) for q in qubits)
        if len(set(qs)) != len(qs) or any(q < 0 or q >= self.num_qubits for q in qs):
            raise ValueError(f"bad qubits {qs} for {name} on {self.num_qubits} qubits")
        self.data.append(Instr(name, qs, tuple(float(p) for p in params)))
        return self

    def __getattr__(self, name):
        if name in _SIMPLE1:
            return lambda q: self.append(name, [q])
        if name in _PAR1:
            return lambda th, q: self.append(name, [q], [th])
        if name in _SIMPLE2:
            return lambda a, b: self.append(name, [a, b])
        if name in _PAR2:
            return lambda th, a, b: self.append(name, [a, b], [th])
        raise AttributeError(name)

    def u(self, th, ph, lam, q):
        return self.append("u", [q], [th, ph, lam])

    u3 = u

    def u2(self, ph, lam, q):
        return self.append("u2", [q], [ph, lam])

    def r(self, th, ph, q):
        return self.append("r", [q], [th, ph])

    def cu(self, th, ph, lam, gam, a, b):
        return self.append("cu", [a, b], [th, ph, lam, gam])

    def xx_plus_yy(self, th, beta, a, b):
        return self.append("xx_plus_yy", [a, b], [th, beta])

    def ccx(self, a, b, c):
        return self.append("ccx", [a, b, c])

    def ccz(self, a, b, c):
        return self.append("ccz", [a, b, c])

    def cswap(self, a, b, c):
        return self.append("cswap", [a, b, c])

    def mcx(self, ctrls, t):
        ctrls = list(ctrls)
        return self.append("mcx", ctrls + [t])

    def mcp(self, lam, ctrls, t):
        return self.append("mcp", list(ctrls) + [t], [lam])

    def barrier(self):
        return self

    def copy(self):
        c = Circuit(self.num_qubits, self.name)
        c.data = list(self.data)
        return c

    def padded(self, n_run):
        """Same gates on a wider register (the reference pads small circuits, host/sv/runtime.py:294-297)."""
        c = Circuit(n_run, self.name)
        c.data = list(self.data)
        return c

    def tuples(self):
        return [(g.name, g.qubits, g.params) for g in self.data]

    def __len__(self):
        return len(self.data)


# ------------------------------------------------------------------ reference families
def ghz(n):
    """circuits/generators.py:13-18"""
    qc = Circuit(n, f"ghz_{n}")
    qc.h(0)
    for i in range(n - 1):
        qc.cx(i, i + 1)
    return qc


def qft(n, swaps=True):
    """circuits/generators.py:21-30"""
    qc = Circuit(n, f"qft_{n}")
    for j in reversed(range(n)):
        qc.h(j)
        for k in reversed(range(j)):
            qc.cp(math.pi / 2 ** (j - k), k, j)
    if swaps:
        for i in range(n // 2):
            qc.swap(i, n - 1 - i)
    return qc


def rqc(n, depth=10, seed=None):
    """circuits/generators.py:33-49 -- brick-wall random circuit (same draws per qubit)."""
    rng = np.random.default_rng(SEED_BASE + 1000 * n + depth if seed is None else seed)
    qc = Circuit(n, f"rqc_{n}_d{depth}")
    for layer in range(depth):
        for q in range(n):
            g = int(rng.integers(5))
            th = float(rng.uniform(0, 2 * math.pi))
            if g == 0:
                qc.rx(th, q)
            elif g == 1:
                qc.ry(th, q)
            elif g == 2:
                qc.rz(th, q)
            elif g == 3:
                qc.t(q)
            else:
                qc.h(q)
        for q in range(layer % 2, n - 1, 2):
            qc.cz(q, q + 1)
    return qc


def vqe(n, layers=2, seed=None):
    """circuits/generators.py:52-65 -- hardware-efficient ansatz."""
    rng = np.random.default_rng(SEED_BASE + 7 * n + layers if seed is None else seed)
    qc = Circuit(n, f"vqe_{n}_L{layers}")
    for _ in range(layers):
        for q in range(n):
            qc.ry(float(rng.uniform(0, 2 * math.pi)), q)
            qc.rz(float(rng.uniform(0, 2 * math.pi)), q)
        for q in range(n - 1):
            qc.cx(q, q + 1)
    for q in range(n):
        qc.ry(float(rng.uniform(0, 2 * math.pi)), q)
        qc.rz(float(rng.uniform(0, 2 * math.pi)), q)
    return qc


def grover(n, marked=None, iters=None):
    """circuits/generators.py:68-82"""
    marked = (2 ** n - 1) if marked is None else marked
    iters = int(math.floor(math.pi / 4 * math.sqrt(2 ** n))) if iters is None else iters
    qc = Circuit(n, f"grover_{n}")
    for q in range(n):
        qc.h(q)
    for _ in range(iters):
        for q in range(n):
            if not (marked >> q) & 1:
                qc.x(q)
        qc.h(n - 1)
        qc.mcx(list(range(n - 1)), n - 1)
        qc.h(n - 1)
        for q in range(n):
            if not (marked >> q) & 1:
                qc.x(q)
        for q in range(n):
            qc.h(q)
        for q in range(n):
            qc.x(q)
        qc.h(n - 1)
        qc.mcx(list(range(n - 1)), n - 1)
        qc.h(n - 1)
        for q in range(n):
            qc.x(q)
        for q in range(n):
            qc.h(q)
    return qc


FAMILIES = {"ghz": ghz, "qft": qft, "rqc": rqc, "vqe": vqe, "grover": grover}


def make(family, n, **kw):
    return FAMILIES[family](n, **kw)


# ------------------------------------------------------------------ measured workloads
def u3_cx_layers(n, depth, seed, force_high=None):
    """Layer recipe of the measured workloads: per layer, U3(theta, phi, lambda) on every qubit with
    angles i.i.d. U[0, 2pi) (drawn as an (n, 3) block), then CX on each pair of a seeded random
    matching (rng.permutation(n) paired consecutively, first element = control; an odd n leaves
    the last qubit unpaired).

    force_high: qubits that must each be paired with a qubit outside the set (the 30-qubit GPU
    workload pairs each of 26..29 with a distinct partner from 0..25, the feasible maximum of
    "CX pairs involving 26..29": 4 of 15 per layer).  Those pairs are drawn first (partner by
    rng.choice without replacement, orientation by rng.integers(2)), the rest as above.
    """
    rng = np.random.default_rng(seed)
    qc = Circuit(n, f"u3cx_{n}_d{depth}_s{seed}")
    for _ in range(depth):
        ang = rng.uniform(0.0, 2 * math.pi, size=(n, 3))
        for q in range(n):
            qc.u(float(ang[q, 0]), float(ang[q, 1]), float(ang[q, 2]), q)
        if force_high:
            hi = list(force_high)
            lo = [q for q in range(n) if q not in hi]
            partners = [lo[int(i)] for i in rng.choice(len(lo), size=len(hi), replace=False)]
            for h, p in zip(hi, partners):
                if int(rng.integers(2)):
                    qc.cx(h, p)
                else:
                    qc.cx(p, h)
            rest = [q for q in lo if q not in partners]
            order = [rest[int(i)] for i in rng.permutation(len(rest))]
        else:
            order = [int(i) for i in rng.permutation(n)]
        for i in range(0, len(order) - 1, 2):
            qc.cx(order[i], order[i + 1])
    return qc


def workload_cpu20():
    """20 qubits, depth 20, seed 0: the CPU golden-model workload."""
    return u3_cx_layers(20, 20, 0)


def workload_gpu30():
    """30 qubits, depth 20, seed 1; each of qubits 26..29 is paired with one of 0..25 every layer."""
    return u3_cx_layers(30, 20, 1, force_high=(26, 27, 28, 29))


def workload_fused(n):
    """The fused-form workload: n qubits (32 or 34), depth 40, seed 2, same layer recipe."""
    return u3_cx_layers(n, 40, 2)


def cp_ladder(n):
    """Controlled-phase ladder: for j = 0..n-1, H(j) then CP(pi / 2^(k-j)) between control k and
    target j for every k > j (n H + n(n-1)/2 CP).  The final bit reversal is a separate
    permutation (perm[i] = n-1-i)."""
    qc = Circuit(n, f"cpladder_{n}")
    for j in range(n):
        qc.h(j)
        for k in range(j + 1, n):
            qc.cp(math.pi / 2 ** (k - j), k, j)
    return qc


def product_state_angles(n, seed):
    """Per-qubit (theta ~ U[0, pi], phi ~ U[0, 2pi)); qubit q = cos(t/2)|0> + e^{i phi} sin(t/2)|1>."""
    rng = np.random.default_rng(seed)
    th = rng.uniform(0.0, math.pi, size=n)
    ph = rng.uniform(0.0, 2 * math.pi, size=n)
    return th, ph
