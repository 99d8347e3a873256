"""Gate-level complex128 statevector simulator -- an independent reference for the MUX2 path.

Gate matrices are derived here from their mathematical definitions (Pauli exponentials via
scipy.linalg.expm, controlled embeddings, products), not from closed-form tables, so the product's
closed-form gate table (the package's gates module) is checked against a different derivation.
Qubit order is little-endian (index bit q = qubit q; multi-qubit gate matrices index
sum_i bit(qubits[i]) << i), the convention of the reference (host/sv/model.py:76,
host/sv/runtime.py:76-78) and of Qiskit.

Pinning: the 1q/2q definitions used by the benchmark configs (U3/U, H, X, CX, CZ, CP, RX/RY/RZ,
T, S) are fixed by textbook identities checked in tests/test_oracle.py (H Z H = X, S^2 = Z,
T^2 = S, U3(pi/2, 0, pi) = H, CX^2 = I, ...).  ECR, RZX and XX+YY follow Qiskit's published
definitions; no stored vector in the reference pins their phase conventions (parity unpinned).
"""
import math

import numpy as np
from scipy.linalg import expm

I2 = np.eye(2, dtype=np.complex128)
PX = np.array([[0, 1], [1, 0]], dtype=np.complex128)
PY = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
PZ = np.array([[1, 0], [0, -1]], dtype=np.complex128)
P0 = np.diag([1, 0]).astype(np.complex128)
P1 = np.diag([0, 1]).astype(np.complex128)


def _rot(P, th):
    return expm(-0.5j * th * P)


def u3(th, ph, lam):
    return np.exp(0.5j * (ph + lam)) * (_rot(PZ, ph) @ _rot(PY, th) @ _rot(PZ, lam))


def phase(lam):
    return np.exp(0.5j * lam) * _rot(PZ, lam)


def kron_le(*mats):
    """Kronecker product in little-endian order: mats[0] acts on the lowest qubit."""
    out = np.eye(1, dtype=np.complex128)
    for m in mats:
        out = np.kron(m, out)
    return out


def controlled(U, nctrl=1):
    """Controls are the first nctrl qubits, the target block is the remaining qubits."""
    g = int(round(math.log2(U.shape[0])))
    dim = 1 << (nctrl + g)
    M = np.eye(dim, dtype=np.complex128)
    on = (1 << nctrl) - 1
    idx = [on | (j << nctrl) for j in range(1 << g)]
    M[np.ix_(idx, idx)] = U
    return M


def _one(name, p):
    if name in ("id", "i"):
        return I2
    if name == "h":
        return (PX + PZ) / math.sqrt(2)
    if name == "x":
        return PX
    if name == "y":
        return PY
    if name == "z":
        return PZ
    if name == "s":
        return phase(math.pi / 2)
    if name == "sdg":
        return phase(-math.pi / 2)
    if name == "t":
        return phase(math.pi / 4)
    if name == "tdg":
        return phase(-math.pi / 4)
    if name == "sx":
        return np.exp(0.25j * math.pi) * _rot(PX, math.pi / 2)
    if name == "sxdg":
        return np.exp(-0.25j * math.pi) * _rot(PX, -math.pi / 2)
    if name == "rx":
        return _rot(PX, p[0])
    if name == "ry":
        return _rot(PY, p[0])
    if name == "rz":
        return _rot(PZ, p[0])
    if name in ("p", "u1"):
        return phase(p[0])
    if name in ("u", "u3"):
        return u3(p[0], p[1], p[2])
    if name == "u2":
        return u3(math.pi / 2, p[0], p[1])
    if name == "r":
        return _rot(math.cos(p[1]) * PX + math.sin(p[1]) * PY, p[0])
    return None


def gate_matrix(name, nq, params=()):
    p = [float(v) for v in (params or ())]
    if nq == 1:
        U = _one(name, p)
        if U is None:
            raise KeyError(name)
        return U
    if name.startswith("c") and name not in ("ccx", "ccz", "cswap", "cu") and nq == 2:
        base = _one(name[1:], p)
        if base is not None:
            return controlled(base)
    if name == "cu":
        return controlled(np.exp(1j * p[3]) * u3(p[0], p[1], p[2]))
    if name == "swap":
        M = np.zeros((4, 4), dtype=np.complex128)
        for i in range(4):
            M[((i & 1) << 1) | (i >> 1), i] = 1
        return M
    if name == "iswap":
        return expm(0.25j * math.pi * (kron_le(PX, PX) + kron_le(PY, PY)))
    if name == "dcx":
        cx01 = controlled(PX)
        swap = gate_matrix("swap", 2)
        cx10 = swap @ cx01 @ swap
        return cx10 @ cx01
    if name == "ecr":
        return (kron_le(PX, I2) - kron_le(PY, PX)) / math.sqrt(2)
    if name == "rxx":
        return expm(-0.5j * p[0] * kron_le(PX, PX))
    if name == "ryy":
        return expm(-0.5j * p[0] * kron_le(PY, PY))
    if name == "rzz":
        return expm(-0.5j * p[0] * kron_le(PZ, PZ))
    if name == "rzx":
        return expm(-0.5j * p[0] * kron_le(PZ, PX))
    if name == "xx_plus_yy":
        th, beta = p[0], p[1]
        rz0 = lambda a: kron_le(_rot(PZ, a), I2)
        core = expm(-0.25j * th * (kron_le(PX, PX) + kron_le(PY, PY)))
        return rz0(-beta) @ core @ rz0(beta)
    if name == "ccx":
        return controlled(PX, 2)
    if name == "ccz":
        return controlled(PZ, 2)
    if name == "cswap":
        return controlled(gate_matrix("swap", 2), 1)
    if name == "mcx":
        return controlled(PX, nq - 1)
    if name in ("mcp", "mcphase"):
        return controlled(phase(p[0]), nq - 1)
    raise KeyError(name)


def apply_gate(state, n, U, qubits):
    """Apply a 2^g x 2^g matrix U (little-endian over `qubits`) to a length-2^n state."""
    g = len(qubits)
    psi = state.reshape((2,) * n)
    Ut = U.reshape((2,) * (2 * g))
    # U's row-major axes: out bit g-1 .. 0, then in bit g-1 .. 0; qubit q is psi axis n-1-q
    in_axes = [n - 1 - qubits[g - 1 - i] for i in range(g)]
    out = np.tensordot(Ut, psi, axes=(list(range(g, 2 * g)), in_axes))
    # out axes: U out bits (g-1..0), then remaining psi axes in order
    rest = [a for a in range(n) if a not in in_axes]
    order = in_axes + rest
    inv = np.argsort(order)
    return np.transpose(out, inv).reshape(-1)


def simulate(gates, n, state=None):
    """gates: iterable of (name, qubits, params). Returns the complex128 final state."""
    if state is None:
        state = np.zeros(1 << n, dtype=np.complex128)
        state[0] = 1
    else:
        state = np.asarray(state, dtype=np.complex128).copy()
    for name, qs, params in gates:
        if name in ("measure", "barrier", "delay", "save_statevector"):
            continue
        U = gate_matrix(name, len(qs), params)
        state = apply_gate(state, n, U, list(qs))
    return state


def mux2_operator(n, t, cmask, M0, M1):
    """Dense 2^n x 2^n operator of MUX2(t, cmask, M0, M1) on physical bits, built index by index."""
    N = 1 << n
    O = np.zeros((N, N), dtype=np.complex128)
    bt = 1 << t
    for i0 in range(N):
        if i0 & bt:
            continue
        M = M1 if (cmask and (i0 & cmask) == cmask) else M0
        i1 = i0 | bt
        O[i0, i0], O[i0, i1] = M[0, 0], M[0, 1]
        O[i1, i0], O[i1, i1] = M[1, 0], M[1, 1]
    return O
