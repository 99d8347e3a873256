"""Closed-form gate matrices in the convention the reference compiler receives from Qiskit
(host/sv/compiler.py:137 calls Operator(...).data): little-endian, a g-qubit gate on
qubits (q0, .., q_{g-1}) is indexed by sum_i bit(q_i) << i; controlled gates take their
control(s) first.  Only closed forms live here; the test suite checks them against the
oracle's independent derivation from Pauli exponentials.
"""
import cmath
import math

import numpy as np

C = np.complex128
I2 = np.eye(2, dtype=C)
X = np.array([[0, 1], [1, 0]], dtype=C)
Y = np.array([[0, -1j], [1j, 0]], dtype=C)
Z = np.array([[1, 0], [0, -1]], dtype=C)
H = np.array([[1, 1], [1, -1]], dtype=C) / math.sqrt(2)


def U(th, ph, lam):
    c, s = math.cos(th / 2), math.sin(th / 2)
    return np.array([[c, -cmath.exp(1j * lam) * s],
                     [cmath.exp(1j * ph) * s, cmath.exp(1j * (ph + lam)) * c]], dtype=C)


def P(lam):
    return np.array([[1, 0], [0, cmath.exp(1j * lam)]], dtype=C)


def RX(th):
    c, s = math.cos(th / 2), math.sin(th / 2)
    return np.array([[c, -1j * s], [-1j * s, c]], dtype=C)


def RY(th):
    c, s = math.cos(th / 2), math.sin(th / 2)
    return np.array([[c, -s], [s, c]], dtype=C)


def RZ(th):
    return np.array([[cmath.exp(-0.5j * th), 0], [0, cmath.exp(0.5j * th)]], dtype=C)


def R(th, ph):
    c, s = math.cos(th / 2), math.sin(th / 2)
    return np.array([[c, -1j * cmath.exp(-1j * ph) * s], [-1j * cmath.exp(1j * ph) * s, c]], dtype=C)


SX = np.array([[1 + 1j, 1 - 1j], [1 - 1j, 1 + 1j]], dtype=C) / 2
SXDG = np.array([[1 - 1j, 1 + 1j], [1 + 1j, 1 - 1j]], dtype=C) / 2

ONE_QUBIT = {
    "id": lambda p: I2, "i": lambda p: I2, "h": lambda p: H, "x": lambda p: X, "y": lambda p: Y,
    "z": lambda p: Z, "s": lambda p: P(math.pi / 2), "sdg": lambda p: P(-math.pi / 2),
    "t": lambda p: P(math.pi / 4), "tdg": lambda p: P(-math.pi / 4), "sx": lambda p: SX,
    "sxdg": lambda p: SXDG, "rx": lambda p: RX(p[0]), "ry": lambda p: RY(p[0]),
    "rz": lambda p: RZ(p[0]), "p": lambda p: P(p[0]), "u1": lambda p: P(p[0]),
    "u": lambda p: U(p[0], p[1], p[2]), "u3": lambda p: U(p[0], p[1], p[2]),
    "u2": lambda p: U(math.pi / 2, p[0], p[1]), "r": lambda p: R(p[0], p[1]),
}


def ctrl(Umat, nctrl=1):
    """Controls on the lowest nctrl gate qubits (all ones selects Umat on the rest)."""
    g = Umat.shape[0].bit_length() - 1
    dim = 1 << (nctrl + g)
    M = np.eye(dim, dtype=C)
    on = (1 << nctrl) - 1
    idx = [on | (j << nctrl) for j in range(1 << g)]
    M[np.ix_(idx, idx)] = Umat
    return M


def _two(name, p):
    if name == "swap":
        return np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], dtype=C)
    if name == "iswap":
        return np.array([[1, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0], [0, 0, 0, 1]], dtype=C)
    if name == "dcx":
        return np.array([[1, 0, 0, 0], [0, 0, 0, 1], [0, 1, 0, 0], [0, 0, 1, 0]], dtype=C)
    if name == "ecr":
        return np.array([[0, 1, 0, 1j], [1, 0, -1j, 0], [0, 1j, 0, 1], [-1j, 0, 1, 0]], dtype=C) / math.sqrt(2)
    if name in ("rxx", "ryy", "rzz", "rzx"):
        th = p[0]
        c, s = math.cos(th / 2), math.sin(th / 2)
        if name == "rzz":
            e0, e1 = cmath.exp(-0.5j * th), cmath.exp(0.5j * th)
            return np.diag([e0, e1, e1, e0]).astype(C)
        if name == "rxx":
            return np.array([[c, 0, 0, -1j * s], [0, c, -1j * s, 0], [0, -1j * s, c, 0], [-1j * s, 0, 0, c]], dtype=C)
        if name == "ryy":
            return np.array([[c, 0, 0, 1j * s], [0, c, -1j * s, 0], [0, -1j * s, c, 0], [1j * s, 0, 0, c]], dtype=C)
        # rzx: Z on the first qubit, X on the second
        return np.array([[c, 0, -1j * s, 0], [0, c, 0, 1j * s], [-1j * s, 0, c, 0], [0, 1j * s, 0, c]], dtype=C)
    if name == "xx_plus_yy":
        th, beta = p[0], p[1]
        c, s = math.cos(th / 2), math.sin(th / 2)
        M = np.eye(4, dtype=C)
        M[1, 1] = M[2, 2] = c
        M[1, 2] = -1j * s * cmath.exp(-1j * beta)
        M[2, 1] = -1j * s * cmath.exp(1j * beta)
        return M
    if name == "cu":
        return ctrl(cmath.exp(1j * p[3]) * U(p[0], p[1], p[2]))
    if name.startswith("c") and name[1:] in ONE_QUBIT:
        return ctrl(ONE_QUBIT[name[1:]](p))
    return None


def matrix(name, nq, params=()):
    """Dense matrix of a named gate on nq qubits (little-endian, controls first)."""
    p = [float(v) for v in (params or ())]
    if nq == 1 and name in ONE_QUBIT:
        return ONE_QUBIT[name](p)
    if nq == 2:
        M = _two(name, p)
        if M is not None:
            return M
    if name == "ccx" or (name == "mcx" and nq >= 2):
        return ctrl(X, nq - 1)
    if name == "ccz":
        return ctrl(Z, nq - 1)
    if name in ("mcp", "mcphase"):
        return ctrl(P(p[0]), nq - 1)
    if name == "cswap":
        return ctrl(_two("swap", p), 1)
    raise KeyError(f"unsupported gate {name!r} on {nq} qubits")
