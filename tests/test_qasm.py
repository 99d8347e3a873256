"""OpenQASM 2.0 reader / writer (the front-end for the reference's circuits/qasm instances).

Pinned by: closed-form expression values; exact round trips of the gate zoo; the reference's own
instance files circuits/qasm/*.qasm (written by circuits/generators.py:100-102 through Qiskit's
qasm2 exporter), which must parse to the same gate lists as the restated generators -- and so
compile to the same QSV2 stream.  The instance files are committed as fixtures under
tests/golden/qasm/ (copied by tests/golden/make_qasm_fixtures.py).
"""
import glob
import math
import os
import re

import numpy as np
import pytest

from u55c_statevector_30q_amd import circuit, compiler, qasm

FIXTURES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "qasm")


@pytest.mark.parametrize("text,val", [
    ("pi/2", math.pi / 2), ("-pi/4", -math.pi / 4), ("2*pi/3", 2 * math.pi / 3), ("1.5e-3", 1.5e-3),
    ("sin(pi/2)", 1.0), ("(pi+1)^2", (math.pi + 1) ** 2), ("2^3^2", 2.0 ** 9), ("-2^2", -4.0),
    ("ln(exp(0.25))", 0.25), ("sqrt(16)/-2", -2.0), ("3-2-1", 0.0), ("8/4/2", 1.0), (".5", 0.5),
])
def test_expressions(text, val):
    assert qasm.eval_expr(text) == pytest.approx(val, rel=0, abs=1e-15)


@pytest.mark.parametrize("bad", ["pi/", "foo", "(1", "1 2", "sin 1"])
def test_bad_expressions(bad):
    with pytest.raises(ValueError):
        qasm.eval_expr(bad)


def _zoo():
    rng = np.random.default_rng(5)
    qc = circuit.Circuit(5, "zoo")
    a = lambda: float(rng.uniform(-7, 7))   # noqa: E731
    for g in ("h", "x", "y", "z", "s", "sdg", "t", "tdg", "sx", "sxdg"):
        getattr(qc, g)(int(rng.integers(5)))
    for g in ("rx", "ry", "rz", "p", "u1"):
        getattr(qc, g)(a(), int(rng.integers(5)))
    qc.u(a(), a(), a(), 2).u2(a(), a(), 3).r(a(), a(), 4)
    for g in ("cx", "cy", "cz", "ch", "swap", "iswap", "ecr", "dcx", "csx", "cs", "csdg"):
        getattr(qc, g)(0, 3)
    for g in ("cp", "crx", "cry", "crz", "rxx", "ryy", "rzz", "rzx", "cu1"):
        getattr(qc, g)(a(), 4, 1)
    qc.cu(a(), a(), a(), a(), 1, 2).cu(a(), a(), a(), 0.0, 2, 1).xx_plus_yy(a(), a(), 3, 0)
    qc.ccx(0, 1, 2).ccz(1, 2, 3).cswap(4, 0, 2).mcx([0, 1, 2], 4).mcx([0, 1, 2, 3], 4)
    return qc


def test_round_trip_exact():
    qc = _zoo()
    back = qasm.loads(qasm.dumps(qc))
    assert back.num_qubits == qc.num_qubits
    assert back.tuples() == qc.tuples()


def test_registers_broadcast_and_gate_definitions():
    src = """OPENQASM 2.0;
    include "qelib1.inc";   // the standard header
    qreg a[2];
    qreg b[2];
    creg c[2];
    gate mygate(theta) x, y { h x; cp(theta/2) x, y; }
    h a;
    cx a, b;
    mygate(pi) a[0], b[1];
    barrier a, b;
    u0(1) b[0];
    cu3(0.5, 0.25, 0.125) a[1], b[0];
    measure a -> c;
    """
    qc = qasm.loads(src)
    assert qc.num_qubits == 4
    assert qc.tuples() == [
        ("h", (0,), ()), ("h", (1,), ()), ("cx", (0, 2), ()), ("cx", (1, 3), ()),
        ("h", (0,), ()), ("cp", (0, 3), (math.pi / 2,)), ("id", (2,), ()),
        ("cu", (1, 2), (0.5, 0.25, 0.125, 0.0)),
    ]


@pytest.mark.parametrize("bad", [
    "OPENQASM 3.0;\nqreg q[1];\nh q[0];",
    "OPENQASM 2.0;\nqreg q[1];\nreset q[0];",
    "OPENQASM 2.0;\nqreg q[1];\nfoo q[0];",
    "OPENQASM 2.0;\nqreg q[1];\nh q[1];",
    "OPENQASM 2.0;\nqreg q[2];\nqreg r[3];\ncx q, r;",
    "OPENQASM 2.0;\nqreg q[1];\nrx q[0];",
    "OPENQASM 2.0;\nqreg q[1];\nh q[0]",
])
def test_errors(bad):
    with pytest.raises(ValueError):
        qasm.loads(bad)


def _family_of(stem):
    m = re.match(r"^(ghz|qft)_(\d+)$", stem)
    if m:
        return circuit.make(m.group(1), int(m.group(2)))
    m = re.match(r"^rqc_(\d+)_d(\d+)$", stem)
    if m:
        return circuit.rqc(int(m.group(1)), depth=int(m.group(2)))
    m = re.match(r"^vqe_(\d+)_L(\d+)$", stem)
    if m:
        return circuit.vqe(int(m.group(1)), layers=int(m.group(2)))
    return None


_FILES = sorted(glob.glob(os.path.join(FIXTURES, "*.qasm")))


def test_fixtures_present():
    assert len(_FILES) == 26, _FILES


@pytest.mark.parametrize("path", _FILES, ids=[os.path.basename(p) for p in _FILES])
def test_reference_instances_match_restated_generators(path):
    qc = qasm.load(path)
    gen = _family_of(os.path.splitext(os.path.basename(path))[0])
    assert gen is not None, path
    assert qc.num_qubits == gen.num_qubits
    got, want = qc.tuples(), gen.tuples()
    assert len(got) == len(want)
    for (gn, gq, gp), (wn, wq, wp) in zip(got, want):
        assert (gn, gq) == (wn, wq)
        assert np.allclose(gp, wp, rtol=0, atol=1e-12), (gn, gq, gp, wp)
    if qc.num_qubits <= 16:
        k = min(12, qc.num_qubits)
        a = compiler.compile_gpu(qc, k=k)[0]
        b = compiler.compile_gpu(gen, k=k)[0]
        assert len(a) == len(b)
        for (ka, wa), (kb, wb) in zip(a, b):
            assert ka == kb
            if ka == "segment":
                assert np.array_equal(wa, wb)
