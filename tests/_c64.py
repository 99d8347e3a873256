"""The complex64 parity bar of the GPU tests: an FP32 rounding bound per amplitude.

A complex64 run of MUX2 records rounds every pair update to FP32, as the reference does
(kernels/sv/sv_engine.cpp:155-162, host/sv/model.py:53-54): each op adds an error of a few units
in the last place of the amplitudes it combines.  Against the complex128 run of the same records,
amplitude i may therefore differ by at most

    |a64_i - a128_i| <= C * nops * 2^-24 * (|a128_i| + rms)

with nops the ops applied (MUX2 or dense records), rms = ||a128|| / sqrt(N) covering amplitudes
that cancel towards zero, and C a small constant.  An arithmetic that is narrower than FP32 (for
instance split FP16, whose error is relative to the tile maximum) fails it on amplitudes far below
their neighbours; the old absolute bar (max|da| < 2e-5) could not see that.  The fidelity must also
be FP32-level (1 - F < 1e-9).
"""
import numpy as np

from oracle import judge, model

EPS32 = 2.0 ** -24
C_BOUND = 4.0


def seg_nops(words):
    """Ops of a QSV2 segment in the bound's units: a MUX2 record counts 1, a dense record on f qubits
    counts f (it stands for at least f one-qubit ops, and each of its outputs sums 2^f products)."""
    from oracle.isa import OP_DENSE
    _, _, passes = model.decode(np.asarray(words))
    return sum(len(op[1]) if op[0] == OP_DENSE else 1 for p in passes for op in p[-1])


def amp_error_ratio(ref, got, nops):
    """max_i |got_i - ref_i| / (nops 2^-24 (|ref_i| + rms)): <= C_BOUND for FP32 arithmetic."""
    ref = np.asarray(ref, dtype=np.complex128).reshape(-1)
    got = np.asarray(got).astype(np.complex128).reshape(-1)
    rms = np.linalg.norm(ref) / np.sqrt(ref.size)
    err = np.abs(got - ref)
    return float(np.max(err / (max(1, nops) * EPS32 * (np.abs(ref) + rms))))


def check_c64(ref, got, what, nops, c=C_BOUND, one_minus_f=1e-9):
    r = amp_error_ratio(ref, got, nops)
    m = judge.metrics(np.asarray(ref, dtype=np.complex128), got)
    print(f"[c64 bound] {what!r} nops={nops} ratio={r:.3f} 1-F={m['one_minus_f']:.2e}")
    assert r <= c and m["one_minus_f"] < one_minus_f, (what, nops, r, m)
    return r


def prog_nops(prog):
    """Op records over every segment of a compiled program [(kind, payload)]."""
    return sum(seg_nops(w) for kind, w in prog if kind == "segment")
