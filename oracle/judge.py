"""Judge metrics in FP64, restating the reference's definitions.

fidelity      -- F = |<ref|psi>|^2 / (<psi|psi> <ref|ref>) accumulated in complex128
                 (tests/functional_suite.py:90-94, bench/fpga/sv_bench.py:75-76).
metrics       -- after optimal global-phase alignment of the normalised states:
                 1 - F = d^2 - d^4 / 4 (no cancellation), l2 = d, max|da|
                 (bench/fpga/sv_maxwidth_ref.py:15-25).
Pinned by the identities F(psi, e^{ia} psi) = 1, d = 0, and by a two-state closed form
(|0> vs cos(e)|0> + sin(e)|1>: 1 - F = sin^2 e) in tests/test_oracle.py.
"""
import numpy as np

CHUNK = 1 << 22


def fidelity(ref, psi):
    ref = np.asarray(ref)
    psi = np.asarray(psi)
    ip = 0j
    nr = 0.0
    ns = 0.0
    for i in range(0, len(ref), CHUNK):
        a = ref[i:i + CHUNK].astype(np.complex128)
        b = psi[i:i + CHUNK].astype(np.complex128)
        ip += complex(np.vdot(a, b))
        nr += float(np.vdot(a, a).real)
        ns += float(np.vdot(b, b).real)
    return abs(ip) ** 2 / (nr * ns)


def metrics(ref, psi):
    ref = np.asarray(ref)
    psi = np.asarray(psi)
    nr = ns = 0.0
    ip = 0j
    for i in range(0, len(ref), CHUNK):
        a = ref[i:i + CHUNK].astype(np.complex128)
        b = psi[i:i + CHUNK].astype(np.complex128)
        nr += float(np.vdot(a, a).real)
        ns += float(np.vdot(b, b).real)
        ip += complex(np.vdot(a, b))
    nr, ns = np.sqrt(nr), np.sqrt(ns)
    ip /= (nr * ns)
    ph = ip / abs(ip) if abs(ip) > 0 else 1.0
    d2 = 0.0
    mx = 0.0
    for i in range(0, len(ref), CHUNK):
        e = psi[i:i + CHUNK].astype(np.complex128) / ns - ph * ref[i:i + CHUNK].astype(np.complex128) / nr
        d2 += float(np.vdot(e, e).real)
        mx = max(mx, float(np.abs(e).max()))
    return dict(fidelity=abs(ip) ** 2, one_minus_f=d2 - d2 * d2 / 4, l2=float(np.sqrt(d2)),
                max_abs_err=mx, norm=float(ns))


def max_abs_diff(a, b):
    return float(np.max(np.abs(np.asarray(a, dtype=np.complex128) - np.asarray(b, dtype=np.complex128))))
