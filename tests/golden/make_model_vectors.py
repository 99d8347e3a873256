"""Write golden vectors of the reference's numpy golden model (host/sv/model.py) for the complex64
rounding that no closed form pins.

Run in a checkout that has the reference at /root/reference (it is read from there, never copied):
    python tests/golden/make_model_vectors.py
The reference's host/sv/model.py (numpy + host/sv/isa.py only) executes each stream; this script
stores the inputs (QSV2 words in the reference format, slice count, final map, initial state) and
the reference's complex64 outputs in model_vectors.npz.  The streams come from this package's
compiler with the reference's plan (k <= 12, at most 4 high qubits) and from random MUX2 records
with controls, so every op kind and tile shape of the reference format appears.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference/host/sv"


def _ours():
    sys.path.insert(0, ROOT)
    import __graft_entry__ as ge
    ge.load_package()
    from u55c_statevector_30q_amd import circuit, compiler
    from oracle.isa import MAGIC, OP_MUX2
    return circuit, compiler, MAGIC, OP_MUX2


def _random_segment(rng, L, k, n_passes, max_ops, max_m, MAGIC, OP_MUX2):
    w = [MAGIC, L, n_passes, k, 0, 0, 0, 0]
    for _ in range(n_passes):
        m = int(rng.integers(0, min(max_m, k) + 1))
        hi = sorted(rng.choice(list(range(k - m, L)), size=m, replace=False).tolist()) if m else []
        h = hi + [0] * (4 - len(hi))
        nops = int(rng.integers(1, max_ops + 1))
        w += [m, h[0], h[1], nops, h[2], h[3], 0, 0]
        for _ in range(nops):
            t = int(rng.integers(0, k))
            cm = 0
            for b in range(k):
                if b != t and rng.random() < 0.25:
                    cm |= 1 << b
            mats = []
            for _ in range(2):
                z = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))
                q, r = np.linalg.qr(z)
                mats.append((q * (np.diag(r) / np.abs(np.diag(r)))).astype(np.complex64))
            fl = np.zeros(16, dtype=np.float32)
            for i, M in enumerate(mats):
                f = M.reshape(-1)
                fl[8 * i:8 * i + 8:2] = f.real
                fl[8 * i + 1:8 * i + 8:2] = f.imag
            w += [OP_MUX2, t, cm, 0, 0, 0, 0, 0] + [int(x) for x in fl.view(np.int32)]
    return (np.asarray(w, dtype=np.int64) & 0xFFFFFFFF).astype(np.uint32).view(np.int32)


def main():
    circuit, compiler, MAGIC, OP_MUX2 = _ours()
    sys.path.insert(0, REF)
    import model as ref_model   # the reference's golden model (numpy only)

    out = {}
    cases = []
    # compiled circuits with the reference's plan: (name, circuit, P, L, k)
    for name, qc, P, k in [("u3cx_10_d6_s0", circuit.u3_cx_layers(10, 6, 0), 4, 6),
                           ("qft_9", circuit.qft(9), 2, 5),
                           ("rqc_10_d10", circuit.rqc(10, depth=10), 4, 6),
                           ("vqe_8_L2", circuit.vqe(8, layers=2), 2, 5),
                           ("ghz_10", circuit.ghz(10), 8, 4)]:
        n = qc.num_qubits
        L = n - (P.bit_length() - 1)
        prog, fmap, _ = compiler.compile_circuit(qc, L, k)
        got = ref_model.run_program(prog, n, P, fmap)
        i = len(cases)
        cases.append(name)
        out[f"c{i}_kind"] = np.array(0)
        out[f"c{i}_n"] = np.array(n)
        out[f"c{i}_P"] = np.array(P)
        out[f"c{i}_fmap"] = np.asarray(fmap, dtype=np.int32)
        out[f"c{i}_nsteps"] = np.array(len(prog))
        for j, (kind, payload) in enumerate(prog):
            out[f"c{i}_s{j}_kind"] = np.array(0 if kind == "segment" else 1)
            out[f"c{i}_s{j}"] = np.asarray(payload, dtype=np.int32)
        out[f"c{i}_out"] = np.asarray(got, dtype=np.complex64)
    # random MUX2 segments with controls on random complex64 states
    rng = np.random.default_rng(2026)
    for (L, k, P, max_m) in [(8, 5, 1, 2), (9, 6, 2, 3), (10, 7, 1, 4), (7, 7, 4, 0), (10, 4, 2, 4)]:
        seg = _random_segment(rng, L, k, n_passes=3, max_ops=10, max_m=max_m, MAGIC=MAGIC, OP_MUX2=OP_MUX2)
        n = L + (P.bit_length() - 1)
        s = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
        s0 = (s / np.linalg.norm(s)).astype(np.complex64)
        got = ref_model.run_segment(s0.copy(), P, seg)
        i = len(cases)
        cases.append(f"random_L{L}_k{k}_P{P}_m{max_m}")
        out[f"c{i}_kind"] = np.array(1)
        out[f"c{i}_n"] = np.array(n)
        out[f"c{i}_P"] = np.array(P)
        out[f"c{i}_seg"] = seg
        out[f"c{i}_in"] = s0
        out[f"c{i}_out"] = np.asarray(got, dtype=np.complex64)
    out["cases"] = np.array(cases)
    path = os.path.join(HERE, "model_vectors.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {len(cases)} cases, numpy {np.__version__}")


if __name__ == "__main__":
    main()
