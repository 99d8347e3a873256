"""Full-size runs at 2^32 and 2^34 amplitudes (32 and 128 GiB of complex64), where 32-bit amplitude
indices overflow, checked against the analytic answers of the reference's suite
(tests/functional_suite.py:74-82): GHZ has |a_0|^2 = |a_(2^n - 1)|^2 = 1/2 and nothing else, and
QFT|0> is the uniform state 2^(-n/2) with zero phases.  Both answers are invariant under qubit
permutations, so they are checked on the physical-order state the program leaves (no unpermute
buffer).  Each runs the plain form (every op on the MUX2 kernel) and the fused form (dense 4-qubit
records on the FP32 matrix cores).  FP64 reductions on the device; complex64 tolerances 1e-6 on the
GHZ probabilities and on the norm, and for QFT|0> an amplitude error relative to 2^(-n/2) below
2 * ops * 2^-24 (the reference's recorded deviations at 30 qubits are 6.8e-8 for GHZ and 2.4e-7 for
QFT, results/raw/phase4/sv_functional_sv_e16w_m4.json)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(qc, form):
    import torch
    from u55c_statevector_30q_amd import compiler, libsvx
    n = qc.num_qubits
    free, _ = torch.cuda.mem_get_info(0)
    if (1 << n) * 8 + (4 << 30) > free:
        pytest.skip(f"a {n}-qubit complex64 state needs {(1 << n) * 8 >> 30} GiB; {free >> 30} GiB free")
    if form == "plain":
        prog, fmap, st = compiler.compile_gpu(qc)
        opts = dict(fuse="off")
    else:
        prog, fmap, st = compiler.compile_fused(qc, width=4)
        opts = {}
    s = torch.empty(1 << n, dtype=torch.complex64, device="cuda")
    libsvx.zero_state(s)
    with libsvx.options(**opts):
        for kind, w in prog:
            assert kind == "segment", "the families compile without remaps at one slice"
            libsvx.run_segment(s, w)
    torch.cuda.synchronize()
    return s, fmap, st


@pytest.mark.parametrize("form", ["plain", "fused"])
@pytest.mark.parametrize("n", [32, 34])
def test_ghz_full_size(gpu, n, form):
    import torch
    from u55c_statevector_30q_amd import circuit, libsvx
    s, fmap, st = _run(circuit.ghz(n), form)
    norm, marg = libsvx.reduce_f64(s, list(fmap[:8]))
    a0, a1 = complex(s[0].item()), complex(s[-1].item())
    print(f"[ghz {n} {form}] passes={st.n_passes} norm-1={norm - 1:.2e} |a0|^2={abs(a0) ** 2:.9f} |a1|^2={abs(a1) ** 2:.9f}")
    assert abs(abs(a0) ** 2 - 0.5) < 1e-6 and abs(abs(a1) ** 2 - 0.5) < 1e-6
    assert norm - abs(a0) ** 2 - abs(a1) ** 2 < 1e-6          # nothing anywhere else
    assert abs(marg[0] - 0.5) < 1e-6 and abs(marg[255] - 0.5) < 1e-6
    del s
    torch.cuda.empty_cache()


@pytest.mark.parametrize("form", ["plain", "fused"])
@pytest.mark.parametrize("n", [32, 34])
def test_qft_of_zero_full_size(gpu, n, form):
    import torch
    from u55c_statevector_30q_amd import circuit, libsvx
    s, fmap, st = _run(circuit.qft(n), form)
    norm, marg = libsvx.reduce_f64(s, list(fmap[:8]))
    c = 2.0 ** (-n / 2)
    worst = 0.0
    step = 1 << 28
    for i in range(0, 1 << n, step):   # every amplitude equals 2^(-n/2) (zero phase)
        worst = max(worst, float((s[i:i + step] - c).abs().max()))
    print(f"[qft {n} {form}] passes={st.n_passes} norm-1={norm - 1:.2e} max|a-2^(-n/2)|/2^(-n/2)={worst / c:.2e} "
          f"marginal dev={np.max(np.abs(marg - 1 / 256)):.2e}")
    assert abs(norm - 1) < 1e-6
    assert np.max(np.abs(marg - 1 / 256)) < 1e-5 / 256
    assert worst / c < 2 * st.n_ops * 2.0 ** -24   # FP32 rounding, at most a few units per op
    del s
    torch.cuda.empty_cache()
