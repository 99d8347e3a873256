"""The oracle and the HIP engine against golden vectors of the reference's own numpy golden model.

tests/golden/model_vectors.npz holds outputs of host/sv/model.py (the reference's golden model,
run by tests/golden/make_model_vectors.py in a checkout with the reference present) on reference-
format streams: compiled circuits with several slices and remaps, and random MUX2 records with
controls on random complex64 states.  The oracle restates the model's complex64 expression order
(host/sv/model.py:43-56), so it must reproduce these outputs bit for bit; the HIP kernels (FMA
contraction, another summation order) are held to the FP32 rounding bound per amplitude of
tests/_c64.py against the complex128 golden model of the same records and, at twice that bound,
against these complex64 vectors (both sides round); the complex128 kernels to 1e-6 max|da| against
these complex64 vectors.
"""
import os

import numpy as np
import pytest

from _c64 import C_BOUND, check_c64, prog_nops, seg_nops

from oracle import judge, model

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "model_vectors.npz")


def _cases():
    z = np.load(PATH, allow_pickle=False)
    out = []
    for i, name in enumerate(z["cases"]):
        c = dict(name=str(name), kind=int(z[f"c{i}_kind"]), n=int(z[f"c{i}_n"]), P=int(z[f"c{i}_P"]),
                 out=z[f"c{i}_out"])
        if c["kind"] == 0:
            c["fmap"] = [int(v) for v in z[f"c{i}_fmap"]]
            c["prog"] = [("segment" if int(z[f"c{i}_s{j}_kind"]) == 0 else "remap",
                          z[f"c{i}_s{j}"] if int(z[f"c{i}_s{j}_kind"]) == 0 else [int(v) for v in z[f"c{i}_s{j}"]])
                         for j in range(int(z[f"c{i}_nsteps"]))]
        else:
            c["seg"] = z[f"c{i}_seg"]
            c["in"] = z[f"c{i}_in"]
        out.append(c)
    return out


CASES = _cases()


@pytest.mark.parametrize("c", CASES, ids=[c["name"] for c in CASES])
def test_oracle_reproduces_reference_model_bit_for_bit(c):
    if c["kind"] == 0:
        got = model.run_program(c["prog"], c["n"], c["P"], c["fmap"], ctype=np.complex64)
    else:
        got = model.run_segment(c["in"].copy(), c["P"], c["seg"], np.complex64)
    assert got.dtype == np.complex64
    assert np.array_equal(got.view(np.uint64), c["out"].view(np.uint64)), c["name"]


def test_golden_vectors_cover_the_format():
    kinds = {c["kind"] for c in CASES}
    assert kinds == {0, 1}
    assert any(c["P"] > 1 for c in CASES)
    assert any(step[0] == "remap" for c in CASES if c["kind"] == 0 for step in c["prog"])


def _run_hip(c, dtype):
    import torch
    from u55c_statevector_30q_amd import libsvx
    td = torch.complex64 if dtype == "c64" else torch.complex128
    n, P = c["n"], c["P"]
    L = n - (P.bit_length() - 1)
    if c["kind"] == 1:
        st = torch.from_numpy(c["in"].astype(np.complex64 if dtype == "c64" else np.complex128)).to("cuda")
        libsvx.run_segment(st, c["seg"], n_slices=P)
        torch.cuda.synchronize()
        return st.cpu().numpy()
    st = torch.empty(1 << n, dtype=td, device="cuda")
    libsvx.zero_state(st)
    for kind, payload in c["prog"]:
        if kind == "segment":
            libsvx.run_segment(st, payload, n_slices=P)
        else:
            tmp = torch.empty_like(st)
            libsvx.permute(tmp, st, payload)
            st = tmp
    torch.cuda.synchronize()
    return model.unpermute(st.cpu().numpy(), n, c["fmap"])


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ["c64", "c128"])
@pytest.mark.parametrize("c", CASES, ids=[c["name"] for c in CASES])
def test_hip_engine_matches_reference_model(gpu, c, dtype):
    got = _run_hip(c, dtype)
    ref = c["out"].astype(np.complex128)
    m = judge.metrics(ref, got)
    if dtype == "c64":
        if c["kind"] == 0:
            gold = model.run_program(c["prog"], c["n"], c["P"], c["fmap"], ctype=np.complex128)
            nops = prog_nops(c["prog"])
        else:
            gold = model.run_segment(c["in"].astype(np.complex128), c["P"], c["seg"], np.complex128)
            nops = seg_nops(c["seg"])
        check_c64(gold, got, c["name"], nops)
        check_c64(ref, got, (c["name"], "reference c64 vector"), nops, c=2 * C_BOUND)
    else:
        assert m["max_abs_err"] < 1e-6, (c["name"], m)
