"""In-place permutations (svx_permute with dst == src) and the runtime's remap fallback.

The reference exchanges bank ranges in place (kernels/sv/sv_swap.cpp:10-39, driven by
host/sv/runtime.py:235-259) and falls back to a host exchange when its device-swap self-test fails
(host/sv/runtime.py:136-150, 192-199).  Here: in-place permutations are bit-exact against
np.transpose (host/sv/model.py:74-78) at every size and dtype; SVRuntime remaps in place when a
second state does not fit, and falls back from the out-of-place to the in-place device path when
the out-of-place path fails -- the failures are injected on the host (a wrapper that raises or
returns a wrong result), never on the device."""
import numpy as np
import pytest

from u55c_statevector_30q_amd import circuit as C

pytestmark = pytest.mark.gpu


def _np_perm(x, perm):
    n = len(perm)
    return x.reshape([2] * n).transpose([n - 1 - perm[n - 1 - i] for i in range(n)]).reshape(-1)


@pytest.mark.parametrize("dtype", ["c64", "c128"])
def test_in_place_permutation_bit_exact(gpu, dtype):
    import torch
    from u55c_statevector_30q_amd import libsvx
    rng = np.random.default_rng(21)
    npdt = np.complex64 if dtype == "c64" else np.complex128
    cases = [(n, [int(x) for x in rng.permutation(n)]) for n in (1, 4, 6, 9, 12, 13, 16, 18, 22) for _ in range(3)]
    cases += [(n, list(range(n))[::-1]) for n in (12, 13, 20, 24)]
    cases += [(20, list(range(20)))]
    for n, perm in cases:
        host = (rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)).astype(npdt)
        t = torch.from_numpy(host).to("cuda")
        libsvx.permute(t, t, perm)
        torch.cuda.synchronize()
        assert np.array_equal(t.cpu().numpy(), _np_perm(host, perm)), (n, perm)
        u = torch.from_numpy(host).to("cuda")
        libsvx.permute(None, u, perm)           # NULL destination: in place as well
        assert np.array_equal(u.cpu().numpy(), _np_perm(host, perm)), (n, perm)


def test_in_place_round_trip_and_out_of_place_agree(gpu):
    import torch
    from u55c_statevector_30q_amd import libsvx
    rng = np.random.default_rng(5)
    n = 23
    perm = [int(x) for x in rng.permutation(n)]
    inv = [perm.index(b) for b in range(n)]
    host = (rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)).astype(np.complex64)
    a = torch.from_numpy(host).to("cuda")
    b = torch.empty_like(a)
    libsvx.permute(b, a, perm)
    libsvx.permute(a, a, perm)
    assert torch.equal(a, b)
    libsvx.permute(a, a, inv)
    assert np.array_equal(a.cpu().numpy(), host)


def test_partial_overlap_is_rejected(gpu):
    import torch
    from u55c_statevector_30q_amd import libsvx
    t = torch.zeros(1 << 11, dtype=torch.complex64, device="cuda")
    with pytest.raises(ValueError):
        libsvx.permute(t[: 1 << 10], t[1 << 9: (1 << 9) + (1 << 10)], list(range(10)))


def _reference_plan_run(rt, qc):
    out = rt.run(qc, shots=0)
    assert out["stats"]["n_remaps"] > 0, "the reference plan remaps this circuit"
    return out["state"]


def test_runtime_falls_back_to_in_place_when_out_of_place_fails(gpu):
    """Injected on the host: the out-of-place path raises at the first remap.  The runtime logs the
    fallback, switches to in-place remaps, and the result is bit-identical to the out-of-place run."""
    from u55c_statevector_30q_amd.runtime import SVRuntime
    qc = C.make("qft", 18)
    base = SVRuntime(target="reference")
    want = _reference_plan_run(base, qc)
    assert base.remap_mode == "out_of_place" and not any("fail" in m for m in base.log)

    rt = SVRuntime(target="reference")
    rt.self_test()
    calls = []

    def broken(dst, src, perm, stream=None):
        calls.append(1)
        raise RuntimeError("injected: permute launch failed")
    rt._permute_oop = broken
    got = _reference_plan_run(rt, qc)
    assert calls and rt.remap_mode == "in_place"
    assert any("out-of-place remap failed" in m for m in rt.log), rt.log
    assert np.array_equal(got, want)


def test_runtime_self_test_detects_a_wrong_out_of_place_result(gpu):
    """Injected on the host: the out-of-place path returns a wrong state (unchanged source copy)."""
    import torch
    from u55c_statevector_30q_amd.runtime import SVRuntime
    rt = SVRuntime(target="reference")

    def wrong(dst, src, perm, stream=None):
        dst.copy_(src)
    rt._permute_oop = wrong
    ok = rt.self_test()
    assert ok == {"out_of_place": False, "in_place": True}
    assert rt.remap_mode == "in_place" and any("self-test failed" in m for m in rt.log)
    qc = C.make("rqc", 18)
    got = _reference_plan_run(rt, qc)
    want = _reference_plan_run(SVRuntime(target="reference"), qc)
    assert np.array_equal(got, want)
    torch.cuda.synchronize()


def test_runtime_remaps_in_place_when_a_second_state_does_not_fit(gpu):
    from u55c_statevector_30q_amd.runtime import SVRuntime
    qc = C.make("vqe", 19)
    want = _reference_plan_run(SVRuntime(target="reference"), qc)
    rt = SVRuntime(target="reference")
    real = rt.admit
    rt.admit = lambda n, dtype=None, buffers=2: buffers == 1 and real(n, dtype, 1)
    got = _reference_plan_run(rt, qc)
    assert rt.tmp is None, "no second state was allocated"
    assert any("does not fit" in m for m in rt.log), rt.log
    assert np.array_equal(got, want)


def test_oversize_state_raises_runtime_error(gpu):
    """The reference's capacity error (host/sv/runtime.py:220-221): RuntimeError before allocating."""
    from u55c_statevector_30q_amd.runtime import SVRuntime
    rt = SVRuntime()
    with pytest.raises(RuntimeError, match="does not fit"):
        rt.run(C.ghz(40), shots=0)
