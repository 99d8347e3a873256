"""The allocation-free reduction (svx_reduce_f64_noalloc) through the C ABI: two calls in a row, the
second on a smaller state with fewer marginal qubits, each equal bit for bit to svx_reduce_f64 (the
same kernels, grid and reduction order) and to the FP64 marginals computed in numpy; n_low beyond
the scratch array's limit is rejected."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _state(rng, n, dtype):
    s = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return (s / np.linalg.norm(s)).astype(dtype)


def _marginals(s, n, fmap):
    p = np.abs(s.astype(np.complex128)) ** 2
    idx = np.arange(1 << n)
    x = np.zeros(1 << n, dtype=np.int64)
    for i, b in enumerate(fmap):
        x |= ((idx >> b) & 1) << i
    return np.bincount(x, weights=p, minlength=1 << len(fmap))


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_noalloc_reduction_matches_reduce_f64(gpu, dtype):
    import torch
    from u55c_statevector_30q_amd import libsvx
    rng = np.random.default_rng(42)
    for n, fmap in ((22, [3, 17, 0, 21, 9, 12, 5, 1]), (12, [11, 2, 7, 4])):   # second call: smaller
        s = _state(rng, n, dtype)
        t = torch.from_numpy(s).to("cuda")
        norm_a, marg_a = libsvx.reduce_f64_noalloc(t, fmap)
        norm_b, marg_b = libsvx.reduce_f64(t, fmap)
        assert norm_a == norm_b and np.array_equal(marg_a, marg_b), (n, fmap)
        ref = _marginals(s, n, fmap)
        assert np.max(np.abs(marg_a - ref)) < 1e-12 and abs(norm_a - ref.sum()) < 1e-12, (n, fmap)


def test_noalloc_reduction_rejects_too_many_bins(gpu):
    import torch
    from u55c_statevector_30q_amd import libsvx
    t = torch.zeros(1 << 12, dtype=torch.complex64, device="cuda")
    with pytest.raises(RuntimeError):
        libsvx.reduce_f64_noalloc(t, list(range(9)))
