"""Bounded descriptor memory and per-thread preparation options.

A long-running server prepares a segment per distinct circuit; the descriptors of released handles
stay cached up to a byte limit (svx_cache_limit) and are evicted least-recently-used first, so 1000
distinct circuits keep the library's device memory bounded.  Options are explicit per call (a
thread-local block in Python), so two threads preparing the same records with different options get
their own settings."""
import threading

import numpy as np
import pytest

from u55c_statevector_30q_amd import circuit as C
from u55c_statevector_30q_amd import compiler as K

pytestmark = pytest.mark.gpu


def test_thousand_distinct_circuits_keep_device_bytes_bounded(gpu):
    import torch
    from u55c_statevector_30q_amd import libsvx
    n = 14
    limit = 8 << 20
    torch.cuda.synchronize()
    # descriptors that earlier callers in this process still hold references to stay resident;
    # a zero limit evicts every unreferenced one, leaving exactly those
    held0 = libsvx.cache_limit(0)
    libsvx.cache_limit(limit)
    st = torch.empty(1 << n, dtype=torch.complex64, device="cuda")
    free0 = None
    held = []
    try:
        for i in range(1000):
            qc = C.u3_cx_layers(n, 3, 1000 + i)          # a distinct circuit (seeded angles) each time
            prog, fmap, stats = K.compile_gpu(qc)
            hs = [libsvx.prepare(w, libsvx.C64, n) for _, w in prog]
            libsvx.zero_state(st)
            for h in hs:
                libsvx.run_prepared(st, h)
            for h in hs:
                libsvx.release_prepared(h)
            if i % 100 == 99:
                torch.cuda.synchronize()
                held.append(libsvx.cache_limit(limit))
                free, _ = torch.cuda.mem_get_info(0)
                if free0 is None:
                    free0 = free
                # device memory taken after the first 100 circuits stays within a few arena slabs
                assert free0 - free <= (64 << 20), (i, free0 - free)
        print(f"[descriptor cache] bytes held by other handles {held0}; after every 100 circuits: {held}")
        assert max(held) - held0 <= limit, (held0, held)
        assert abs(float(st.abs().square().sum()) - 1) < 1e-4
    finally:
        libsvx.cache_limit(256 << 20)


def test_released_handles_are_reused_and_unknown_handles_raise(gpu):
    from u55c_statevector_30q_amd import libsvx
    prog, fmap, stats = K.compile_gpu(C.make("qft", 12))
    w = prog[0][1]
    h1 = libsvx.prepare(w, libsvx.C64, 12)
    h2 = libsvx.prepare(w, libsvx.C64, 12)
    assert h1 == h2, "same records and options share one cached handle"
    libsvx.release_prepared(h1)
    libsvx.release_prepared(h2)
    with pytest.raises(ValueError):
        libsvx.release_prepared(10 ** 12)


def test_two_threads_with_different_options_get_their_own_settings(gpu):
    """Thread A prepares with fusion off, thread B with the defaults (FP32 dense units), at the same
    time and on the same records: A's passes run on the MUX2 kernel, B's on dense-unit kernels, and
    each thread's options block leaves the other's untouched."""
    from u55c_statevector_30q_amd import libsvx
    prog, fmap, stats = K.compile_gpu(C.u3_cx_layers(20, 6, 3))
    words = [w for _, w in prog]
    barrier = threading.Barrier(2)
    out = {}

    def worker(name, kw):
        with libsvx.options(**kw):
            barrier.wait()
            for _ in range(5):
                hs = [libsvx.prepare(w, libsvx.C64, 20) for w in words]
                kernels = set()
                for h in hs:
                    npass, _ = libsvx.prepared_info(h)
                    kernels |= {libsvx.prepared_kernel(h, p) for p in range(npass)}
                units = sum(libsvx.prepared_pass_stats(h, p)["dense4"] for h in hs
                            for p in range(libsvx.prepared_info(h)[0]))
                for h in hs:
                    libsvx.release_prepared(h)
            out[name] = (kernels, units, libsvx.current_options().fuse)

    ta = threading.Thread(target=worker, args=("off", dict(fuse="off")))
    tb = threading.Thread(target=worker, args=("auto", {}))
    ta.start(), tb.start()
    ta.join(), tb.join()
    print(f"[options per thread] {out}")
    assert out["off"][1] == 0 and all("mux" in k for k in out["off"][0]), out
    assert out["auto"][1] > 0, out
    assert out["off"][2] == libsvx.FUSE["off"] and out["auto"][2] == libsvx.FUSE["auto"]
    assert libsvx.current_options() is None
