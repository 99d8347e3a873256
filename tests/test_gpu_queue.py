"""The persistent queue of whole-state circuits (svx_queue_*): many small circuits per launch.

Every circuit of up to 2^13 amplitudes compiles to one whole-state pass (serving.compile_request);
the queue's resident kernel runs each with the MUX2 kernel's device code, so its states must be
bit-identical to svx_run_prepared of the same handle on a |0...0> state -- plain MUX2 records and
FP32 dense units alike -- and match the complex128 golden model to the FP32 bound.  The queue also
relaunches after its kernel has left on idling, and rejects jobs it cannot run."""
import time

import numpy as np
import pytest

from _c64 import check_c64, prog_nops
from oracle import model
from u55c_statevector_30q_amd import circuit as C

pytestmark = pytest.mark.gpu


def _circuits(rng, count, bits):
    from u55c_statevector_30q_amd import serving
    out = []
    for i in range(count):
        qc = C.u3_cx_layers(bits, int(rng.integers(5, 40)), 500 + 17 * i + bits)
        prog, fmap, st = serving.compile_request(qc)
        assert len(prog) == 1 and st.n_passes == 1
        out.append((prog, prog[0][1]))
    return out


@pytest.mark.parametrize("bits", [12, 13])
def test_queue_matches_single_runs_bit_identical(gpu, bits):
    import torch
    from u55c_statevector_30q_amd import libsvx
    rng = np.random.default_rng(bits)
    circs = _circuits(rng, 40, bits)
    q = libsvx.Queue(bits, capacity=64, workgroups=16)
    try:
        handles, states, tickets = [], [], []
        for i, (prog, w) in enumerate(circs):
            with libsvx.options(**({} if i % 2 == 0 else dict(fuse="off"))):   # dense units / plain MUX2
                h = libsvx.prepare(w, libsvx.C64, bits)
            s = torch.full((1 << bits,), 7.0, dtype=torch.complex64, device="cuda")   # overwritten: zero_init
            handles.append(h)
            states.append(s)
            tickets.append(q.submit(s, h))
        for t in tickets:
            q.wait(t)
        kinds = set()
        for i, ((prog, w), h, s) in enumerate(zip(circs, handles, states)):
            alone = torch.empty(1 << bits, dtype=torch.complex64, device="cuda")
            libsvx.zero_state(alone)
            libsvx.run_prepared(alone, h)
            torch.cuda.synchronize()
            assert torch.equal(s, alone), i
            kinds.add(libsvx.prepared_pass_stats(h, 0)["dense4"] > 0)
            if i < 6:
                ref = model.run_program(prog, bits, 1, list(range(bits)), ctype=np.complex128)
                check_c64(ref, s.cpu().numpy(), ("queue", bits, i), prog_nops(prog))
        assert kinds == {True, False}, "both dense-unit and plain MUX2 jobs ran"
        for h in handles:
            libsvx.release_prepared(h)
    finally:
        q.close()


def test_queue_relaunches_after_idling_and_reads_a_given_state(gpu):
    """idle_us = 200: the kernel leaves between the two batches of jobs; the second batch relaunches
    it.  zero_init False: the job applies its pass to the state it is given."""
    import torch
    from u55c_statevector_30q_amd import libsvx
    bits = 12
    rng = np.random.default_rng(3)
    circs = _circuits(rng, 6, bits)
    q = libsvx.Queue(bits, capacity=8, workgroups=4, idle_us=200)
    try:
        hs = [libsvx.prepare(w, libsvx.C64, bits) for _, w in circs]
        for rnd in range(2):
            states = [torch.empty(1 << bits, dtype=torch.complex64, device="cuda") for _ in hs]
            for s in states:
                libsvx.zero_state(s)
            torch.cuda.synchronize()
            ts = [q.submit(s, h, zero_init=False) for s, h in zip(states, hs)]
            for t in ts:
                q.wait(t)
                assert q.done(t)
            for s, h in zip(states, hs):
                alone = torch.empty_like(s)
                libsvx.zero_state(alone)
                libsvx.run_prepared(alone, h)
                torch.cuda.synchronize()
                assert torch.equal(s, alone)
            time.sleep(0.05)   # far beyond the idle time: the resident kernel has left
        for h in hs:
            libsvx.release_prepared(h)
    finally:
        q.close()


def test_queue_rejects_jobs_it_cannot_run(gpu):
    import torch
    from u55c_statevector_30q_amd import compiler, libsvx
    q = libsvx.Queue(12, capacity=4, workgroups=2)
    try:
        prog, fmap, st = compiler.compile_gpu(C.u3_cx_layers(14, 10, 1))      # 2^14: not the queue's size
        h = libsvx.prepare(prog[0][1], libsvx.C64, 14)
        s = torch.empty(1 << 14, dtype=torch.complex64, device="cuda")
        with pytest.raises(ValueError):
            q.submit(s, h)
        libsvx.release_prepared(h)
        prog, fmap, st = compiler.compile_gpu(C.u3_cx_layers(12, 30, 1), k=10)   # several passes
        assert st.n_passes > 1
        h = libsvx.prepare(prog[0][1], libsvx.C64, 12)
        s = torch.empty(1 << 12, dtype=torch.complex64, device="cuda")
        with pytest.raises(ValueError):
            q.submit(s, h)
        libsvx.release_prepared(h)
    finally:
        q.close()
    with pytest.raises(ValueError):
        libsvx.Queue(14)
