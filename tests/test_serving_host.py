"""Host side of the served mix (no GPU): the request stream generator and latency statistics."""
import numpy as np

from u55c_statevector_30q_amd import serving


def test_stream_is_seeded_poisson_and_compiled():
    a = serving.make_stream(rate=5.0, duration=6.0, seed=3, max_qubits=16)
    b = serving.make_stream(rate=5.0, duration=6.0, seed=3, max_qubits=16)
    assert [(r.n, r.depth, r.arrival) for r in a] == [(r.n, r.depth, r.arrival) for r in b]
    t = np.array([r.arrival for r in a])
    assert np.all(np.diff(t) > 0) and t[-1] <= 6.0 and 10 <= len(a) <= 60
    for r in a:
        assert 12 <= r.n <= 16 and 10 <= r.depth <= 60
        assert all(kind == "segment" for kind, _ in r.prog)
        assert sorted(r.fmap) == list(range(r.n))


def test_small_circuits_compile_to_one_whole_state_pass():
    from u55c_statevector_30q_amd import circuit as C
    from u55c_statevector_30q_amd.compiler import split_passes
    for n in (12, 13):
        prog, fmap, st = serving.compile_request(C.u3_cx_layers(n, 40, 1))
        assert len(prog) == 1 and len(split_passes(prog[0][1])) == 1 and st.n_passes == 1
    prog, fmap, st = serving.compile_request(C.u3_cx_layers(16, 10, 1))
    assert st.n_passes > 1


def test_latency_stats():
    rs = []
    for i in range(100):
        r = serving.Request(rid=i, n=12, prog=[], fmap=list(range(12)), arrival=0.1 * i)
        r.t_done = r.arrival + 0.01 + 0.001 * (i % 10)
        rs.append(r)
    s = serving.latency_stats(rs)
    assert s["n"] == 100 and abs(s["p50_ms"] - 14.5) < 0.6 and s["p99_ms"] <= 19.0 + 1e-9
    assert abs(s["circuits_per_s"] - 100 / (9.9 + 0.019)) < 0.05
