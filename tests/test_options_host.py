"""Preparation options on the host (no GPU): explicit svx_options values, thread-local `options`
blocks that nest and never leak between threads, and no dependence on environment variables."""
import os
import threading

from u55c_statevector_30q_amd import libsvx


def test_make_options_maps_names_to_the_header_constants():
    o = libsvx.make_options(fuse="off", dense_math="split16", kernel="single", tile_order="identity",
                            mux_engine="valu", fuse_width=5, fuse_cost4=20.0)
    assert (o.fuse, o.dense_math, o.kernel, o.tile_order, o.mux_engine) == (1, 1, 1, 1, 1)
    assert o.size == 4 * 16 and o.fuse_width == 5 and abs(o.fuse_cost4 - 20.0) < 1e-6
    d = libsvx.make_options()
    assert (d.fuse, d.dense_math, d.kernel) == (0, 0, 0), "FP32 dense units are the default"


def test_options_blocks_nest_and_are_per_thread():
    seen = {}
    gate = threading.Barrier(2)

    def worker(name, kw, inner):
        with libsvx.options(**kw):
            gate.wait()
            a = libsvx.current_options()
            with libsvx.options(**inner):
                b = libsvx.current_options()
            gate.wait()
            c = libsvx.current_options()
            seen[name] = ((a.fuse, a.dense_math), (b.fuse, b.dense_math), (c.fuse, c.dense_math))
        seen[name + "/after"] = libsvx.current_options()

    t1 = threading.Thread(target=worker, args=("t1", dict(fuse="off"), dict(dense_math="split16")))
    t2 = threading.Thread(target=worker, args=("t2", dict(dense_math="split16"), dict(fuse="dense4")))
    t1.start(), t2.start()
    t1.join(), t2.join()
    assert seen["t1"] == ((1, 0), (1, 1), (1, 0))
    assert seen["t2"] == ((0, 1), (2, 1), (0, 1))
    assert seen["t1/after"] is None and seen["t2/after"] is None
    assert libsvx.current_options() is None


def test_environment_variables_are_not_read():
    os.environ["SVX_FUSE"] = "off"
    os.environ["SVX_DENSE_MATH"] = "split16"
    try:
        with libsvx.options() as o:
            assert (o.fuse, o.dense_math) == (0, 0)
    finally:
        del os.environ["SVX_FUSE"], os.environ["SVX_DENSE_MATH"]
    src = open(os.path.join(os.path.dirname(libsvx.__file__), "csrc", "svx_api.hip")).read()
    assert "getenv" not in src
