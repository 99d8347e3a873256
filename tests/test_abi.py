"""The C-ABI library loads and exports every function include/svx.h declares (no GPU needed:
nothing here launches a kernel)."""
import ctypes
import os
import re

from u55c_statevector_30q_amd import libsvx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "svx.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(svx_\w+)\s*\(", txt, re.M)))


def test_header_declares_the_binding_surface():
    decl = _declared()
    assert decl == sorted(libsvx.SYMBOLS)
    assert "svx_run_segment" in decl and "svx_engine" in decl and "svx_permute" in decl


def test_library_exports_every_declared_symbol():
    assert os.path.exists(libsvx.LIB_PATH), "libsvx.so not built (run __graft_entry__.build())"
    lib = libsvx.load()
    for name in _declared():
        assert isinstance(getattr(lib, name), ctypes._CFuncPtr), name
    assert lib.svx_version().decode().startswith("libsvx")


def test_errors_map_without_a_device():
    lib = libsvx.load()
    import torch
    if torch.cuda.is_available():
        return
    rc = lib.svx_init(0, None)
    assert rc == libsvx.SVX_E_NODEV
    assert lib.svx_last_error()


def test_bench_is_one_gpu_only():
    """bench.py measures one GPU: any other --gpus exits with status 2 and says so, before touching a
    device."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr, (r.returncode, r.stderr)
