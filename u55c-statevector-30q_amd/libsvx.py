"""ctypes binding of libsvx.so (include/svx.h) -- the Python side of the drop-in boundary.

Device buffers are torch tensors (complex64 / complex128 on a ROCm device); only their
data_ptr() and the current HIP stream cross the boundary.  Negative return codes become the
exceptions the reference raises at the same conditions (AssertionError for a bad magic,
host/sv/model.py:9; ValueError for bad arguments; RuntimeError for capacity / runtime failures,
host/sv/runtime.py:220-221).  The library must be built (python csrc/build.py); there is no
fallback path.
"""
import contextlib
import ctypes
import os
import threading

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libsvx.so")

C64, C128 = 0, 1
SVX_E_ARG, SVX_E_MAGIC, SVX_E_FORMAT, SVX_E_CAPACITY, SVX_E_HIP, SVX_E_NODEV = -1, -2, -3, -4, -5, -6

SYMBOLS = ["svx_init", "svx_run_segment", "svx_run_segment_ex", "svx_engine", "svx_default_options",
           "svx_prepare", "svx_prepare_ex", "svx_release_prepared", "svx_cache_limit", "svx_prepared_shape",
           "svx_run_prepared", "svx_tune_prepared", "svx_prepared_info", "svx_prepared_kernel",
           "svx_prepared_pass_stats", "svx_run_batch", "svx_queue_open", "svx_queue_submit", "svx_queue_done",
           "svx_queue_wait", "svx_queue_close", "svx_permute", "svx_permute_plan", "svx_swap_words", "svx_zero_state",
           "svx_product_state", "svx_reduce_f64", "svx_reduce_f64_noalloc", "svx_fidelity_f64", "svx_convert", "svx_sync",
           "svx_last_error", "svx_version"]

FUSE = {"auto": 0, "on": 0, "off": 1, "dense4": 2}
DENSE_MATH = {"f32": 0, "split16": 1}
KERNEL = {"auto": 0, "single": 1}
TILE_ORDER = {"tunable": 0, "identity": 1, "xcd": 2}
MUX_ENGINE = {"auto": 0, "valu": 1}


class Options(ctypes.Structure):
    """svx_options: how a segment is prepared (include/svx.h)."""
    _fields_ = [("size", ctypes.c_int32), ("fuse", ctypes.c_int32), ("dense_math", ctypes.c_int32),
                ("kernel", ctypes.c_int32), ("diag_tables", ctypes.c_int32), ("dense_only", ctypes.c_int32),
                ("tile_scale", ctypes.c_int32), ("tile_order", ctypes.c_int32), ("mux_engine", ctypes.c_int32),
                ("fuse_width", ctypes.c_int32), ("fuse_cost4", ctypes.c_float), ("fuse_cost5", ctypes.c_float),
                ("reserved", ctypes.c_int32 * 4)]


def make_options(fuse="auto", dense_math="f32", kernel="auto", diag_tables=True, dense_only=True,
                 tile_scale=True, tile_order="tunable", mux_engine="auto", fuse_width=0, fuse_cost4=0.0,
                 fuse_cost5=0.0):
    """An Options value from names (defaults = the library's defaults: FP32 everywhere)."""
    o = Options()
    o.size = ctypes.sizeof(Options)
    o.fuse = FUSE[fuse]
    o.dense_math = DENSE_MATH[dense_math]
    o.kernel = KERNEL[kernel]
    o.diag_tables = int(bool(diag_tables))
    o.dense_only = int(bool(dense_only))
    o.tile_scale = int(bool(tile_scale))
    o.tile_order = TILE_ORDER[tile_order]
    o.mux_engine = MUX_ENGINE[mux_engine]
    o.fuse_width = int(fuse_width)
    o.fuse_cost4 = float(fuse_cost4)
    o.fuse_cost5 = float(fuse_cost5)
    return o


_tls = threading.local()


def current_options():
    """The calling thread's options (innermost `options(...)` block), or None for the defaults."""
    st = getattr(_tls, "stack", None)
    return st[-1][1] if st else None


@contextlib.contextmanager
def options(**kw):
    """Thread-local preparation options for prepare / run_segment calls inside the block, layered
    over the enclosing block's settings (no environment variables, no process-wide state)."""
    st = getattr(_tls, "stack", None)
    if st is None:
        st = _tls.stack = []
    merged = dict(st[-1][0]) if st else {}
    merged.update(kw)
    st.append((merged, make_options(**merged)))
    try:
        yield st[-1][1]
    finally:
        st.pop()


class SvxCaps(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("cu_count", ctypes.c_int32),
                ("lds_per_block", ctypes.c_int64), ("hbm_bytes", ctypes.c_int64),
                ("wavefront", ctypes.c_int32), ("max_k_c64", ctypes.c_int32),
                ("max_k_c128", ctypes.c_int32), ("max_high", ctypes.c_int32),
                ("arch", ctypes.c_char * 64)]


_lib = None

_VP, _I32, _I64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
_PI32 = ctypes.POINTER(ctypes.c_int32)
_PD = ctypes.POINTER(ctypes.c_double)
_SIGS = {
    "svx_init": ([_I32, ctypes.POINTER(SvxCaps)], _I32),
    "svx_run_segment": ([_VP, _I64, _I32, _I32, _PI32, _I64, _VP], _I32),
    "svx_run_segment_ex": ([_VP, _I64, _I32, _I32, _PI32, _I64, ctypes.POINTER(Options), _VP], _I32),
    "svx_engine": ([_VP, _I64, _VP], _I32),
    "svx_default_options": ([ctypes.POINTER(Options)], _I32),
    "svx_prepare": ([_PI32, _I64, _I32, _I32, ctypes.POINTER(_I64)], _I32),
    "svx_prepare_ex": ([_PI32, _I64, _I32, _I32, ctypes.POINTER(Options), ctypes.POINTER(_I64)], _I32),
    "svx_release_prepared": ([_I64], _I32),
    "svx_cache_limit": ([_I64, ctypes.POINTER(_I64)], _I32),
    "svx_prepared_shape": ([_I64, _PI32, _PI32], _I32),
    "svx_run_prepared": ([_VP, _I64, _VP], _I32),
    "svx_tune_prepared": ([_VP, _I64, _I32, _VP], _I32),
    "svx_prepared_info": ([_I64, _PI32, _PI32], _I32),
    "svx_prepared_kernel": ([_I64, _I32, ctypes.c_char_p, _I32], _I32),
    "svx_prepared_pass_stats": ([_I64, _I32, _PI32, ], _I32),
    "svx_run_batch": ([ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_int64), _PI32, _I32, _VP], _I32),
    "svx_queue_open": ([_I32, _I32, _I32, _I32, ctypes.POINTER(_I64)], _I32),
    "svx_queue_submit": ([_I64, _VP, _I64, _I32, ctypes.POINTER(_I64)], _I32),
    "svx_queue_done": ([_I64, _I64, _PI32], _I32),
    "svx_queue_wait": ([_I64, _I64], _I32),
    "svx_queue_close": ([_I64], _I32),
    "svx_permute": ([_VP, _VP, _I32, _PI32, _I32, _VP], _I32),
    "svx_permute_plan": ([_I32, _PI32, _I32, _PI32, _I32, _PI32], _I32),
    "svx_swap_words": ([_VP, _VP, _I64, _I64, _I64, _VP], _I32),
    "svx_zero_state": ([_VP, _I32, _I32, _VP], _I32),
    "svx_product_state": ([_VP, _I32, _I32, _PD, _VP], _I32),
    "svx_reduce_f64": ([_VP, _I32, _I32, _PI32, _I32, _PD, _PD, _VP], _I32),
    "svx_reduce_f64_noalloc": ([_VP, _I32, _I32, _PI32, _I32, _PD, _PD, _VP], _I32),
    "svx_fidelity_f64": ([_VP, _I32, _VP, _I32, _I32, _PD, _VP], _I32),
    "svx_convert": ([_VP, _I32, _VP, _I32, _I64, _VP], _I32),
    "svx_sync": ([_VP], _I32),
    "svx_last_error": ([], ctypes.c_char_p),
    "svx_version": ([], ctypes.c_char_p),
}


def load(path=LIB_PATH):
    """Load libsvx.so (torch is imported first so one HIP runtime serves both)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run `python u55c-statevector-30q_amd/csrc/build.py` "
                           "(or __graft_entry__.build()); there is no non-HIP fallback")
    import torch  # noqa: F401  (binds libamdhip64 before libsvx resolves it)
    lib = ctypes.CDLL(path)
    for name, (args, res) in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def _check(rc):
    if rc == 0:
        return
    msg = _lib.svx_last_error().decode(errors="replace")
    if rc == SVX_E_MAGIC:
        raise AssertionError(msg)
    if rc in (SVX_E_ARG, SVX_E_FORMAT):
        raise ValueError(msg)
    raise RuntimeError(f"libsvx error {rc}: {msg}")


def _i32(a):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.int32))
    return a, a.ctypes.data_as(_PI32)


def dtype_code(t):
    import torch
    if t.dtype == torch.complex64:
        return C64
    if t.dtype == torch.complex128:
        return C128
    raise ValueError(f"state must be complex64 or complex128, got {t.dtype}")


def _stream(stream):
    if stream is not None:
        return ctypes.c_void_p(int(stream))
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _nbits(t):
    n = t.numel().bit_length() - 1
    if t.numel() != 1 << n:
        raise ValueError("state length must be a power of two")
    return n


def _dev(t):
    if not t.is_cuda or not t.is_contiguous():
        raise ValueError("state must be a contiguous device tensor")
    return ctypes.c_void_p(t.data_ptr())


def init(device=0):
    lib = load()
    caps = SvxCaps()
    _check(lib.svx_init(device, ctypes.byref(caps)))
    return caps


def version():
    return load().svx_version().decode()


def _opts_ptr(opts):
    o = opts if opts is not None else current_options()
    return None if o is None else ctypes.byref(o)


def run_segment(state, words, n_slices=1, stream=None, opts=None):
    """svx_run_segment: every 2^L slice of `state` (n_slices of them) runs the QSV2 segment, with
    `opts` (an Options) or the thread's current options."""
    lib = load()
    w, wp = _i32(words)
    n_slice = state.numel() // n_slices
    _check(lib.svx_run_segment_ex(_dev(state), n_slice, n_slices, dtype_code(state), wp, w.size, _opts_ptr(opts),
                                  _stream(stream)))


def engine(buf_u8, ins_off, stream=None):
    """svx_engine: the reference kernel ABI -- a uint8 device buffer holding the 2^L complex64
    slice followed by the instruction region at 256-bit word ins_off."""
    lib = load()
    _check(lib.svx_engine(_dev(buf_u8), int(ins_off), _stream(stream)))


def fusion(mode):
    """Options block selecting the library's pass fusion ("auto" clusters MUX2 ops into dense units
    where that is cheaper, "off" runs every op as MUX2)."""
    if mode not in ("auto", "on", "off", "dense4"):
        raise ValueError("fusion mode must be auto, on, off or dense4")
    return options(fuse=mode)


def prepare(words, dtype, state_bits, opts=None):
    """svx_prepare_ex: a handle holding one reference (give it back with release_prepared)."""
    lib = load()
    w, wp = _i32(words)
    h = ctypes.c_int64(0)
    _check(lib.svx_prepare_ex(wp, w.size, dtype, state_bits, _opts_ptr(opts), ctypes.byref(h)))
    return h.value


def release_prepared(handle):
    _check(load().svx_release_prepared(int(handle)))


def cache_limit(limit_bytes):
    """Set the byte limit of the cache of unreferenced descriptors; returns the descriptor bytes held."""
    held = ctypes.c_int64(0)
    _check(load().svx_cache_limit(int(limit_bytes), ctypes.byref(held)))
    return held.value


def prepared_shape(handle):
    lib = load()
    dt, nb = ctypes.c_int32(0), ctypes.c_int32(0)
    _check(lib.svx_prepared_shape(int(handle), ctypes.byref(dt), ctypes.byref(nb)))
    return dt.value, nb.value


def _check_state(state, handle):
    dt, nb = prepared_shape(handle)
    if dtype_code(state) != dt:
        raise ValueError(f"state dtype {state.dtype} does not match the prepared segment's dtype")
    if state.numel() != 1 << nb:
        raise ValueError(f"state holds {state.numel()} amplitudes; the segment was prepared for 2^{nb}")


def run_prepared(state, handle, stream=None):
    lib = load()
    _check_state(state, handle)
    _check(lib.svx_run_prepared(_dev(state), handle, _stream(stream)))


def tune_prepared(state, handle, reps=2, stream=None):
    """svx_tune_prepared: pick each pass's workgroup-to-tile order by timing both candidates on
    `state` (synchronous; the passes are applied 2 (1 + reps) times to it, so use a state whose
    contents may change).  Results of later runs stay bit-identical."""
    lib = load()
    _check_state(state, handle)
    _check(lib.svx_tune_prepared(_dev(state), handle, int(reps), _stream(stream)))


def prepared_info(handle):
    lib = load()
    a, b = ctypes.c_int32(0), ctypes.c_int32(0)
    _check(lib.svx_prepared_info(handle, ctypes.byref(a), ctypes.byref(b)))
    return a.value, b.value


def prepared_kernel(handle, pass_index):
    lib = load()
    buf = ctypes.create_string_buffer(64)
    _check(lib.svx_prepared_kernel(handle, pass_index, buf, 64))
    return buf.value.decode()


def run_batch(states, handles, passes, stream=None):
    """svx_run_batch: one pass of each (state, prepared handle, pass index) item, batched into one
    launch per kernel shape (complex64 states)."""
    lib = load()
    n = len(states)
    if not (len(handles) == len(passes) == n):
        raise ValueError("states, handles and passes must have the same length")
    for t, h in zip(states, handles):
        if dtype_code(t) != C64:
            raise ValueError("batched launches run complex64 states")
        _check_state(t, h)
    if len({t.data_ptr() for t in states}) != n:
        raise ValueError("batch items must touch distinct states")
    sp = (ctypes.c_void_p * n)(*[_dev(t) for t in states])
    hp = (ctypes.c_int64 * n)(*[int(h) for h in handles])
    pp = (ctypes.c_int32 * n)(*[int(p) for p in passes])
    _check(lib.svx_run_batch(sp, hp, pp, n, _stream(stream)))


def prepared_pass_stats(handle, pass_index):
    """Work of one pass after classification and fusion: MUX2 groups, MUX2 ops, dense units on 4
    and 5 qubits, dense units on the FP32 matrix cores (all of them in the default FP32 mode),
    MUX2 ops on the matrix cores, diagonal tables (runs of diagonal ops applied as one table of
    phases), and the workgroup-to-tile order the pass runs with (-2: XCD-contiguous, 0: identity)."""
    lib = load()
    out = (ctypes.c_int32 * 8)()
    _check(lib.svx_prepared_pass_stats(handle, pass_index, out))
    return dict(mux_groups=out[0], mux_ops=out[1], dense4=out[2], dense5=out[3], dense_exact=out[4],
                mux_mfma=out[5], diag_tables=out[6], tile_order=out[7])


class Queue:
    """svx_queue_*: a persistent kernel running whole-state circuits of 2^state_bits amplitudes
    (one prepared whole-state pass each) from a host-visible ring, many circuits per launch."""

    def __init__(self, state_bits, capacity=1024, workgroups=64, idle_us=20000):
        self.bits = state_bits
        h = ctypes.c_int64(0)
        _check(load().svx_queue_open(int(state_bits), int(capacity), int(workgroups), int(idle_us), ctypes.byref(h)))
        self.h = h.value

    def submit(self, state, handle, zero_init=True):
        _check_state(state, handle)
        t = ctypes.c_int64(0)
        _check(load().svx_queue_submit(self.h, _dev(state), int(handle), int(bool(zero_init)), ctypes.byref(t)))
        return t.value

    def done(self, ticket):
        d = ctypes.c_int32(0)
        _check(load().svx_queue_done(self.h, int(ticket), ctypes.byref(d)))
        return bool(d.value)

    def wait(self, ticket):
        _check(load().svx_queue_wait(self.h, int(ticket)))

    def close(self):
        if self.h:
            _check(load().svx_queue_close(self.h))
            self.h = 0


def permute(dst, src, perm, stream=None):
    """svx_permute: dst = src with perm[new_bit] = old_bit (bit-exact).  dst None or dst is src:
    in place (a few sweeps, no second state)."""
    lib = load()
    if dst is None:
        dst = src
    if dst.dtype != src.dtype or dst.numel() != src.numel():
        raise ValueError("dst and src must have the same dtype and size")
    n = _nbits(src)
    p, pp = _i32(perm)
    if p.size != n:
        raise ValueError("perm must have one entry per qubit")
    _check(lib.svx_permute(_dev(dst), _dev(src), n, pp, dtype_code(src), _stream(stream)))


def permute_plan(perm, in_place=True):
    """svx_permute_plan: the step permutations of svx_permute (one per HBM sweep), as an int array
    of shape (steps, n)."""
    p, pp = _i32(perm)
    n = p.size
    cnt = ctypes.c_int32(0)
    lib = load()
    _check(lib.svx_permute_plan(n, pp, int(bool(in_place)), None, 0, ctypes.byref(cnt)))
    out = np.zeros((cnt.value, n), dtype=np.int32)
    _check(lib.svx_permute_plan(n, pp, int(bool(in_place)), out.ctypes.data_as(_PI32), cnt.value, ctypes.byref(cnt)))
    return out


def swap_words(a, b, offa, offb, nwords, stream=None):
    lib = load()
    _check(lib.svx_swap_words(_dev(a), _dev(b), offa, offb, nwords, _stream(stream)))


def zero_state(state, stream=None):
    lib = load()
    _check(lib.svx_zero_state(_dev(state), _nbits(state), dtype_code(state), _stream(stream)))


def product_state(state, amps, stream=None):
    """amps: (n, 2) complex -- qubit q = amps[q, 0] |0> + amps[q, 1] |1>."""
    lib = load()
    n = _nbits(state)
    a = np.ascontiguousarray(np.asarray(amps, dtype=np.complex128).reshape(n, 2).view(np.float64).reshape(n, 4))
    _check(lib.svx_product_state(_dev(state), n, dtype_code(state), a.ctypes.data_as(_PD), _stream(stream)))


def reduce_f64(state, fmap=(), stream=None):
    """FP64 norm and the marginal distribution of logical qubits 0..len(fmap)-1 (at physical fmap[i])."""
    lib = load()
    n = _nbits(state)
    f, fp = _i32(list(fmap) if len(fmap) else [0])
    nl = len(fmap)
    norm = ctypes.c_double(0)
    marg = np.zeros(1 << nl, dtype=np.float64)
    _check(lib.svx_reduce_f64(_dev(state), n, dtype_code(state), fp, nl, ctypes.byref(norm),
                              marg.ctypes.data_as(_PD), _stream(stream)))
    return norm.value, marg


def reduce_f64_noalloc(state, fmap=(), stream=None):
    """svx_reduce_f64_noalloc: reduce_f64 working in the library's static scratch array (no device
    allocation; at most 8 marginal qubits; overlapping calls on one GPU race)."""
    lib = load()
    n = _nbits(state)
    f, fp = _i32(list(fmap) if len(fmap) else [0])
    nl = len(fmap)
    norm = ctypes.c_double(0)
    marg = np.zeros(1 << nl, dtype=np.float64)
    _check(lib.svx_reduce_f64_noalloc(_dev(state), n, dtype_code(state), fp, nl, ctypes.byref(norm),
                                      marg.ctypes.data_as(_PD), _stream(stream)))
    return norm.value, marg


def fidelity_f64(ref, psi, stream=None):
    """(F, 1-F without cancellation, max|da|, l2) of psi against the reference ref."""
    lib = load()
    if ref.numel() != psi.numel():
        raise ValueError("states differ in size")
    out = np.zeros(4, dtype=np.float64)
    _check(lib.svx_fidelity_f64(_dev(ref), dtype_code(ref), _dev(psi), dtype_code(psi), _nbits(ref),
                                out.ctypes.data_as(_PD), _stream(stream)))
    return dict(fidelity=float(out[0]), one_minus_f=float(out[1]), max_abs_err=float(out[2]), l2=float(out[3]))


def convert(dst, src, stream=None):
    lib = load()
    if dst.numel() != src.numel():
        raise ValueError("size mismatch")
    _check(lib.svx_convert(_dev(dst), dtype_code(dst), _dev(src), dtype_code(src), src.numel(), _stream(stream)))


def sync(stream=None):
    lib = load()
    _check(lib.svx_sync(_stream(stream)))
