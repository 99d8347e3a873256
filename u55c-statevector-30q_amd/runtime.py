"""SVRuntime: the reference runtime's interface (host/sv/runtime.py:83-334) on one MI355X.

`run(qc, ...)` compiles, zeroes the state on the device, launches every segment through libsvx
(the drop-in for the XRT run call), executes remaps on the device (instead of the reference's
transposition loop over sv_swap / host copies, host/sv/runtime.py:235-284), reads the state back,
samples shots and returns the reference's dict shape: dict(timing, P, L, k, n_run, stats, counts[, state]).

Remaps run out of place (one sweep into a second state buffer) when a second state fits in device
memory, otherwise in place (a few sweeps, no second buffer: libsvx.permute with dst = src).  Like the
reference's device-swap self-test with its host-exchange fallback (host/sv/runtime.py:136-150,
192-199), the runtime checks both device permutation paths once on a small state and falls back from
the out-of-place path to the in-place one when its self-test fails or its buffer cannot be allocated;
each fallback is recorded in `log`.  A state that does not fit raises RuntimeError (the reference's
capacity error, host/sv/runtime.py:220-221).

target="mi355x" (default) compiles with the MI355X retarget (one slice, 2^13-amplitude LDS tile, up
to 7 gathered high qubits, DAG-aware pass packing); target="reference" keeps the reference's
plan (P slices of 2^L, k <= 12, m <= max_high) and runs its stream unchanged -- P slices are one
device buffer whose top log2(P) index bits are the reference's bank qubits.
"""
import json
import os
import time

import numpy as np

from . import circuit as circuit_mod
from . import compiler, libsvx
from .compiler import compile_circuit, compile_fused, compile_gpu

KMAX = 12          # host/sv/runtime.py:26-32
LPREF = 17
LMIN = 12
AMP = 8


def sample_probabilities(state, shots, fmap, n, seed=None, chunk_bits=12):
    """Born sampling (host/sv/runtime.py:47-80): per-chunk sums (|a|^2 in the state's precision,
    accumulated in FP64 across chunks), choose chunks, then positions; physical indices map to
    logical bitstrings with fmap[l] = physical position of logical qubit l."""
    if shots <= 0:
        return {}
    rng = np.random.default_rng(seed)
    N = state.size
    cb = min(chunk_bits, N.bit_length() - 1)
    f = state.view(np.float32 if state.dtype == np.complex64 else np.float64).reshape(-1, 2 << cb)
    sums = np.einsum("ij,ij->i", f, f).astype(np.float64)
    cs = np.cumsum(sums)
    u = rng.random(shots) * cs[-1]
    ch = np.searchsorted(cs, u, side="right").clip(0, sums.size - 1)
    rem = u - np.where(ch > 0, cs[ch - 1], 0.0)
    phys = np.empty(shots, dtype=np.int64)
    for c in np.unique(ch):
        sel = np.nonzero(ch == c)[0]
        off = int(c) << cb
        p = np.cumsum(np.abs(state[off:off + (1 << cb)].astype(np.complex128)) ** 2)
        phys[sel] = off + np.searchsorted(p, rem[sel], side="right").clip(0, (1 << cb) - 1)
    logical = np.zeros(shots, dtype=np.int64)
    for l in range(n):
        logical |= ((phys >> fmap[l]) & 1) << l
    vals, cnt = np.unique(logical, return_counts=True)
    return {int(v): int(c) for v, c in zip(vals, cnt)}


FORMS = ("plain", "lib", "w4", "w5")
TUNE_KS = (12, 13, 14)


def compile_form(qc, k, form, max_high=None):
    """Compile for one execution form: "plain" (MUX2 records scheduled op by op, every op on the
    vector units), "lib" (MUX2 records scheduled cluster by cluster; the library fuses each pass's
    clusters into dense units), "w4" / "w5"
    (fused-form records: the compiler emits dense 4- / 5-qubit units)."""
    if form == "plain":
        return compile_gpu(qc, k=k, max_high=max_high, cluster=0)
    if form == "lib":
        return compile_gpu(qc, k=k, max_high=max_high)
    if form in ("w4", "w5"):
        return compile_fused(qc, k=k, width=int(form[1]), max_high=max_high)
    raise ValueError(f"unknown form {form!r}")


def form_fusion(form):
    """Library pass-fusion mode for a form (plain records run unfused for "plain")."""
    return "off" if form == "plain" else "auto"


class Autotuner:
    """Per (qubit count, dtype, device) tuning record: the tile width k and execution form with the
    shortest sweep time on a calibration circuit of the layer recipe (depth `depth`, its own seed),
    measured with HIP events on the resident state.  Records persist as JSON when `path` is set."""

    def __init__(self, path=None, ks=TUNE_KS, forms=FORMS, depth=4, seed=7):
        self.path = path
        self.ks = tuple(ks)
        self.forms = tuple(forms)
        self.depth = depth
        self.seed = seed
        self.records = {}
        if path and os.path.exists(path):
            with open(path) as f:
                self.records = json.load(f)

    @staticmethod
    def key(n, dtype, arch):
        return f"{arch}/{dtype}/{n}"

    def lookup(self, n, dtype, arch):
        return self.records.get(self.key(n, dtype, arch))

    def tune(self, state, arch, max_high=None):
        """Time every (k, form) candidate on `state` (a complex64 device tensor of 2^n amplitudes,
        overwritten) and keep the fastest.  complex128 states are not tuned: they run plain."""
        import torch
        n = state.numel().bit_length() - 1
        dt = "c64" if state.dtype == torch.complex64 else "c128"
        if dt != "c64":
            rec = dict(k=compiler.GPU_K, form="plain", ms=None, table={})
            self.records[self.key(n, dt, arch)] = rec
            return rec
        qc = circuit_mod.u3_cx_layers(n, self.depth, self.seed)
        stream = torch.cuda.current_stream()
        table = {}
        for k in self.ks:
            if k > n:
                continue
            for form in self.forms:
                prog, fmap, st = compile_form(qc, k, form, max_high)
                if any(kind != "segment" for kind, _ in prog):
                    continue
                with libsvx.fusion(form_fusion(form)):
                    hs = [libsvx.prepare(w, libsvx.C64, n) for _, w in prog]
                libsvx.zero_state(state)
                for h in hs:                      # warm-up (first launch of each kernel)
                    libsvx.run_prepared(state, h)
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                for h in hs:
                    libsvx.run_prepared(state, h)
                e1.record(stream)
                torch.cuda.synchronize()
                table[f"k{k}/{form}"] = round(e0.elapsed_time(e1), 3)
                for h in hs:
                    libsvx.release_prepared(h)
        best = min(table, key=table.get)
        kk, form = best.split("/")
        rec = dict(k=int(kk[1:]), form=form, ms=table[best], table=table,
                   calibration=f"u3cx layers, depth {self.depth}, seed {self.seed}")
        self.records[self.key(n, dt, arch)] = rec
        if self.path:
            with open(self.path, "w") as f:
                json.dump(self.records, f, indent=1, sort_keys=True)
        return rec


class SVRuntime:
    def __init__(self, device=0, n_engines=16, max_high=None, target="mi355x", dtype="c64", tuner=None):
        import torch
        self.torch = torch
        self.device = device
        self.caps = libsvx.init(device)
        self.arch = self.caps.arch.decode().split(":")[0]
        self.tuner = tuner
        self.form = "lib"
        self.n_engines = n_engines
        self.target = target
        self.max_high = max_high
        self.dtype = torch.complex64 if dtype == "c64" else torch.complex128
        self.st = None
        self.tmp = None
        self.log = []
        self.remap_mode = None      # "out_of_place" / "in_place", set by the self-test on first use

    # ------------------------------------------------------------------ planning
    def plan(self, n, p_max=None):
        """Reference slice planning (host/sv/runtime.py:158-171) for target="reference"."""
        if self.target != "reference":
            return 1, n, n
        cap = p_max or self.n_engines
        P = 1
        while P * 2 <= min(cap, self.n_engines) and n - (P * 2).bit_length() + 1 >= LPREF:
            P *= 2

        def fits(P):
            L = max(n, LMIN + P.bit_length() - 1) - (P.bit_length() - 1)
            return (1 << L) * AMP + (4 << 20) <= (512 << 20)
        while not fits(P) and P * 2 <= cap and n - (P * 2).bit_length() + 1 >= LMIN:
            P *= 2
        n_run = max(n, LMIN + P.bit_length() - 1)
        return P, n_run, n_run - (P.bit_length() - 1)

    def _held(self):
        return sum(t.numel() * t.element_size() for t in (self.st, self.tmp) if t is not None)

    def admit(self, n, dtype=None, buffers=2):
        """Footprint admission: `buffers` states of 2^n amplitudes (2: the state and an out-of-place
        remap buffer; 1: the state alone, remaps in place) fit in the free device memory plus what
        this runtime already holds."""
        dt = self.dtype if dtype is None else dtype
        es = 8 if dt == self.torch.complex64 else 16
        need = buffers * (1 << n) * es
        free, total = self.torch.cuda.mem_get_info(self.device)
        return need <= free + self._held()

    def _alloc(self, n):
        torch = self.torch
        if self.st is None or self.st.numel() != (1 << n) or self.st.dtype != self.dtype:
            if not self.admit(n, buffers=1):
                es = 8 if self.dtype == torch.complex64 else 16
                raise RuntimeError(f"state of 2^{n} amplitudes ({((1 << n) * es) >> 20} MiB) does not fit "
                                   "in device memory")
            self.st = None
            self.tmp = None
            self.st = torch.empty(1 << n, dtype=self.dtype, device=f"cuda:{self.device}")

    def _tmp(self):
        if self.tmp is None or self.tmp.numel() != self.st.numel() or self.tmp.dtype != self.st.dtype:
            self.tmp = None
            self.tmp = self.torch.empty_like(self.st)
        return self.tmp

    def self_test(self):
        """Both device permutation paths on a seeded 14-qubit state against the same permutation of
        the host copy (a bit reversal: data movement only, so equality is exact).  Sets `remap_mode`:
        out of place when that path passes, else in place when that one passes; RuntimeError when
        neither does."""
        torch = self.torch
        n = 14
        rng = np.random.default_rng(14)
        host = (rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)).astype(np.complex64)
        perm = list(range(n))[::-1]
        want = host.reshape([2] * n).transpose([n - 1 - perm[n - 1 - i] for i in range(n)]).reshape(-1)
        src = torch.from_numpy(host).to(f"cuda:{self.device}")
        ok = {}
        try:
            dst = torch.empty_like(src)
            self._permute_oop(dst, src, perm)
            ok["out_of_place"] = bool(np.array_equal(dst.cpu().numpy(), want))
        except RuntimeError as e:
            ok["out_of_place"] = False
            self.log.append(f"out-of-place permutation self-test raised: {e}")
        try:
            self._permute_inplace(src, perm)
            ok["in_place"] = bool(np.array_equal(src.cpu().numpy(), want))
        except RuntimeError as e:
            ok["in_place"] = False
            self.log.append(f"in-place permutation self-test raised: {e}")
        if ok["out_of_place"]:
            self.remap_mode = "out_of_place"
        elif ok["in_place"]:
            self.remap_mode = "in_place"
            self.log.append("out-of-place permutation self-test failed: in-place remaps")
        else:
            raise RuntimeError("device permutation self-test failed on both paths")
        return ok

    # the two device paths (separate methods so a failure can be injected in one of them)
    def _permute_oop(self, dst, src, perm, stream=None):
        libsvx.permute(dst, src, perm, stream=stream)

    def _permute_inplace(self, st, perm, stream=None):
        libsvx.permute(st, st, perm, stream=stream)

    def remap(self, perm, stream=None):
        """Apply perm[new] = old to the resident state: out of place when a second state fits, else
        (or after an out-of-place failure) in place."""
        if self.remap_mode is None:
            self.self_test()
        if self.remap_mode == "out_of_place":
            if self.tmp is None and not self.admit(self.st.numel().bit_length() - 1, self.st.dtype):
                self.log.append(f"second state of 2^{self.st.numel().bit_length() - 1} amplitudes does not fit: "
                                "in-place remap")
            else:
                try:
                    tmp = self._tmp()
                    self._permute_oop(tmp, self.st, perm, stream=stream)
                    self.st, self.tmp = tmp, self.st
                    return "out_of_place"
                except RuntimeError as e:   # includes torch.cuda.OutOfMemoryError
                    self.tmp = None
                    self.remap_mode = "in_place"
                    self.log.append(f"out-of-place remap failed ({e}): in-place remaps from here on")
        self._permute_inplace(self.st, perm, stream=stream)
        return "in_place"

    def compile(self, qc, k=None, fuse=True, tiling=True, initial_map=None, p_max=None):
        n = qc.num_qubits
        if self.target == "reference":
            P, n_run, L = self.plan(n, p_max)
            kk = min(k or KMAX, L)
            qrun = qc.padded(n_run) if n_run != n else qc
            mh = 4 if self.max_high is None else self.max_high
            prog, fmap, stats = compile_circuit(qrun, L, kk, fuse_1q=fuse, tiling=tiling,
                                                initial_map=initial_map, max_high=mh)
            return prog, fmap, stats, P, n_run, L, kk
        kk, form = k or compiler.GPU_K, "lib"
        dt = "c64" if self.dtype == self.torch.complex64 else "c128"
        rec = self.tuner.lookup(n, dt, self.arch) if self.tuner is not None else None
        if rec is not None and k is None:
            kk, form = rec["k"], rec["form"]
        if dt == "c128" or not fuse:
            form = "plain" if form in ("w4", "w5") else form
        self.form = form
        if form in ("w4", "w5"):
            prog, fmap, stats = compile_fused(qc, k=kk, width=int(form[1]), max_high=self.max_high, fuse_1q=fuse)
        else:
            prog, fmap, stats = compile_gpu(qc, k=kk, max_high=self.max_high, fuse_1q=fuse,
                                            cluster=0 if form == "plain" else compiler.GPU_CLUSTER)
        return prog, fmap, stats, 1, n, n, min(kk, n)

    def execute(self, prog, n_run, P, stream=None, timing=None):
        """Run a compiled program on the resident state (already initialised)."""
        torch = self.torch
        started = False
        T = timing if timing is not None else {}
        for kind, payload in prog:
            if kind == "remap" and not started:
                continue          # |0...0> is invariant under qubit permutations (host/sv/runtime.py:308-309)
            a = time.perf_counter()
            if kind == "segment":
                started = True
                with libsvx.fusion(form_fusion(self.form) if self.target == "mi355x" else "auto"):
                    libsvx.run_segment(self.st, payload, n_slices=P, stream=stream)
                torch.cuda.synchronize(self.device)
                T["T_kernel"] = T.get("T_kernel", 0.0) + time.perf_counter() - a
            else:
                self.remap(payload, stream=stream)
                torch.cuda.synchronize(self.device)
                T["T_exchange"] = T.get("T_exchange", 0.0) + time.perf_counter() - a
        return T

    def logical_state(self, fmap, n_run):
        """State in logical order (host/sv/runtime.py:327-330), permuted on the device (in place
        when a second state does not fit: the resident state is then left in logical order)."""
        if list(fmap) == list(range(n_run)):
            return self.st
        perm = [fmap[l] for l in range(n_run)]
        if self.remap_mode == "in_place" or (self.tmp is None and not self.admit(n_run, self.st.dtype)):
            self._permute_inplace(self.st, perm)
            return self.st
        tmp = self._tmp()
        self._permute_oop(tmp, self.st, perm)
        return tmp

    def run(self, qc, shots=1024, k=None, p_max=None, fuse=True, tiling=True, initial_map=None,
            return_state=True, seed=None):
        torch = self.torch
        n = qc.num_qubits
        t0 = time.perf_counter()
        prog, fmap, stats, P, n_run, L, kk = self.compile(qc, k, fuse, tiling, initial_map, p_max)
        t1 = time.perf_counter()
        self._alloc(n_run)
        t1b = time.perf_counter()
        libsvx.zero_state(self.st)
        torch.cuda.synchronize(self.device)
        t2 = time.perf_counter()
        T = dict(T_compile=t1 - t0, T_h2d=t2 - t1b, T_alloc=t1b - t1, T_kernel=0.0, T_exchange=0.0)
        self.execute(prog, n_run, P, timing=T)
        a = time.perf_counter()
        host = self.st.cpu().numpy()
        b = time.perf_counter()
        T["T_d2h"] = b - a
        counts = sample_probabilities(host, shots, fmap, n, seed)
        c = time.perf_counter()
        T["T_post"] = c - b
        T["T_e2e"] = T["T_compile"] + T["T_h2d"] + T["T_kernel"] + T["T_exchange"] + T["T_d2h"] + T["T_post"]
        out = dict(timing=T, P=P, L=L, k=kk, n_run=n_run, stats=stats.__dict__, counts=counts)
        if return_state:
            from_dev = self.logical_state(fmap, n_run).cpu().numpy()
            out["state"] = from_dev[: 1 << n]
            T["T_state_reorder"] = time.perf_counter() - c
        return out
