"""Served mix on one MI355X: a request loop over compiled circuits with continuous batching.

Requests carry compiled QSV2 programs (the client compiles; `compile_request`).  The server
- admits a request when its state footprint 2^n x 8 B fits the HBM budget (288 GB, or the device's
  memory, minus a reserve) and keeps its complex64 state resident until the result is read;
- runs circuits of up to 2^13 amplitudes as one whole-state pass (one workgroup holds the state in
  LDS for the whole circuit) on the library's persistent queue kernel (svx_queue_*: a resident launch
  takes each such circuit from a host-visible ring, many circuits per launch, no launch per circuit),
  and larger ones pass by pass;
- advances every active batched circuit by one pass per step with `libsvx.run_batch` (one launch per
  kernel shape, the tiles of all circuits side by side in one grid), so requests that arrive while
  others run join at the next step and their work shares the device;
- runs large circuits (>= `graph_min_qubits`) as a captured HIP graph of their whole pass sequence,
  replayed on a second stream next to the batched steps;
- reduces the 2^8 marginal probabilities of logical qubits 0..7 in FP64 on a third stream once a
  circuit's last pass has completed (host/sv/runtime.py:47-80 samples from the same marginals); the
  reduction and its read-back run on a collector thread, so the request loop keeps issuing steps
  while a large state is being read.

Results do not depend on the mix: a batched pass runs the same device code as a single run
(svx_run_batch), so each state is bit-identical alone and in the mix.
"""
import collections
import threading
import sys
import time
from dataclasses import dataclass, field

import numpy as np

from . import circuit as circuit_mod
from . import compiler, libsvx

WHOLE_STATE_MAX = 13     # 2^13 complex64 amplitudes = 64 KiB of LDS: one workgroup holds the state
N_MARGINAL = 8           # marginals of logical qubits 0..7 (256 probabilities)
HBM_BYTES = 288 * 10 ** 9


def compile_request(qc):
    """Client-side compilation: one whole-state pass up to 2^13 amplitudes, the MI355X tile stream
    (2^13-amplitude tiles, gathered high qubits) above.  Returns (prog, fmap, stats)."""
    n = qc.num_qubits
    if n <= WHOLE_STATE_MAX:
        return compiler.compile_gpu(qc, k=n, max_ops=1 << 30)
    return compiler.compile_gpu(qc)


@dataclass
class Request:
    rid: int
    n: int
    prog: list
    fmap: list
    arrival: float = 0.0               # seconds from the start of the stream
    depth: int = 0
    # server state
    steps: list = field(default_factory=list)     # [(handle, pass index)]
    state: object = None
    next_step: int = 0
    mode: str = ""                     # "batch" or "graph"
    graph: object = None
    done_event: object = None
    t_admit: float = None
    t_done: float = None
    marginals: np.ndarray = None
    norm: float = None
    queue: object = None               # the persistent queue running it (mode "queue")
    ticket: int = -1

    @property
    def footprint(self):
        return (1 << self.n) * 8


class SVServer:
    """Continuous-batching request loop (see the module docstring)."""

    def __init__(self, device=0, budget_bytes=None, reserve_bytes=16 << 30, graph_min_qubits=30,
                 max_inflight_steps=2, keep_states=False, queue_small=True):
        import torch
        self.torch = torch
        self.caps = libsvx.init(device)
        self.device = device
        total = min(int(self.caps.hbm_bytes), HBM_BYTES)
        self.budget = (total - reserve_bytes) if budget_bytes is None else int(budget_bytes)
        self.in_use = 0
        self.peak_in_use = 0
        self.graph_min_qubits = graph_min_qubits
        self.max_inflight = max_inflight_steps
        self.keep_states = keep_states
        self.queue_small = queue_small
        self.queues = {}          # state bits -> libsvx.Queue (persistent whole-state kernel)
        self.n_queued = 0
        self.s_batch = torch.cuda.Stream(device)
        self.s_graph = torch.cuda.Stream(device)
        self.s_red = torch.cuda.Stream(device)
        self.queue = []           # arrived, not admitted (FIFO)
        self.active = []          # batched requests with passes left
        self.graphs = []          # admitted graph requests not launched yet
        self.finishing = []       # last pass issued, result pending
        self.done = []
        self.step_events = []
        self.t0 = None
        self.n_steps = 0
        self.n_launch_items = 0
        # collector thread: FP64 reductions of finished states (one at a time: the library's
        # reduction scratch is shared), results handed back to the loop through `reduced`
        self._to_reduce = collections.deque()
        self.reduced = collections.deque()
        self._cv = threading.Condition()
        # the collector's HIP calls (reduction, read-back, scratch growth) and the loop thread's graph
        # captures never overlap: a capture is invalidated by another thread's allocating or
        # synchronising call on some HIP versions, whatever the capture mode
        self._hip_lock = threading.Lock()
        self._stop = False
        self._collector = None
        self.collector_error = None
        self._handed = 0
        self._taken = 0

    def _collector_loop(self):
        torch = self.torch
        torch.cuda.set_device(self.device)
        while True:
            with self._cv:
                while not self._to_reduce and not self._stop:
                    self._cv.wait()
                if not self._to_reduce:
                    return
                r = self._to_reduce.popleft()
            try:
                fm = list(r.fmap[:N_MARGINAL])
                with self._hip_lock:
                    if r.done_event is not None:
                        self.s_red.wait_event(r.done_event)
                    r.norm, r.marginals = libsvx.reduce_f64(r.state, fm, stream=self.s_red.cuda_stream)
                r.t_done = self.now()
            except Exception as e:   # reported by the loop thread
                self.collector_error = e
            self.reduced.append(r)

    def _start_collector(self):
        if self._collector is None:
            self._stop = False
            self._collector = threading.Thread(target=self._collector_loop, daemon=True)
            self._collector.start()

    def _stop_collector(self):
        if self._collector is not None:
            with self._cv:
                self._stop = True
                self._cv.notify()
            self._collector.join()
            self._collector = None

    def now(self):
        return time.perf_counter() - self.t0

    # ------------------------------------------------------------------ admission
    def submit(self, req):
        if req.footprint > self.budget:
            raise ValueError(f"request {req.rid}: a {req.n}-qubit state does not fit the HBM budget")
        self.queue.append(req)

    def _prepare(self, req):
        steps = []
        for kind, w in req.prog:
            if kind != "segment":
                raise ValueError("served programs are segments only (remaps are not batched)")
            h = libsvx.prepare(w, libsvx.C64, req.n)
            npass, _ = libsvx.prepared_info(h)
            steps += [(h, p) for p in range(npass)]
        req.steps = steps

    def _admit(self):
        torch = self.torch
        while self.queue and self.in_use + self.queue[0].footprint <= self.budget:
            req = self.queue.pop(0)
            self._prepare(req)
            req.state = torch.empty(1 << req.n, dtype=torch.complex64, device=f"cuda:{self.device}")
            self.in_use += req.footprint
            self.peak_in_use = max(self.peak_in_use, self.in_use)
            req.t_admit = self.now()
            if self._queueable(req):
                req.mode = "queue"
                q = self.queues.get(req.n)
                if q is None:
                    q = self.queues[req.n] = libsvx.Queue(req.n, capacity=4096, workgroups=64)
                req.queue = q
                req.ticket = q.submit(req.state, req.steps[0][0], zero_init=True)
                self.n_queued += 1
                self.finishing.append(req)
            elif req.n >= self.graph_min_qubits:
                req.mode = "graph"
                g = torch.cuda.CUDAGraph()
                cap = torch.cuda.Stream(self.device)
                # the collector thread makes no HIP call while this thread captures
                with self._hip_lock, torch.cuda.graph(g, stream=cap, capture_error_mode="thread_local"):
                    s = torch.cuda.current_stream().cuda_stream
                    libsvx.zero_state(req.state, stream=s)
                    for h, p in self._segments(req):
                        libsvx.run_prepared(req.state, h, stream=s)
                req.graph = g
                self.graphs.append(req)
            else:
                req.mode = "batch"
                with torch.cuda.stream(self.s_batch):
                    libsvx.zero_state(req.state, stream=self.s_batch.cuda_stream)
                self.active.append(req)

    def _queueable(self, req):
        """One whole-state pass on the MUX2 kernel, 2^10 .. 2^13 amplitudes: the queue kernel runs it."""
        if not self.queue_small or not 10 <= req.n <= WHOLE_STATE_MAX or len(req.steps) != 1:
            return False
        return libsvx.prepared_kernel(req.steps[0][0], 0).startswith("k_mux<")

    def close(self):
        for q in self.queues.values():
            q.close()
        self.queues = {}

    @staticmethod
    def _release(req):
        """Give back the request's prepared handles (their descriptors become evictable)."""
        for h in {h for h, _ in req.steps}:
            libsvx.release_prepared(h)
        req.steps = []

    @staticmethod
    def _segments(req):
        seen = []
        for h, p in req.steps:
            if not seen or seen[-1][0] != h:
                seen.append((h, p))
        return seen

    # ------------------------------------------------------------------ execution
    def _step(self):
        torch = self.torch
        # bound the steps in flight so that new arrivals join soon
        while len(self.step_events) >= self.max_inflight:
            self.step_events.pop(0).synchronize()
        if self.active:
            items = [(r.state, r.steps[r.next_step][0], r.steps[r.next_step][1]) for r in self.active]
            libsvx.run_batch([i[0] for i in items], [i[1] for i in items], [i[2] for i in items],
                             stream=self.s_batch.cuda_stream)
            self.n_steps += 1
            self.n_launch_items += len(items)
            ev = torch.cuda.Event()
            ev.record(self.s_batch)
            self.step_events.append(ev)
            still = []
            for r in self.active:
                r.next_step += 1
                if r.next_step == len(r.steps):
                    r.done_event = ev
                    self.finishing.append(r)
                else:
                    still.append(r)
            self.active = still
        for r in self.graphs:
            with torch.cuda.stream(self.s_graph):
                r.graph.replay()
                ev = torch.cuda.Event()
                ev.record(self.s_graph)
            r.done_event = ev
            self.finishing.append(r)
        self.graphs = []

    def _collect(self, block=False):
        """Hand finished circuits (last pass complete) to the collector thread, and take back the
        reduced ones: free their states and give back their handles.  `block`: wait for one."""
        left = []
        for r in self.finishing:
            if r.mode == "queue":
                ready = r.queue.done(r.ticket)
                if block and not ready:
                    r.queue.wait(r.ticket)
                    ready = True
            else:
                if block:
                    r.done_event.synchronize()
                ready = r.done_event.query()
            if ready:
                with self._cv:
                    self._to_reduce.append(r)
                    self._handed += 1
                    self._cv.notify()
            else:
                left.append(r)
        self.finishing = left
        if block:
            while not self.reduced and self.outstanding():
                time.sleep(0.0002)
        while self.reduced:
            r = self.reduced.popleft()
            self._taken += 1
            if self.collector_error is not None:
                raise self.collector_error
            if not self.keep_states:
                r.state = None
                self.in_use -= r.footprint
            r.graph = None            # replayed once; its handles are given back below
            self._release(r)
            self.done.append(r)
            if getattr(self, "progress_every", 0) and len(self.done) % self.progress_every == 0:
                print(f"[serve {self.now():7.1f} s] {len(self.done)} circuits done, {self.n_steps} batched steps",
                      file=sys.stderr, flush=True)

    def outstanding(self):
        """Circuits handed to the collector whose results the loop has not taken back yet."""
        return self._handed - self._taken

    def serve(self, requests, realtime=True, progress_every=0):
        """Run a request stream: each request is submitted at its arrival time (seconds from the
        start; all at once when realtime is False) and the loop steps until every result is in.
        progress_every > 0: one line on stderr each time that many more results are in.  Returns
        the completed requests."""
        self.progress_every = int(progress_every)
        pending = sorted(requests, key=lambda r: r.arrival)
        self.t0 = time.perf_counter()
        self._start_collector()
        try:
            self._loop(pending, realtime)
        finally:
            self._stop_collector()
            self.close()
        self.torch.cuda.synchronize(self.device)
        return self.done

    def _loop(self, pending, realtime):
        i = 0
        while (i < len(pending) or self.queue or self.active or self.graphs or self.finishing
               or self.outstanding()):
            t = self.now()
            while i < len(pending) and (not realtime or pending[i].arrival <= t):
                self.submit(pending[i])
                i += 1
            self._admit()
            if self.active or self.graphs:
                self._step()
            idle = not (self.active or self.graphs)
            # with nothing to step, wait for a result unless an arrival is due first
            self._collect(block=idle and bool(self.finishing or self.outstanding()) and i >= len(pending))
            if idle and i < len(pending):
                wait = pending[i].arrival - self.now()
                if wait > 0:
                    time.sleep(min(wait, 0.0005 if (self.finishing or self.outstanding()) else 0.002))
            if self.queue and not (self.active or self.graphs or self.finishing or self.outstanding()):
                if self.in_use + self.queue[0].footprint > self.budget:
                    raise RuntimeError("a queued request can never be admitted")


# ---------------------------------------------------------------------- request streams
def mix_circuit(rng, rid):
    """One request of the served mix: n ~ U{12..28} w.p. 0.95, U{30..32} w.p. 0.05; depth ~ U{10..60};
    the layer recipe with the circuit seed 3000 + 100 n + depth, so equal (n, depth) requests repeat
    a circuit (odd n leaves one qubit unpaired per layer)."""
    n = int(rng.integers(12, 29)) if rng.random() < 0.95 else int(rng.integers(30, 33))
    depth = int(rng.integers(10, 61))
    return circuit_mod.u3_cx_layers(n, depth, 3000 + 100 * n + depth), n, depth


def make_stream(n_requests=None, rate=None, duration=None, seed=3, max_qubits=32, compile_cache=None,
                progress_every=0):
    """Poisson arrivals at `rate` per second for `duration` seconds (or `n_requests` requests at
    time 0 when rate is None), each a compiled circuit of the mix.  Requests above max_qubits are
    redrawn (for boxes with less memory).  progress_every > 0: a line on stderr per that many
    requests made."""
    rng = np.random.default_rng(seed)
    reqs = []
    t = 0.0
    rid = 0
    while True:
        if rate is not None:
            t += float(rng.exponential(1.0 / rate))
            if t > duration:
                break
        elif rid >= n_requests:
            break
        qc, n, depth = mix_circuit(rng, rid)
        while n > max_qubits:
            qc, n, depth = mix_circuit(rng, rid)
        key = (n, depth)
        if compile_cache is not None and key in compile_cache:
            prog, fmap = compile_cache[key]
        else:
            prog, fmap, _ = compile_request(qc)
            if compile_cache is not None:
                compile_cache[key] = (prog, fmap)
        reqs.append(Request(rid=rid, n=n, prog=prog, fmap=fmap, arrival=t if rate is not None else 0.0,
                            depth=depth))
        rid += 1
        if progress_every and rid % progress_every == 0:
            print(f"[stream] {rid} requests compiled ({len(compile_cache or {})} distinct)", file=sys.stderr, flush=True)
    return reqs


def latency_stats(done):
    lat = np.array([r.t_done - r.arrival for r in done])
    span = max(r.t_done for r in done) - min(r.arrival for r in done)
    return dict(n=len(done), circuits_per_s=round(len(done) / span, 3) if span > 0 else None,
                p50_ms=round(float(np.percentile(lat, 50)) * 1e3, 2),
                p99_ms=round(float(np.percentile(lat, 99)) * 1e3, 2),
                max_ms=round(float(lat.max()) * 1e3, 2), span_s=round(span, 3))


def judge_c128(reqs, every=10, max_qubits=24):
    """Complex128 plain fallback on a sample of the served requests: the same records, every op on
    the vector units in FP64, marginals in FP64; returns the largest marginal difference."""
    import torch
    worst = 0.0
    count = 0
    for r in reqs[::every]:
        if r.n > max_qubits or r.marginals is None:
            continue
        s = torch.empty(1 << r.n, dtype=torch.complex128, device="cuda")
        libsvx.zero_state(s)
        with libsvx.fusion("off"):
            for _, w in r.prog:
                libsvx.run_segment(s, w)
        _, m = libsvx.reduce_f64(s, list(r.fmap[:N_MARGINAL]))
        worst = max(worst, float(np.max(np.abs(m - r.marginals))))
        count += 1
        del s
    return dict(sampled=count, max_abs_marginal_diff=worst)
