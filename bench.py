"""Benchmark: 30-qubit complex64 MUX2 pass bandwidth on one MI355X, fused vs plain time at 32 and
34 qubits.

Workload (value): the 30-qubit layer circuit (U3 on every qubit with angles U[0, 2pi), then CX on a
seeded matching in which each of qubits 26..29 is paired with one of 0..25; depth 20, seed 1),
compiled to plain MUX2 records by this project's compiler for the MI355X tile (2^13 amplitudes, up
to 7 gathered high qubits).  The library runs the records with its default pass fusion (each
pass's ops clustered into dense units on the matrix cores).  A step applies every pass once to the
8 GiB state resident in HBM (the state is a valid input for the next step, so nothing is
re-initialised inside the timed region).

value = passes x 2 x 2^30 x 8 B / time per step (the reference's sweep-bandwidth definition,
analysis/make_tables.py:320-321).  roofline: the pass kernel's algorithmic bytes per launch over
its mean duration from HIP events on the launch stream.  cpu_baseline: the numpy golden model
(oracle/model.py, the reference's CPU path) on the 20-qubit workload, one pinned core, in a
child process.  fused_vs_plain: the 32- and 34-qubit depth-40 workloads, plain MUX2 (vector
units only) against the autotuned form, with the complex128 plain run as judge at 32 qubits.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_GBS = 8000.0
_T0 = time.perf_counter()


def progress(what):
    """One line on stderr as each section of the run completes (the JSON line stays alone on stdout)."""
    print(f"[bench {time.perf_counter() - _T0:7.1f} s] {what}", file=sys.stderr, flush=True)


def cpu_baseline(max_seconds=30.0):
    """Child-process entry: time the golden model on the 20-qubit workload (reference stream:
    P = 8 slices of 2^17, k = 12, MAXH 4 -- the reference runtime's plan for 20 qubits)."""
    os.environ["OMP_NUM_THREADS"] = "1"
    os.environ["OPENBLAS_NUM_THREADS"] = "1"
    os.environ["MKL_NUM_THREADS"] = "1"
    core = sorted(os.sched_getaffinity(0))[0]
    os.sched_setaffinity(0, {core})
    import numpy as np
    import __graft_entry__ as ge
    ge.load_package()
    from u55c_statevector_30q_amd import circuit, compiler
    from oracle import model
    qc = circuit.workload_cpu20()
    t0 = time.perf_counter()
    prog, fmap, st = compiler.compile_circuit(qc, 17, 12, max_high=4)
    t_compile = time.perf_counter() - t0
    times = []
    t_start = time.perf_counter()
    for rep in range(4):
        a = time.perf_counter()
        model.run_program(prog, 20, 8, fmap, ctype=np.complex64)
        times.append(time.perf_counter() - a)
        if time.perf_counter() - t_start > max_seconds and rep >= 1:
            break
    timed = times[1:] if len(times) > 1 else times
    med = float(np.median(timed))
    sweeps = st.n_passes + st.n_remaps
    gbs = sweeps * 2 * (1 << 20) * 8 / med / 1e9
    # the complex128 variant of the same records (the judge's arithmetic), bounded the same way
    t128 = []
    t_start = time.perf_counter()
    for rep in range(3):
        a = time.perf_counter()
        model.run_program(prog, 20, 8, fmap, ctype=np.complex128)
        t128.append(time.perf_counter() - a)
        if time.perf_counter() - t_start > max_seconds / 2 and rep >= 1:
            break
    med128 = float(np.median(t128[1:] if len(t128) > 1 else t128))
    return dict(value=round(gbs, 4), unit="GB/s", cores=1, kind="port",
                sample=(f"golden model (numpy c64) on the 20-qubit workload (depth 20, seed 0): "
                        f"{st.n_ops} ops, {st.n_passes} passes + {st.n_remaps} remaps at P=8, L=17, k=12; "
                        f"median of {len(timed)} runs after 1 warm-up = {med:.3f} s; compile {t_compile:.3f} s; "
                        f"pinned to core {core} of {os.cpu_count()}"),
                c64_s=round(med, 4),
                c128={"value": round(sweeps * 2 * (1 << 20) * 16 / med128 / 1e9, 4), "unit": "GB/s",
                      "s": round(med128, 4), "runs": len(t128)})


def gpu_same_records(libsvx, circuit, compiler, torch, np, reps=5):
    """The cpu_baseline's records (20 qubits, reference stream: P = 8 slices of 2^17, k = 12, MAXH 4;
    passes and remaps) on the device: svx_run_segment with P slices and one permutation sweep per
    remap, complex64 and complex128, median of `reps` runs after one warm-up (HIP events)."""
    qc = circuit.workload_cpu20()
    prog, fmap, st = compiler.compile_circuit(qc, 17, 12, max_high=4)
    out = {}
    stream = torch.cuda.current_stream()
    for name, dt in (("c64", torch.complex64), ("c128", torch.complex128)):
        a = torch.empty(1 << 20, dtype=dt, device="cuda:0")
        b = torch.empty_like(a)
        ms = []
        for rep in range(reps + 1):
            libsvx.zero_state(a)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            cur, oth = a, b
            for kind, payload in prog:
                if kind == "segment":
                    libsvx.run_segment(cur, payload, n_slices=8)
                else:
                    libsvx.permute(oth, cur, payload)
                    cur, oth = oth, cur
            e1.record(stream)
            torch.cuda.synchronize()
            if rep:
                ms.append(e0.elapsed_time(e1))
        med = float(np.median(ms))
        sweeps = st.n_passes + st.n_remaps
        out[name] = {"ms": round(med, 3),
                     "gbs": round(sweeps * 2 * (1 << 20) * (8 if name == "c64" else 16) / (med * 1e-3) / 1e9, 1)}
    out["records"] = f"{st.n_passes} passes + {st.n_remaps} remaps, P=8, L=17, k=12 (the cpu_baseline's stream)"
    return out


U55C_FLAGSHIP = {"ghz": (141, 1.3), "qft": (102, 15.8), "rqc": (67, 13.3), "vqe": (80, 4.7)}


def families_section(runtime, circuit, torch, np, n=30, reps=3):
    """The reference's 30-qubit flagship families (GHZ, QFT, RQC depth 10, VQE 2 layers) through
    SVRuntime.run on the default FP32 path, timed like the reference's records (T_kernel = time of the
    segment launches, T_e2e = the run call; median of `reps` after one warm-up run), and the sweep
    bandwidth by the reference's definition passes x 2^30 x 16 B / T_kernel
    (analysis/make_tables.py:320-321).  The U55C's medians (paper/generated/macros.tex:33-57) are
    context: a different device, and the reference's plan has more passes and remaps."""
    rt = runtime.SVRuntime()
    out = {}
    for fam in ("ghz", "qft", "rqc", "vqe"):
        qc = circuit.make(fam, n)
        rt.run(qc, shots=1024, return_state=False, seed=1)
        runs = [rt.run(qc, shots=1024, return_state=False, seed=1) for _ in range(reps)]
        tk = float(np.median([r["timing"]["T_kernel"] for r in runs]))
        te = float(np.median([r["timing"]["T_e2e"] for r in runs]))
        stt = runs[0]["stats"]
        passes = stt["n_passes"]
        u_bw, u_tk = U55C_FLAGSHIP[fam]
        out[fam] = {"ops": stt["n_ops"], "passes": passes, "remaps": stt["n_remaps"],
                    "T_kernel_s": round(tk, 4), "T_e2e_s": round(te, 4),
                    "sweep_gbs": round(passes * (1 << 30) * 16 / tk / 1e9, 1),
                    "u55c_sweep_gbs": u_bw, "u55c_T_kernel_s": u_tk}
    del rt
    torch.cuda.empty_cache()
    return out


def latency_section(runtime, circuit, np, runs=2000, warmup=20):
    """Serial latency (bench/latency_jitter.py:24-58): RQC-16 (depth 10) through SVRuntime.run with
    1024 shots and no state return, `runs` timed calls after `warmup`; p50 / p99 / max of the
    end-to-end time and of T_kernel, and the spread (p99 - p50) / p50.  The U55C's record is
    p50 / p99 = 32.4 / 36.9 ms (paper/generated/macros.tex:114-115)."""
    rt = runtime.SVRuntime()
    qc = circuit.make("rqc", 16)
    e2e, ker = [], []
    for i in range(runs + warmup):
        t0 = time.perf_counter()
        r = rt.run(qc, shots=1024, return_state=False)
        t = time.perf_counter() - t0
        if i >= warmup:
            e2e.append(t * 1e3)
            ker.append(r["timing"]["T_kernel"] * 1e3)
    x, k = np.array(e2e), np.array(ker)
    p50, p99 = float(np.percentile(x, 50)), float(np.percentile(x, 99))
    return {"workload": f"rqc_16_d10 through SVRuntime.run, 1024 shots, {runs} runs after {warmup}",
            "e2e_p50_ms": round(p50, 3), "e2e_p99_ms": round(p99, 3), "e2e_max_ms": round(float(x.max()), 3),
            "spread_pct": round(100 * (p99 - p50) / p50, 1),
            "kernel_p50_ms": round(float(np.percentile(k, 50)), 3),
            "kernel_p99_ms": round(float(np.percentile(k, 99)), 3),
            "u55c_p50_p99_ms": [32.4, 36.9]}


def hbm_roof_section(libsvx, torch, reps=10):
    """Measured HBM roof: svx_convert of an 8 GiB complex64 state into another (a same-type copy:
    the library's streaming copy kernel, 16-byte non-temporal accesses), mean of `reps` launches."""
    a = torch.zeros(1 << 30, dtype=torch.complex64, device="cuda:0")
    b = torch.empty_like(a)
    stream = torch.cuda.current_stream()
    for _ in range(2):
        libsvx.convert(b, a)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for e0, e1 in ev:
        e0.record(stream)
        libsvx.convert(b, a)
        e1.record(stream)
    torch.cuda.synchronize()
    ms = [e0.elapsed_time(e1) for e0, e1 in ev]
    del a, b
    torch.cuda.empty_cache()
    m = sum(ms) / len(ms)
    return {"copy_bytes": 2 * (1 << 33), "mean_ms": round(m, 4), "best_ms": round(min(ms), 4),
            "gbs_mean": round(2 * (1 << 33) / (m * 1e-3) / 1e9, 1),
            "gbs_best": round(2 * (1 << 33) / (min(ms) * 1e-3) / 1e9, 1)}


def cp_ladder_section(libsvx, circuit, compiler, torch, np, n=30, seed=300):
    """Controlled-phase ladder (H(j), CP(pi/2^(k-j)) control k > j, j = 0..n-1) on a seeded random
    product state, complex64 and complex128 runs of the same records, error growth per pass, then
    the bit reversal perm[i] = n-1-i as one permutation sweep (checked bit-exact by reversing back)."""
    qc = circuit.cp_ladder(n)
    prog, fmap, st = compiler.compile_gpu(qc)
    assert fmap == list(range(n))
    th, ph = circuit.product_state_angles(n, seed)
    amps = np.stack([np.cos(th / 2), np.exp(1j * ph) * np.sin(th / 2)], axis=1)
    s64 = torch.empty(1 << n, dtype=torch.complex64, device="cuda:0")
    s128 = torch.empty(1 << n, dtype=torch.complex128, device="cuda:0")
    single = [seg for _, w in prog for seg in compiler.split_passes(w)]
    h64 = [libsvx.prepare(w, libsvx.C64, n) for w in single]
    h128 = [libsvx.prepare(w, libsvx.C128, n) for w in single]
    growth = []
    libsvx.product_state(s64, amps)
    libsvx.product_state(s128, amps)
    for i, (a, b) in enumerate(zip(h64, h128)):
        libsvx.run_prepared(s64, a)
        libsvx.run_prepared(s128, b)
        r = libsvx.fidelity_f64(s128, s64)
        growth.append([i + 1, r["one_minus_f"], r["max_abs_err"]])
    stream = torch.cuda.current_stream()
    times = {}
    for name, s, hs in (("c64", s64, h64), ("c128", s128, h128)):
        for h in hs:   # workgroup-to-tile order per pass by measurement (untimed)
            libsvx.tune_prepared(s, h)
        libsvx.product_state(s, amps)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for h in hs:
            libsvx.run_prepared(s, h)
        e1.record(stream)
        torch.cuda.synchronize()
        times[name] = e0.elapsed_time(e1)
    rev = list(range(n))[::-1]
    t64 = torch.empty_like(s64)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    libsvx.permute(t64, s64, rev)
    e0.record(stream)
    libsvx.permute(t64, s64, rev)
    e1.record(stream)
    torch.cuda.synchronize()
    perm_ms = e0.elapsed_time(e1)
    back = torch.empty_like(s64)
    libsvx.permute(back, t64, rev)
    exact = bool(torch.equal(back, s64))
    # the same reversal in place (no second state): a few tile-local sweeps
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    libsvx.permute(back, back, rev)
    e1.record(stream)
    torch.cuda.synchronize()
    inplace_ms = e0.elapsed_time(e1)
    inplace_exact = bool(torch.equal(back, t64))
    sweeps = len(libsvx.permute_plan(rev))
    t128 = torch.empty_like(s128)
    libsvx.permute(t128, s128, rev)
    final = libsvx.fidelity_f64(t128, t64)
    del s64, s128, t64, t128, back
    torch.cuda.empty_cache()
    nb = 2 * (1 << n) * 8
    return {"workload": f"cp ladder, {n} qubits: {sum(g.name == 'h' for g in qc.data)} H + "
                        f"{sum(g.name == 'cp' for g in qc.data)} CP, product state seed {seed}, then bit reversal",
            "passes": len(single), "c64_ms": round(times["c64"], 3), "c128_ms": round(times["c128"], 3),
            "c64_pass_gbs": round(len(single) * nb / (times["c64"] * 1e-3) / 1e9, 1),
            "bitrev_ms": round(perm_ms, 3), "bitrev_gbs": round(nb / (perm_ms * 1e-3) / 1e9, 1),
            "bitrev_roundtrip_bit_exact": exact,
            "bitrev_in_place": {"ms": round(inplace_ms, 3), "sweeps": sweeps,
                                "gbs_per_sweep": round(sweeps * nb / (inplace_ms * 1e-3) / 1e9, 1),
                                "equals_out_of_place": inplace_exact},
            "final_vs_c128": final,
            "error_growth_pass_1mF_maxabs": [[g[0], float(f"{g[1]:.3e}"), float(f"{g[2]:.3e}")] for g in growth]}


def fused_section(libsvx, circuit, runtime, torch, np, n, arch, judge=True):
    """Fused vs plain on the n-qubit fused-form workload (layer recipe, depth 40, seed 2),
    complex64: the plain MUX2 records with every op on the vector units against the autotuned
    form (tile width x {plain, library fusion, fused 4-/5-qubit records}, timed on a depth-4
    calibration circuit).  With `judge`, the complex128 run of the plain records judges the
    fused state (FP64 fidelity on the device)."""
    qc = circuit.workload_fused(n)
    free, _ = torch.cuda.mem_get_info(0)
    need = (1 << n) * (8 + (16 if judge else 0))
    if need > free * 0.95:
        return {"skipped": f"needs {need / 2**30:.0f} GiB of device memory, {free / 2**30:.0f} GiB free"}
    state = torch.empty(1 << n, dtype=torch.complex64, device="cuda:0")
    tuner = runtime.Autotuner()
    t0 = time.perf_counter()
    rec = tuner.tune(state, arch)
    t_tune = time.perf_counter() - t0
    stream = torch.cuda.current_stream()

    def timed(k, form, dense_math="f32"):
        prog, fmap, st = runtime.compile_form(qc, k, form)
        with libsvx.options(fuse=runtime.form_fusion(form), dense_math=dense_math):
            hs = [libsvx.prepare(w, libsvx.C64, n) for _, w in prog]
        libsvx.zero_state(state)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for h in hs:
            libsvx.run_prepared(state, h)
        e1.record(stream)
        torch.cuda.synchronize()
        for h in hs:
            libsvx.release_prepared(h)
        return e0.elapsed_time(e1), st, fmap, prog

    plain_ms, st_p, fmap_p, prog_p = timed(13, "plain")
    split_ms = None
    if rec["form"] != "plain":
        split_ms, _, _, _ = timed(rec["k"], rec["form"], "split16")   # opt-in split-FP16 dense units
    fused_ms, st_f, fmap_f, prog_f = timed(rec["k"], rec["form"])
    out = {"workload": f"u3cx_{n}q_d40_s2: {n} qubits, complex64, depth 40, seed 2 ({st_p.n_ops} MUX2 ops)",
           "plain_ms": round(plain_ms, 1), "plain_passes": st_p.n_passes,
           "fused_ms": round(fused_ms, 1), "fused_passes": st_f.n_passes,
           "fused_dense_records": getattr(st_f, "n_dense", 0),
           "speedup": round(plain_ms / fused_ms, 3),
           "fused_split16_optin_ms": None if split_ms is None else round(split_ms, 1),
           "tuned": {"k": rec["k"], "form": rec["form"], "tune_s": round(t_tune, 1),
                     "calibration_ms": rec["table"]}}
    norm, marg_hi = libsvx.reduce_f64(state, list(range(n - 1, n - 9, -1)))
    out["fused_norm_f64"] = norm
    if not judge:
        # bit reversal of the whole state in place (no second state fits beside it at 34 qubits):
        # the marginals of qubits n-1..n-8 before equal those of qubits 0..7 after (FP64 sums in
        # another order: within 1e-12), and reversing twice restores the reductions bit for bit
        rev = list(range(n))[::-1]
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        libsvx.permute(state, state, rev)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        _, marg_lo = libsvx.reduce_f64(state, list(range(8)))
        libsvx.permute(state, state, rev)
        norm2, marg_hi2 = libsvx.reduce_f64(state, list(range(n - 1, n - 9, -1)))
        sweeps = len(libsvx.permute_plan(rev))
        out["bitrev_in_place"] = {"ms": round(ms, 2), "sweeps": sweeps,
                                  "gbs_per_sweep": round(sweeps * 2 * (1 << n) * 8 / (ms * 1e-3) / 1e9, 1),
                                  "marginal_match": float(np.max(np.abs(marg_lo - marg_hi))),
                                  "twice_restores_bit_exact": bool(norm2 == norm and np.array_equal(marg_hi2, marg_hi))}
    if judge:
        assert fmap_p == fmap_f == list(range(n)), "the layer recipe compiles without relabelling"
        s128 = torch.empty(1 << n, dtype=torch.complex128, device="cuda:0")
        libsvx.zero_state(s128)
        for _, w in prog_p:
            libsvx.run_segment(s128, w)
        out["fused_vs_c128_plain"] = libsvx.fidelity_f64(s128, state)
        del s128
    del state
    torch.cuda.empty_cache()
    return out


def served_section(serving, torch, seconds=20.0, seed=3, burst=60, progress_every=0):
    """Served mix: capacity from a calibration burst (requests all at time 0, another seed), then a
    Poisson stream at 75% of it for `seconds`; circuits/s, p50/p99 latency (arrival to FP64
    marginals on the host), and the complex128 plain fallback as judge on a sample."""
    cache = {}
    t0 = time.perf_counter()
    cal = serving.make_stream(n_requests=burst, seed=seed + 1000, compile_cache=cache)
    srv = serving.SVServer()
    done = srv.serve(cal, realtime=False)
    cap = serving.latency_stats(done)["circuits_per_s"]
    rate = 0.75 * cap
    reqs = serving.make_stream(rate=rate, duration=seconds, seed=seed, compile_cache=cache,
                               progress_every=10 * progress_every)
    t_compile = time.perf_counter() - t0
    if progress_every:
        progress(f"served mix: {len(reqs)} requests over {seconds:.0f} s at {rate:.2f}/s, compiled in {t_compile:.1f} s")
    srv = serving.SVServer()
    done = srv.serve(reqs, realtime=True, progress_every=progress_every)
    st = serving.latency_stats(done)
    judge = serving.judge_c128(done, every=10, max_qubits=24)
    big = [r for r in done if r.n >= 30]
    return {"workload": f"Poisson stream, seed {seed}, {seconds:.0f} s at {rate:.2f}/s (75% of the {cap:.2f}/s "
                        f"measured on a {burst}-request burst); n ~ U{{12..28}} w.p. 0.95, U{{30..32}} w.p. 0.05; "
                        f"depth ~ U{{10..60}}; marginals of logical qubits 0..7 in FP64",
            "rate_per_s": round(rate, 3), "capacity_burst_per_s": cap, **st,
            "batched_steps": srv.n_steps, "mean_items_per_step": round(srv.n_launch_items / max(1, srv.n_steps), 2),
            "graph_requests": len(big), "peak_resident_gib": round(srv.peak_in_use / 2 ** 30, 2),
            "judge_c128_sampled": judge, "client_compile_s": round(t_compile, 1)}


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    (profiles/traffic_*.md, scaled by the copy-kernel calibration there), or None."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "traffic_*.md"))):
        fields = {}
        with open(path) as f:
            for line in f:
                if ":" in line and line.split(":")[0] in ("traffic_bytes_per_launch", "kernel"):
                    k, v = line.split(":", 1)
                    fields[k] = v.strip()
        if fields.get("kernel") == kernel and "traffic_bytes_per_launch" in fields:
            return int(fields["traffic_bytes_per_launch"]), os.path.relpath(path, ROOT)
    return None, None


def _run_cpu_child():
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-baseline-child"],
                       capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        return dict(error=r.stderr[-400:])
    return json.loads(r.stdout.strip().splitlines()[-1])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-judge", action="store_true")
    ap.add_argument("--no-fused", action="store_true", help="skip the 32/34-qubit fused-vs-plain section")
    ap.add_argument("--no-serve", action="store_true", help="skip the served-mix section")
    ap.add_argument("--no-families", action="store_true", help="skip the 30-qubit families section")
    ap.add_argument("--no-latency", action="store_true", help="skip the RQC-16 serial-latency section")
    ap.add_argument("--latency-runs", type=int, default=2000)
    ap.add_argument("--serve-seconds", type=float, default=20.0)
    ap.add_argument("--serve-only", action="store_true",
                    help="only the served-mix section (e.g. --serve-seconds 600 for the 10-minute stream)")
    ap.add_argument("--plain", action="store_true",
                    help="with --traffic-probe: the workload's passes unfused (every op on the MUX2 kernel)")
    ap.add_argument("--cpu-baseline-child", action="store_true")
    ap.add_argument("--ladder-probe", action="store_true",
                    help="for rocprofv3 kernel traces: the 30-qubit controlled-phase ladder's complex64 passes "
                         "(tile orders tuned first), once, then the bit reversal twice out of place and once in "
                         "place; prints per-pass work")
    ap.add_argument("--fused-probe", type=int, default=0, metavar="N",
                    help="for rocprofv3 kernel traces: the N-qubit fused-form workload (depth 40, seed 2), "
                         "complex64, run once in each of three forms at k = 13: plain MUX2 records (fusion "
                         "off), MUX2 records with library fusion (the form the autotuner picks at 32 and 34 "
                         "qubits) and fused 4-qubit records; prints the pass counts and the per-pass kernels")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="for rocprofv3 --pmc runs: an 8 GiB complex64 copy (svx_convert, calibration) twice, "
                         "the tile-order tuning launches (on the copy's destination), then one step of the "
                         "30-qubit workload; prints the kernel names")
    a = ap.parse_args()
    if a.cpu_baseline_child:
        print(json.dumps(cpu_baseline()))
        return
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (single-GPU build)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import __graft_entry__ as ge
    ge.load_package()
    from u55c_statevector_30q_amd import circuit, compiler, libsvx, runtime, serving

    caps = libsvx.init(0)
    if a.serve_only:
        print(json.dumps({"served_mix": served_section(serving, torch, seconds=a.serve_seconds,
                                                       progress_every=250 if a.serve_seconds > 60 else 0)}))
        return
    n = 30
    qc = circuit.workload_gpu30()
    if a.ladder_probe:
        lq = circuit.cp_ladder(n)
        prog, fmap, st = compiler.compile_gpu(lq)
        single = [seg for _, w in prog for seg in compiler.split_passes(w)]
        hs = [libsvx.prepare(w, libsvx.C64, n) for w in single]
        s64 = torch.empty(1 << n, dtype=torch.complex64, device="cuda:0")
        t64 = torch.empty_like(s64)
        libsvx.zero_state(s64)
        for h in hs:
            libsvx.tune_prepared(s64, h)
        th, ph = circuit.product_state_angles(n, 300)
        libsvx.product_state(s64, np.stack([np.cos(th / 2), np.exp(1j * ph) * np.sin(th / 2)], axis=1))
        for h in hs:
            libsvx.run_prepared(s64, h)
        for _ in range(2):
            libsvx.permute(t64, s64, list(range(n))[::-1])
        libsvx.permute(t64, t64, list(range(n))[::-1])   # in place: a planned sequence of tile-local sweeps
        torch.cuda.synchronize()
        work = [dict(libsvx.prepared_pass_stats(h, 0), kernel=libsvx.prepared_kernel(h, 0)) for h in hs]
        print(json.dumps({"passes": len(hs), "per_pass": work,
                          "in_place_sweeps": len(libsvx.permute_plan(list(range(n))[::-1]))}))
        return
    if a.fused_probe:
        nf = a.fused_probe
        qf = circuit.workload_fused(nf)
        sf = torch.empty(1 << nf, dtype=torch.complex64, device="cuda:0")
        res = {}
        for form in ("plain", "lib", "w4"):
            prog, fmap, st = runtime.compile_form(qf, 13, form)
            with libsvx.options(fuse=runtime.form_fusion(form)):
                hs = [libsvx.prepare(w, libsvx.C64, nf) for _, w in prog]
            libsvx.zero_state(sf)
            for h in hs:
                libsvx.run_prepared(sf, h)
            torch.cuda.synchronize()
            kinds = {}
            for h in hs:
                for p in range(libsvx.prepared_info(h)[1]):
                    kn = libsvx.prepared_kernel(h, p)
                    kinds[kn] = kinds.get(kn, 0) + 1
            res[form] = {"passes": st.n_passes, "launches": sum(kinds.values()), "kernels": kinds,
                         "norm_f64": libsvx.reduce_f64(sf, [])[0]}
            for h in hs:
                libsvx.release_prepared(h)
        print(json.dumps({"fused_probe": nf, **res}))
        return
    if a.traffic_probe:
        prog, fmap, st = compiler.compile_gpu(qc, cluster=0 if a.plain else compiler.GPU_CLUSTER)
        with libsvx.options(**(dict(fuse="off") if a.plain else {})):
            hs = [libsvx.prepare(w, libsvx.C64, n) for _, w in prog]
        src = torch.empty(1 << n, dtype=torch.complex64, device="cuda:0")
        dst = torch.empty_like(src)
        libsvx.zero_state(src)
        for _ in range(2):
            libsvx.convert(dst, src)
        for h in hs:
            libsvx.tune_prepared(dst, h)
        for h in hs:
            libsvx.run_prepared(src, h)
        torch.cuda.synchronize()
        names = sorted({libsvx.prepared_kernel(h, p) for h in hs for p in range(libsvx.prepared_info(h)[1])})
        print(json.dumps({"copy_bytes_read": (1 << n) * 8, "copy_bytes_written": (1 << n) * 8,
                          "pass_kernels": names, "passes": st.n_passes}))
        return
    t0 = time.perf_counter()
    prog, fmap, st = compiler.compile_gpu(qc)
    t_compile = time.perf_counter() - t0
    assert all(kind == "segment" for kind, _ in prog), "the 30-qubit workload compiles without remaps"
    handles = [libsvx.prepare(w, libsvx.C64, n) for _, w in prog]
    n_launch = sum(libsvx.prepared_info(h)[1] for h in handles)
    state = torch.empty(1 << n, dtype=torch.complex64, device="cuda:0")
    libsvx.zero_state(state)
    stream = torch.cuda.current_stream()
    # each pass's workgroup-to-tile order by measurement (untimed; bit-identical results), then |0..0>
    for h in handles:
        libsvx.tune_prepared(state, h)
    libsvx.zero_state(state)
    orders = [libsvx.prepared_pass_stats(h, p)["tile_order"] for h in handles
              for p in range(libsvx.prepared_info(h)[1])]

    def step():
        for h in handles:
            libsvx.run_prepared(state, h)

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step()
    torch.cuda.synchronize()
    t_total = time.perf_counter() - t0
    ms_per_step = t_total / a.steps * 1e3
    bytes_per_pass = 2 * (1 << n) * 8
    value = n_launch * bytes_per_pass / (t_total / a.steps) / 1e9

    # per-pass durations: each pass as its own prepared segment, HIP events on the launch stream
    single = [libsvx.prepare(seg, libsvx.C64, n) for _, w in prog for seg in compiler.split_passes(w)]
    for h in single:
        libsvx.tune_prepared(state, h)
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in single]
    for (e0, e1), h in zip(evs, single):
        e0.record(stream)
        libsvx.run_prepared(state, h)
        e1.record(stream)
    torch.cuda.synchronize()
    per_pass_ms = [e0.elapsed_time(e1) for e0, e1 in evs]
    progress(f"30-qubit workload: {ms_per_step:.3f} ms per step, {value:.1f} GB/s; per-pass timing done")
    kernels = [libsvx.prepared_kernel(h, 0) for h in single]
    by_kernel = {}
    for kname, t in zip(kernels, per_pass_ms):
        by_kernel.setdefault(kname, []).append(t)
    dominant = max(by_kernel, key=lambda kk: sum(by_kernel[kk]))
    mean_ms = float(np.mean(by_kernel[dominant]))
    achieved = bytes_per_pass / (mean_ms * 1e-3) / 1e9
    per_pass_gbs = [bytes_per_pass / (t * 1e-3) / 1e9 for t in per_pass_ms]

    traffic, traffic_src = profiled_traffic(dominant)
    judge = None
    if not a.no_judge:
        s64 = torch.empty(1 << n, dtype=torch.complex64, device="cuda:0")
        s128 = torch.empty(1 << n, dtype=torch.complex128, device="cuda:0")
        for s in (s64, s128):
            libsvx.zero_state(s)
            for _, w in prog:
                libsvx.run_segment(s, w)
        torch.cuda.synchronize()
        judge = libsvx.fidelity_f64(s128, s64)
        norm, _ = libsvx.reduce_f64(s64)
        judge["norm_c64"] = norm
        del s64, s128
        torch.cuda.empty_cache()

    progress(f"judge vs complex128: {judge}")
    ladder = None if a.no_judge else cp_ladder_section(libsvx, circuit, compiler, torch, np)
    progress("controlled-phase ladder done")
    arch = caps.arch.decode().split(":")[0]
    del state
    torch.cuda.empty_cache()
    fused = None
    if not a.no_fused:
        fused = {}
        for nq in (32, 34):
            fused[str(nq)] = fused_section(libsvx, circuit, runtime, torch, np, nq, arch,
                                           judge=(nq == 32 and not a.no_judge))
            progress(f"fused vs plain at {nq} qubits done")
    families = None if a.no_families else families_section(runtime, circuit, torch, np)
    progress("30-qubit families done")
    latency = None if a.no_latency else latency_section(runtime, circuit, np, runs=a.latency_runs)
    progress("serial latency done")
    roof = hbm_roof_section(libsvx, torch)
    served = None if a.no_serve else served_section(serving, torch, seconds=a.serve_seconds)
    progress("served mix done")
    cpu = None
    if not a.no_cpu:
        cpu = _run_cpu_child()
        progress("cpu baseline done")
        cpu["gpu_same_records"] = gpu_same_records(libsvx, circuit, compiler, torch, np)
    out = {
        "metric": "30q c64 pass GB/s (% of 8 TB/s); fused vs plain time; served circuits/s+p99",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {"workload": "u3cx_30q_d20_s1: 30 qubits, complex64, depth 20, seed 1, U3 + CX matching "
                               "(26..29 paired with 0..25)",
                   "n_qubits": n, "ops": st.n_ops, "passes": st.n_passes, "launches": n_launch,
                   "tile_k": compiler.GPU_K, "compile_s": round(t_compile, 3),
                   "schedule": "clusters" if st.clustered else "ops",
                   "tile_order_tuned": {"xcd_contiguous": orders.count(-2), "identity": orders.count(0)}},
        "frac_of_peak": round(value / PEAK_GBS, 4),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": dominant, "bytes_per_launch": bytes_per_pass,
                     "mean_launch_ms": round(mean_ms, 4), "launches": len(by_kernel[dominant]),
                     "per_kernel": {kk: {"launches": len(v), "mean_ms": round(float(np.mean(v)), 4),
                                         "total_ms": round(float(np.sum(v)), 3)} for kk, v in by_kernel.items()}},
        "per_pass_gbs": {"min": round(min(per_pass_gbs), 1), "median": round(float(np.median(per_pass_gbs)), 1),
                         "max": round(max(per_pass_gbs), 1),
                         "n_at_or_above_70pct": int(sum(g >= 0.7 * PEAK_GBS for g in per_pass_gbs)),
                         "n": len(per_pass_gbs)},
        "judge_vs_c128": judge,
        "cp_ladder_30q": ladder,
        "fused_vs_plain": fused,
        "families_30q": families,
        "serial_latency_rqc16": latency,
        "hbm_roof_copy": roof,
        "served_mix": served,
        "cpu_baseline": cpu,
        "device": caps.arch.decode(),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
