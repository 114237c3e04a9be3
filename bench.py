"""Benchmark: microseconds per exact modular squaring of a Proth residue on
one MI355X, with the residue resident in HBM.

Workload (default): the n = 2^21 single large test -- k drawn odd in
[3, 2^20) from PCG64(seed=2) (gate survivors only), base a by the reference's
rule, seed a^k mod N built on the device.  A "step" is one squaring
x <- x^2 mod N.  Before anything is timed the launch-plan autotuner picks
the plan (shape, radices, outer pass) for the workload's length.  W untimed
warm-up squarings run on the same context from the seed (the default, 512,
is two 256-squaring chunks: the first runs partly in normal form while the
seed's exact value is small, the second is whole and captures the chunk
graph before the clock starts); then exactly the next K squarings are timed
with the device synchronised on both sides.  Res64 after 10^4 squarings from
the seed (an untimed second run) is checked against the GMP fixture.

The same line reports the n = 2^24 test (1024 squarings timed, Res64 checked
after 10^3), the per-kernel roofline (HIP events on the engine's stream), the
batched (config 4, scaled) and serving (config 5, scaled) legs, and the CPU
legs (the reference's own sources and the C restatement, timed on this
host's cores for bounded samples).

  python bench.py [--gpus 1] [--steps 10000] [--warmup 512] [--no-cpu] [--no-tune]
  python bench.py --c4-full OUT.json     config 4 at its stated size
  python bench.py --c5-full S NS RATE OUT.json    segment S of NS of config 5 at its stated size
  python bench.py --plain-keys OUT.json  the plain loop's time per (length, carry run) key of configs 4/5
  python bench.py --c5-combine PLAIN.json GMP.jsonl SEG.json...   summary of config-5 segments (CPU only)
  python bench.py --cpu-only             the CPU legs alone
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
METRIC = "us/squaring at n=2^21 and 2^24; batched tests/hour vs plain loop, Res64 exact"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def golden_res64(k, n, i):
    p = ROOT / "tests" / "golden" / "large_cases.json"
    if not p.exists():
        return None
    for c in json.loads(p.read_text())["cases"]:
        if c["k"] == k and c["n"] == n and str(i) in c["res64_after"]:
            return int(c["res64_after"][str(i)], 16)
    return None


def seed_context(core, k, n):
    """Context holding a^k mod N, powered on the device (left-to-right)."""
    a = core.choose_base(k, n)
    c = core.Context(k, n)
    c.set_residue(a)
    for bit in bin(k)[3:]:
        c.square(1)
        if bit == "1":
            c.mul_small(a)
    return c, a


def time_squarings(core, k, n, steps, warmup, verify_at=None):
    """Times `steps` squarings that follow `warmup` untimed ones from the
    seed a^k (the warm-up takes the test past its first squarings, whose
    exact values are still small and run in normal form, and captures the
    chunk graph); Res64 is read after them, or, with verify_at, after that
    many squarings from the seed in a second, untimed run (a count the GMP
    fixtures hold)."""
    c, a = seed_context(core, k, n)
    seed = c.get_residue()
    if warmup:
        c.square(warmup)
    c.square(0)
    c.loop_time(reset=True)
    t0 = time.perf_counter()
    emax = c.square(steps)
    t1 = time.perf_counter()
    # the resident launches of the timed squarings, timed by HIP events
    # around each launch on the context's stream as they ran
    loop = c.loop_time(reset=True)
    if verify_at is not None:
        c.set_residue(seed)
        emax = max(emax, c.square(verify_at))
    r64 = c.res64()
    kt, _ = c.kernel_times(reps=20)
    info = dict(fft_len=c.fft_len, rows=c.info.rows, cols=c.info.cols, bits_per_digit=round(c.info.bits_per_digit, 4),
                kernels_per_squaring=c.info.passes, row_radix=c.info.row_radix, col_radix=c.info.col_radix,
                run=c.info.run)
    c.close()
    return dict(seconds=t1 - t0, max_err=emax, res64=r64, kernel_ms=kt, a=a, nt=info["run"], loop=loop, **info)


def _plan_str(t):
    return (f"{t['rows']}x{t['cols']} r{t['row_radix']}/c{t['col_radix']} run{t['run']} "
            f"{'one-launch' if t['one_launch'] else 'split'}")


TUNE = True  # bench.py --no-tune clears it


def tune_single(core, k, n):
    """The autotuner's run for a single context of (k, n) (untimed set-up,
    as the reference's auto_tune_fft_len runs before its tests): every
    trial's plan and device time, and the choice later contexts take."""
    if not TUNE:
        return dict(chosen="not tuned (--no-tune)")
    r = core.autotune_plan(k, n)
    return dict(chosen=_plan_str(r["best"]), chosen_us=round(r["best"]["us_per_squaring"], 3),
                trials={_plan_str(t): round(t["us_per_squaring"], 3) for t in r["trials"]},
                all_exact=all(t["exact"] for t in r["trials"]))


def tune_batch(core, pool, cands):
    """Untimed set-up of a batch: the pool's group plans tuned for every
    length group the candidates form (keys already tuned are not timed
    again), and the single-context plan of each transform length (the plain
    loop the batch is compared with)."""
    groups, plain = {}, {}
    if not TUNE:
        return dict(note="not tuned (--no-tune)")
    for k, n in cands:
        r = pool.autotune(k, n)
        if r["trials"]:
            groups[f"L={core.plan_fft_len(n, k)} run{r['trials'][0]['run']}"] = _plan_str(r["best"])
        L = core.plan_fft_len(n, k)
        if L not in plain:
            plain[L] = tune_single(core, k, n)["chosen"]
    return dict(pool_groups=groups, plain_contexts={str(L): v for L, v in sorted(plain.items())})


def kernel_bytes(L: int, nt: int) -> dict:
    """Algorithmic HBM bytes one launch of each kernel moves for a length-L
    residue: complex FP64 points are 16 B (read and/or written once); the
    split outer pass hands 4 B int32 digits and 24 B of run carries per run
    of nt digits from its first to its second kernel; the one-launch outer
    pass reads and writes the points and writes and reads the 24 B of run
    carries."""
    run = 24.0 / nt if nt else 0.0
    return {"fwd_col": 32 * L, "row_square": 32 * L, "inv_col": 32 * L, "carry": 32 * L,
            "inv_carry": int((16 + 4 + run) * L), "adjust_fwd": int((4 + run + 16) * L),
            "carry_turn": int((32 + 2 * run) * L)}


TURN = ("carry_turn", "row_square")
STEADY = ("inv_carry", "adjust_fwd", "row_square")
UNFUSED = ("inv_col", "carry", "fwd_col", "row_square")


def steady_kernels(kt: dict, nt: int):
    """The kernels of one steady-state squaring, as the context ran them."""
    if not nt:
        return UNFUSED
    return TURN if kt.get("carry_turn", 0) else STEADY


def roofline(r: dict):
    """Dominant steady-state kernel's achieved algorithmic GB/s vs the HBM peak.

    With the resident loop (passes 0) that kernel is k_loop: one launch per
    chunk carries every steady squaring of the chunk (the outer pass and the
    row pass of each, 70*L bytes per squaring with runs of 8), and its
    duration is the HIP-event time around its launches in the timed region.
    Otherwise it is the slower of the two-kernel steady state's kernels, timed
    by pl_kernel_times."""
    kt, L, nt = r["kernel_ms"], r["fft_len"], r["nt"]
    kb = kernel_bytes(L, nt)
    names = steady_kernels(kt, nt)
    per_sq = sum(kb[k] for k in names)
    lp = r.get("loop") or {}
    if lp.get("launches"):
        dom = "k_loop"
        sq_per_launch = lp["squarings"] / lp["launches"]
        dom_bytes = int(round(per_sq * sq_per_launch))
        ms = lp["ms"] / lp["launches"]
        kt = dict(kt, k_loop=ms, k_loop_per_squaring=lp["ms"] / lp["squarings"])
        r["kernel_ms"] = kt
    else:
        dom = max(names, key=lambda k: kt[k])
        dom_bytes = kb[dom]
        ms = kt[dom]
    achieved = dom_bytes / (ms * 1e-3) / 1e9
    return dom, dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                     frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None), dom_bytes, per_sq


def fp64_rate(L: int, us_per_squaring: float) -> dict:
    """Achieved FP64 rate of one squaring against the vector peak: about
    2*5*L*log2(L) flops for the forward and inverse transforms plus 20*L for
    weights, pointwise square and rounding."""
    flops = 2 * 5 * L * (L.bit_length() - 1) + 20 * L
    tf = flops / (us_per_squaring * 1e-6) / 1e12
    return dict(flops_per_squaring=flops, achieved_tflops=round(tf, 3), peak_tflops=FP64_PEAK_TFLOPS,
                frac=round(tf / FP64_PEAK_TFLOPS, 4))


def measured_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel`, from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE per dispatch, scaled by a copy kernel of known
    size); None when no summary exists for this kernel and workload."""
    p = ROOT / "profiles" / f"traffic_{kernel}_{workload}.txt"
    if not p.exists():
        return None
    for line in p.read_text().splitlines():
        if line.startswith("bytes_per_launch:"):
            try:
                return int(line.split(":", 1)[1].strip())
            except ValueError:
                return None
    return None


def _cpu_length(k, n):
    """The reference's starting length for n (n/L <= 18, auto_tune_fft_len),
    raised until its limb layout holds the residue."""
    sys.path.insert(0, str(ROOT / "oracle"))
    from oracle import Oracle
    o = Oracle()
    L = 16
    while not o.fits(L, n, k) or n / L > 18.0:
        L *= 2
    return L


def cpu_sample(kind, k, n, seconds):
    """One process's bounded CPU sample: squarings of a random residue mod
    k*2^n+1 with fft_square_karatsuba for about `seconds`.  kind: reference
    (oracle/_ref, the reference's own sources), reference-omp (the same built
    with -fopenmp; OMP_NUM_THREADS threads) or port (the C restatement)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import random
    import numpy as np
    from oracle import Oracle, RefLib, int_to_limbs, _ptr
    L = _cpu_length(k, n)
    nl = L // 4
    N = k * (1 << n) + 1
    w = np.zeros(nl + 4, dtype=np.uint64)
    w[:nl] = int_to_limbs(random.Random(1).randrange(N), nl)
    if kind == "port":
        oc = Oracle().ctx(L)
        step = lambda: oc.o.lib.or_square(oc.h, _ptr(w), nl, n, k)
    else:
        ref = RefLib(ROOT / "oracle" / "_ref" / ("libprothref_omp.so" if kind == "reference-omp" else "libprothref.so"))
        h = ref.ctx(L)
        step = lambda: ref.lib.ref_fft_square(h, _ptr(w), nl, n, k)
    step()
    cnt = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        step()
        cnt += 1
    return dict(count=cnt, seconds=time.perf_counter() - t0, fft_len=L)


def _cpu_run(kind, k, n, seconds, procs=1, threads=1):
    """`procs` concurrent child processes, each running cpu_sample."""
    import subprocess
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    cmd = [sys.executable, str(Path(__file__).resolve()), "--cpu-sample", kind, str(k), str(n), str(seconds)]
    kids = [subprocess.Popen(cmd, stdout=subprocess.PIPE, env=env) for _ in range(procs)]
    outs = []
    for c in kids:
        out, _ = c.communicate()
        if c.returncode != 0:
            raise RuntimeError(f"{kind} sample exited with {c.returncode}")
        outs.append(json.loads(out.decode().strip().splitlines()[-1]))
    rate = sum(o["count"] / o["seconds"] for o in outs)  # squarings per second, all processes
    return dict(us_per_squaring=round(1e6 * procs / rate, 1), squarings_per_s=round(rate, 2),
                squarings=sum(o["count"] for o in outs), procs=procs, threads=threads, fft_len=outs[0]["fft_len"])


def host_cores():
    """Cores this process may use: the affinity mask, capped by
    OMP_NUM_THREADS where that is set (the share a batch host grants)."""
    n = len(os.sched_getaffinity(0))
    cap = os.environ.get("OMP_NUM_THREADS")
    return min(n, int(cap)) if cap and cap.isdigit() and int(cap) > 0 else n


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(k, n, seconds=6.0, kat=(843, 120002)):
    """The CPU legs, each a bounded sample in child processes:
      reference-omp  the reference's OpenMP path, one test on every host core
                     (the headline CPU figure)
      reference      the reference as published (no -fopenmp), one core
      port           the C restatement, one core and one process per core
      workers        tests/hour of the reference's dispatch model (one
                     single-threaded test per core) at the reference's own
                     test size n=120002, from squarings/s of every core"""
    cores = host_cores()
    legs, errors = {}, {}
    for name, kind, kk, nn, procs, threads in (
            ("reference_omp", "reference-omp", k, n, 1, cores),
            ("reference_1core", "reference", k, n, 1, 1),
            ("port_1core", "port", k, n, 1, 1),
            ("port_all_cores", "port", k, n, cores, 1),
            ("reference_workers_n120002", "reference", kat[0], kat[1], cores, 1)):
        try:
            legs[name] = _cpu_run(kind, kk, nn, seconds, procs, threads)
        except Exception as e:  # not built here, or no AVX-512 on this host
            errors[name] = str(e)
    if "reference_workers_n120002" in legs:
        wk = legs["reference_workers_n120002"]
        wk["tests_per_hour"] = round(wk["squarings_per_s"] * 3600.0 / (kat[1] + kat[0].bit_length() - 2), 1)
    if "reference_omp" in legs:
        head, kind, used = legs["reference_omp"], "reference", cores
    elif "reference_1core" in legs:
        head, kind, used = legs["reference_1core"], "reference", 1
    else:
        head, kind, used = legs["port_1core"], "port", 1
    out = dict(value=head["us_per_squaring"], unit="us/squaring", cores=used, kind=kind,
               sample=(f"{head['squarings']} squarings of a random residue mod {k}*2^{n}+1 with "
                       f"fft_square_karatsuba at L={head['fft_len']} (the reference's starting length, raised "
                       f"until its limb layout holds the residue), {used} thread(s), ~{seconds:g} s"),
               cpu_model=cpu_model(), host_cores=cores, legs=legs)
    if "reference_1core" in legs and "port_1core" in legs:
        out["port_over_reference_1core"] = round(legs["port_1core"]["us_per_squaring"]
                                                 / legs["reference_1core"]["us_per_squaring"], 2)
    if errors:
        out["errors"] = errors
    return out


def _topbit(k: int) -> int:
    return k.bit_length() - 1


def plain_per_squaring(core, k, n, iters=1536):
    """Steady-state seconds per squaring of one candidate alone on the GPU
    (the plain one-candidate-at-a-time loop: one context, graph-captured).
    Two untimed chunks (the first runs partly in normal form while a^3 is
    small, as direct launches; the second is whole and captures the chunk
    graph, a one-time cost that a full test of 10^5..10^6 squarings
    amortises); whole chunks are timed."""
    c = core.Context(k, n)
    c.set_residue(core.choose_base(k, n) ** 3)
    c.square(512)
    t0 = time.perf_counter()
    c.square(iters)
    dt = time.perf_counter() - t0
    L = c.fft_len
    c.close()
    return dt / iters, L


def batch_c4(core, workloads, steps=30000):
    """Config 4: 256 candidates, n uniform in [2^18, 2^20], advanced together
    by the pool (one launch per kernel kind per length group per squaring
    step).  Each test is run for its seed powering plus `steps` squarings;
    tests/hour extrapolates the measured squaring throughput to full tests
    (topbit(k) + n - 1 squarings each).  The timed region includes creating
    the length groups, capturing their graphs, admitting and retiring every
    candidate, and each candidate's first squarings in normal form (while
    its value is small, a few dozen device products): 3*10^4 steps keep that
    one-time cost near the share it has in a full test (2.6*10^5..10^6
    squarings).  The full-length run is batch_c4_full.  The plain loop runs one candidate at a time: its
    per-squaring time is measured per transform length."""
    w = workloads.c4()
    cands = w.candidates
    full = [_topbit(k) + n - 1 for k, n in cands]
    with core.Pool(k_max=1 << 16) as pool:
        tuned = tune_batch(core, pool, cands)
        t0 = time.perf_counter()
        for i, (k, n) in enumerate(cands):
            pool.submit(k, n, tag=i, stop_after=steps)
        res = pool.drain(steps=64)
        t1 = time.perf_counter()
    done_sq = sum(r["squarings"] for r in res)
    thr = done_sq / (t1 - t0)
    batched_tph = 3600.0 * len(cands) / (sum(full) / thr)
    per_len = {}
    for (k, n) in cands:
        L = core.plan_fft_len(n, k)
        if L not in per_len:
            per_len[L] = plain_per_squaring(core, k, n)[0]
    plain_time = sum(per_len[core.plan_fft_len(n, k)] * f for (k, n), f in zip(cands, full))
    plain_tph = 3600.0 * len(cands) / plain_time
    return dict(workload="c4_batch: 256 candidates, k odd in [3,2^16), n uniform in [2^18,2^20] (PCG64 seed 4)",
                gate_redraws=w.redraws, measured_squarings=done_sq, wall_s=round(t1 - t0, 3),
                squarings_per_s=round(thr, 1), batched_tests_per_hour=round(batched_tph, 1),
                plain_tests_per_hour=round(plain_tph, 1), speedup_vs_plain=round(batched_tph / plain_tph, 2),
                plain_us_per_squaring_by_fft_len={str(L): round(v * 1e6, 2) for L, v in sorted(per_len.items())},
                max_err=max(r["max_err"] for r in res), tuned_plans=tuned,
                note=f"scaled: each test ran its seed powering + {steps} squarings; tests/hour extrapolates to n-1")


def batch_c4_full(core, workloads, out_path=None):
    """Config 4 at its stated size: the 256 candidates (k odd in [3,2^16), n
    uniform in [2^18,2^20], PCG64 seed 4) plus the known prime 3*2^303093+1,
    every one run to its verdict by the pool.  tests/hour counts completed
    verdicts over the wall time of the whole batch (group creation, graph
    captures, admissions and retirements included).  The plain loop's figure
    is its measured steady per-squaring time per transform length times each
    test's squarings.  Every (k, n, verdict, Res64) goes to out_path."""
    w = workloads.c4()
    cands = list(w.candidates) + [(3, 303093)]
    full = [_topbit(k) + n - 1 for k, n in cands]
    res = []
    with core.Pool(k_max=1 << 16) as pool:
        tuned = tune_batch(core, pool, cands)
        t0 = time.perf_counter()
        for i, (k, n) in enumerate(cands):
            pool.submit(k, n, tag=i)
        done = [0]

        def progress(r):
            done[0] += 1
            if done[0] % 32 == 0:
                log(f"[bench] c4 full: {done[0]}/{len(cands)} verdicts, {time.perf_counter() - t0:.1f} s")
        res = pool.drain(steps=64, on_result=progress)
        t1 = time.perf_counter()
    by_tag = {r["tag"]: r for r in res}
    prime_r = by_tag[len(cands) - 1]
    per_len = {}
    for (k, n) in cands:
        L = core.plan_fft_len(n, k)
        if L not in per_len:
            per_len[L] = plain_per_squaring(core, k, n)[0]
    plain_time = sum(per_len[core.plan_fft_len(n, k)] * f for (k, n), f in zip(cands, full))
    tph = 3600.0 * len(cands) / (t1 - t0)
    plain_tph = 3600.0 * len(cands) / plain_time
    if out_path:
        Path(out_path).write_text(json.dumps([dict(k=k, n=n, verdict=by_tag[i]["verdict"], error=by_tag[i]["error"],
                                                   res64=format(by_tag[i]["res64"], "016X"),
                                                   max_err=by_tag[i]["max_err"])
                                              for i, (k, n) in enumerate(cands)], indent=0))
    return dict(workload="c4_batch (full): 256 candidates + 3*2^303093+1, every test run to its verdict",
                tests=len(cands), squarings=sum(full), wall_s=round(t1 - t0, 2),
                squarings_per_s=round(sum(full) / (t1 - t0), 1), batched_tests_per_hour=round(tph, 1),
                plain_tests_per_hour=round(plain_tph, 1), speedup_vs_plain=round(tph / plain_tph, 2),
                plain_us_per_squaring_by_fft_len={str(L): round(v * 1e6, 2) for L, v in sorted(per_len.items())},
                verdicts={"prime": sum(1 for r in res if r["verdict"] == 1),
                          "composite": sum(1 for r in res if r["verdict"] == 0),
                          "void": sum(1 for r in res if r["error"] != 0)},
                known_prime={"k": 3, "n": 303093, "verdict": prime_r["verdict"],
                             "res64": format(prime_r["res64"], "016X")},
                max_err=max(r["max_err"] for r in res), tuned_plans=tuned)


def serve_c5_full(core, workloads, seg, nseg, rate, out_path=None, seed=5):
    """Config 5 at its stated size, one segment of it: the 2000 candidates
    (k odd in [3,2^20), n log-uniform in [5e4,4e6], PCG64 seed 5), every test
    run to its verdict, arriving as one Poisson process of `rate` tests per
    second (exponential gaps, PCG64(seed)).  The whole run does not fit one
    bounded GPU session, so it is cut into `nseg` consecutive segments of the
    arrival sequence; segment `seg` replays its arrivals (times shifted to
    start at 0) into a fresh pool.  Segment 0 also receives the known prime
    843*2^120002+1 right after its first arrival.  Every test's (k, n,
    arrival, time to verdict, verdict, Res64, max round-off) goes to
    out_path; the segments are summarised together afterwards."""
    import numpy as np
    w = workloads.c5()
    cands = list(w.candidates)
    rng = np.random.Generator(np.random.PCG64(seed))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=len(cands)))
    lo, hi = seg * len(cands) // nseg, (seg + 1) * len(cands) // nseg
    todo = [(i, cands[i][0], cands[i][1], float(arrivals[i] - arrivals[lo])) for i in range(lo, hi)]
    if seg == 0:
        todo.insert(1, (len(cands), 843, 120002, todo[1][3]))
    meta = {t[0]: t for t in todo}
    out = {}
    with core.Pool(k_max=1 << 20) as pool:
        tuned = tune_batch(core, pool, [(k, n) for _, k, n, _ in todo])
        log(f"[bench] c5 full segment {seg}/{nseg}: {len(todo)} arrivals at {rate}/s, tuned {len(tuned.get('pool_groups', {}))} groups")
        t0 = time.perf_counter()
        nxt, last_log = 0, t0
        while len(out) < len(todo):
            now = time.perf_counter() - t0
            while nxt < len(todo) and todo[nxt][3] <= now:
                tag, k, n, _ = todo[nxt]
                pool.submit(k, n, tag=tag)
                nxt += 1
            st = pool.status()
            if st["active"] == 0 and st["queued"] == 0 and nxt < len(todo):
                time.sleep(max(0.0, min(0.01, todo[nxt][3] - now)))
            got = pool.run(steps=64)
            for r in got:
                out[r["tag"]] = r
            if got and out_path:
                # results so far, so that a run cut short still leaves them
                with open(out_path + ".partial", "a") as f:
                    for r in got:
                        f.write(json.dumps(dict(tag=r["tag"], k=r["k"], n=r["n"], latency_s=r["t_done"] - r["t_submit"],
                                                verdict=r["verdict"], error=r["error"],
                                                res64=format(r["res64"], "016X"), max_err=r["max_err"])) + "\n")
            if time.perf_counter() - last_log > 30:
                last_log = time.perf_counter()
                log(f"[bench] c5 full segment {seg}: {now:.0f} s, {nxt}/{len(todo)} arrived, {len(out)} verdicts, "
                    f"{st['active']} active, {st['queued']} queued")
        makespan = time.perf_counter() - t0
    rows = []
    for tag, k, n, arr in todo:
        r = out[tag]
        rows.append(dict(tag=tag, k=k, n=n, arrival_s=round(arr, 4), latency_s=round(r["t_done"] - r["t_submit"], 4),
                         verdict=r["verdict"], error=r["error"], res64=format(r["res64"], "016X"),
                         max_err=r["max_err"]))
    if out_path:
        Path(out_path).write_text(json.dumps(dict(segment=seg, segments=nseg, rate=rate, makespan_s=makespan,
                                                  tuned_plans=tuned, tests=rows)))
    # the plain one-at-a-time loop's steady time per squaring at each length
    # of the segment (untimed against the serving run, after it)
    plain = {}
    for _, k, n, _ in todo:
        L = core.plan_fft_len(n, k)
        if L not in plain:
            plain[L] = plain_per_squaring(core, k, n, iters=768)[0]
    if out_path:
        d = json.loads(Path(out_path).read_text())
        d["plain_s_per_squaring_by_fft_len"] = {str(L): v for L, v in sorted(plain.items())}
        Path(out_path).write_text(json.dumps(d))
    lat = np.array([r["latency_s"] for r in rows])
    return dict(workload=f"c5_serving (full, segment {seg} of {nseg}): arrivals {lo}..{hi - 1} of 2000"
                         + (" + 843*2^120002+1" if seg == 0 else ""),
                rate_per_s=rate, tests=len(rows), makespan_s=round(makespan, 2),
                p50_time_to_verdict_s=round(float(np.percentile(lat, 50)), 3),
                p99_time_to_verdict_s=round(float(np.percentile(lat, 99)), 3),
                verdicts={"prime": sum(1 for r in rows if r["verdict"] == 1),
                          "composite": sum(1 for r in rows if r["verdict"] == 0),
                          "void": sum(1 for r in rows if r["error"] != 0)},
                max_err=max(r["max_err"] for r in rows),
                plain_us_per_squaring_by_fft_len={str(L): round(v * 1e6, 2) for L, v in sorted(plain.items())})


def serve_c5(core, workloads, count=200, scale=100, load=2.0, seed=5):
    """Config 5 (scaled): Poisson arrivals of the first `count` config-5
    candidates (k odd in [3,2^20), n log-uniform in [5e4,4e6]), each test cut
    to (n-1)/scale squarings after its seed so the run stays bounded.
    Arrival rate = `load` x the plain sequential loop's capacity.  The pool
    admits arrivals as slots free up (continuous batching); the plain loop's
    latencies come from a FIFO single-server replay of the same arrivals with
    its measured per-squaring times."""
    import numpy as np
    w = workloads.c5(count)
    cands = w.candidates
    sq = [_topbit(k) + max(1, (n - 1) // scale) for k, n in cands]
    lat = {}
    with core.Pool(k_max=1 << 20) as pool:
        tuned = tune_batch(core, pool, cands)
        per_len = {}
        for (k, n) in cands:
            L = core.plan_fft_len(n, k)
            if L not in per_len:
                per_len[L] = plain_per_squaring(core, k, n, iters=768)[0]
        svc = [per_len[core.plan_fft_len(n, k)] * s for (k, n), s in zip(cands, sq)]
        lam = load / (sum(svc) / len(svc))
        rng = np.random.Generator(np.random.PCG64(seed))
        arrivals = np.cumsum(rng.exponential(1.0 / lam, size=len(cands)))
        t0 = time.perf_counter()
        nxt = 0
        while len(lat) < len(cands):
            now = time.perf_counter() - t0
            while nxt < len(cands) and arrivals[nxt] <= now:
                k, n = cands[nxt]
                pool.submit(k, n, tag=nxt, stop_after=max(1, (n - 1) // scale))
                nxt += 1
            st = pool.status()
            if st["active"] == 0 and st["queued"] == 0 and nxt < len(cands):
                time.sleep(max(0.0, min(0.01, arrivals[nxt] - now)))
            for r in pool.run(steps=64):
                lat[r["tag"]] = r["t_done"] - r["t_submit"]
        makespan = time.perf_counter() - t0
    # plain sequential loop on the same arrivals (FIFO, one test at a time)
    free_at, plat = 0.0, []
    for a, s in zip(arrivals, svc):
        start = max(a, free_at)
        free_at = start + s
        plat.append(free_at - a)
    L_b = np.array([lat[i] for i in range(len(cands))])
    L_p = np.array(plat)
    return dict(workload=f"c5_serving (scaled): first {count} config-5 candidates, each cut to (n-1)/{scale} "
                         "squarings after the seed, Poisson arrivals",
                arrival_rate_per_s=round(lam, 3), offered_load_vs_plain_capacity=load,
                p50_time_to_verdict_s=round(float(np.percentile(L_b, 50)), 4),
                p99_time_to_verdict_s=round(float(np.percentile(L_b, 99)), 4),
                plain_p50_s=round(float(np.percentile(L_p, 50)), 4),
                plain_p99_s=round(float(np.percentile(L_p, 99)), 4),
                tests_per_hour=round(3600.0 * len(cands) / makespan, 1),
                plain_tests_per_hour=round(3600.0 * len(cands) / max(float(free_at), 1e-9), 1),
                makespan_s=round(makespan, 3), tuned_plans=tuned)


def plain_keys(core, workloads, out_path):
    """The plain one-at-a-time loop's steady time per squaring for every
    (transform length, carry run) key that the config-4 and config-5
    candidates form (a single context's speed depends on both: runs of 16
    digits take the split outer pass).  Each candidate's key is read from a
    context; each key's first candidate is tuned and timed alone, best of two
    1536-squaring runs.  Writes {"keys": [key of each config-5 candidate, by
    its index, 843*2^120002+1 last], "keys_c4": [the same for config 4,
    3*2^303093+1 last], "plain": {key: {k, n, s, plan}}}; keys read "L/run"."""
    c5 = list(workloads.c5().candidates) + [(843, 120002)]
    c4 = list(workloads.c4().candidates) + [(3, 303093)]
    keys = []
    for k, n in c5 + c4:
        c = core.Context(k, n)
        keys.append(f"{c.fft_len}/{c.info.run}")
        c.close()
    first = {}
    for i, key in enumerate(keys):
        first.setdefault(key, i)
    plain = {}
    for key, i in sorted(first.items(), key=lambda x: tuple(map(int, x[0].split("/")))):
        k, n = (c5 + c4)[i]
        r = core.autotune_plan(k, n) if TUNE else None
        s = min(plain_per_squaring(core, k, n, iters=1536)[0] for _ in range(2))
        plain[key] = dict(k=k, n=n, s=s, plan=_plan_str(r["best"]) if r else "rule")
        log(f"[bench] plain {key}: {s * 1e6:.2f} us per squaring ({plain[key]['plan']})")
    d = dict(keys=keys[:len(c5)], keys_c4=keys[len(c5):], plain=plain)
    Path(out_path).write_text(json.dumps(d))
    return {key: round(v["s"] * 1e6, 2) for key, v in plain.items()}


def c5_combine(plain_path, gmp_path, seg_paths):
    """Text summary of full-length config-5 segments (`--c5-full` outputs):
    time to verdict, verdicts, throughput, and the plain one-at-a-time loop
    replayed on the same arrivals (FIFO, one test at a time, each taking its
    squarings times the steady per-squaring time of its key from
    `--plain-keys`), plus the GMP check lines of tests/gmp_check_results.py
    when gmp_path names a file ("-": none)."""
    import numpy as np
    data = sorted((json.loads(Path(p).read_text()) for p in seg_paths), key=lambda d: d["segment"])
    keyed = json.loads(Path(plain_path).read_text())
    nseg, rate = data[0]["segments"], data[0]["rate"]
    rows = [r for d in data for r in d["tests"]]
    lat = np.array([r["latency_s"] for r in rows])

    def squarings(r):
        return r["n"] - 1 + r["k"].bit_length() - 1

    def plain_s(r):
        return keyed["plain"][keyed["keys"][r["tag"]]]["s"]

    plat = []
    for d in data:
        free = 0.0
        for r in d["tests"]:
            free = max(r["arrival_s"], free) + plain_s(r) * squarings(r)
            plat.append(free - r["arrival_s"])
    plat = np.array(plat)
    mk = sum(d["makespan_s"] for d in data)
    busy = sum(plain_s(r) * squarings(r) for r in rows)
    tph, ptph = 3600 * len(rows) / mk, 3600 * len(rows) / busy
    ver = dict(prime=sum(r["verdict"] == 1 for r in rows), composite=sum(r["verdict"] == 0 for r in rows),
               void=sum(r["error"] != 0 for r in rows))
    kp = [r for r in rows if (r["k"], r["n"]) == (843, 120002)]
    by_key = ", ".join(f"L={key.replace('/', ' run')}: {v['s'] * 1e6:.2f}" for key, v in
                       sorted(keyed["plain"].items(), key=lambda x: tuple(map(int, x[0].split("/")))))
    out = [
        "# Config 5 at its stated size: 2000 Poisson arrivals, every test run to its verdict (serving)",
        "",
        "commands (from the repository root; the first two on the GPU host, one bounded session per segment):",
        "",
        f"    python3 bench.py --no-cpu --c5-full S {nseg} {rate} gpurun_out/c5seg/segS_of_{nseg}.json     "
        f"(S = {', '.join(str(d['segment']) for d in data)})",
        "    python3 bench.py --no-cpu --plain-keys gpurun_out/plain_keys.json",
        "    python3 tests/gmp_check_results.py --nmax N --out gpurun_out/c5_gmp.jsonl gpurun_out/c5seg/seg*_of_8.json"
        "     (in the container; N below)",
        "    python3 bench.py --c5-combine gpurun_out/plain_keys.json gpurun_out/c5_gmp.jsonl "
        "gpurun_out/c5seg/seg*_of_8.json",
        "",
        "Workload: the 2000 config-5 candidates (k odd in [3,2^20), n log-uniform in [5e4,4e6], PCG64 seed 5, gate",
        f"survivors) arriving as one Poisson process at {rate} tests/s (exponential gaps, PCG64 seed 5), plus the",
        f"known prime 843*2^120002+1 right after the first arrival.  The arrival sequence is cut into {nseg} "
        "consecutive",
        "segments of 250 arrivals; each segment replays its arrivals (times shifted to start at 0) into a fresh "
        "pool in",
        f"its own GPU session.  Segments run: {[d['segment'] for d in data]} of 0..{nseg - 1} ({len(rows)} of the "
        "2001 tests).",
        "Time to verdict = verdict time - submission time of each test.  The plain comparison replays the same",
        "arrivals through one test at a time (FIFO, one context), each test taking its squarings times the",
        "steady per-squaring time of a single context of its (length, carry run) key (--plain-keys).",
        "",
        f"tests                      {len(rows)}",
        f"verdicts                   {ver}",
        f"known prime                843*2^120002+1: "
        + (f"verdict {kp[0]['verdict']}, Res64 {kp[0]['res64']}" if kp else "not in these segments"),
        f"max round-off              {max(r['max_err'] for r in rows)}",
        f"pool p50 / p99 / max       {np.percentile(lat, 50):.2f} / {np.percentile(lat, 99):.2f} / "
        f"{lat.max():.2f} s time to verdict",
        f"plain p50 / p99 / max      {np.percentile(plat, 50):.1f} / {np.percentile(plat, 99):.1f} / "
        f"{plat.max():.1f} s",
        f"makespan (sum of segments) {mk:.1f} s  -> {tph:.1f} tests/h completed (the pool, not the arrivals, "
        f"bounds it:",
        f"                           it serves this mix at about {len(rows) / mk:.2f} tests/s against "
        f"{rate} arriving)",
        f"plain loop busy time       {busy:.1f} s  -> its capacity {ptph:.1f} tests/h",
        f"pool / plain throughput    {tph / ptph:.2f}x",
        f"plain us/squaring by key   {by_key}",
        "",
    ]
    if gmp_path != "-" and Path(gmp_path).exists():
        got = {(r["k"], r["n"]): r for r in rows}
        gmp = [g for g in map(json.loads, Path(gmp_path).read_text().splitlines()) if (g["k"], g["n"]) in got]
        if gmp:
            ok = sum(g["match"] and g["device"] == got[(g["k"], g["n"])]["res64"] for g in gmp)
            top = max(g["n"] for g in gmp)
            checked = {(g["k"], g["n"]) for g in gmp}
            below = [r for r in rows if r["n"] <= top]
            out += ["GMP check (tests/gmp_check_results.py: mpz_powm, the reference's base rule):",
                    f"    {len(gmp)} tests, n {min(g['n'] for g in gmp)}..{top}, k {min(g['k'] for g in gmp)}.."
                    f"{max(g['k'] for g in gmp)}: {sum((r['k'], r['n']) in checked for r in below)} of the "
                    f"{len(below)} tests with n <= {top}",
                    f"    Res64 and verdict equal GMP's: {ok} of {len(gmp)}", ""]
    out.append("per segment:")
    for d in data:
        sl = np.array([r["latency_s"] for r in d["tests"]])
        out.append(f"    segment {d['segment']}: {len(d['tests'])} tests, makespan {d['makespan_s']:.1f} s, "
                   f"p50 {np.percentile(sl, 50):.2f} s, p99 {np.percentile(sl, 99):.2f} s")
    return "\n".join(out) + "\n"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-c3", action="store_true", help="skip the n=2^24 line")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched (config 4) leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving (config 5) leg")
    ap.add_argument("--c4-full", metavar="OUT.json", default=None,
                    help="only run config 4 at its stated size (every test to its verdict); per-test results to OUT")
    ap.add_argument("--c5-full", nargs=4, metavar=("SEG", "NSEG", "RATE", "OUT.json"), default=None,
                    help="only run segment SEG of NSEG of config 5 at its stated size (Poisson arrivals at RATE "
                         "tests/s, every test to its verdict); per-test results to OUT")
    ap.add_argument("--plain-keys", metavar="OUT.json", default=None,
                    help="only time the plain loop per (transform length, carry run) key of the config-4 and "
                         "config-5 candidates; to OUT")
    ap.add_argument("--c5-combine", nargs="+", metavar="FILE", default=None,
                    help="PLAIN.json GMP.jsonl|- SEG.json...: print the summary of full-length config-5 segments "
                         "(no GPU needed)")
    ap.add_argument("--cpu-only", action="store_true", help="only the CPU legs (no GPU needed)")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the launch-plan autotuner (plans from --plans, else the built-in rule); for "
                         "profiler runs, whose per-dispatch tracing makes the tuner's timings meaningless")
    ap.add_argument("--plans", metavar="FILE", default=None, help="load a tuned-plan table first")
    ap.add_argument("--save-plans", metavar="FILE", default=None, help="write the tuned-plan table at the end")
    ap.add_argument("--cpu-sample", nargs=4, metavar=("KIND", "K", "N", "SECONDS"), help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_sample:
        kind, k, n, sec = args.cpu_sample
        print(json.dumps(cpu_sample(kind, int(k), int(n), float(sec))), flush=True)
        return
    if args.c5_combine:
        if len(args.c5_combine) < 3:
            ap.error("--c5-combine needs PLAIN.json, GMP.jsonl (or -) and at least one segment file")
        sys.stdout.write(c5_combine(args.c5_combine[0], args.c5_combine[1], args.c5_combine[2:]))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: every configuration fits one MI355X and candidates are "
              "independent (they shard by candidate with no communication); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    from prothlogic_amd import workloads
    (k2, n2), = workloads.c2().candidates
    cpu = None
    if args.cpu_only or not args.no_cpu:
        # before the GPU is initialised: the legs run in child processes
        try:
            cpu = cpu_baseline(k2, n2)
            log(f"[bench] cpu baseline: {cpu}")
        except Exception as e:
            log(f"[bench] cpu baseline failed: {e}")
        if args.cpu_only:
            print(json.dumps(cpu), flush=True)
            return

    from prothlogic_amd import core
    if core.device_count() < 1:
        log("no HIP device visible")
        sys.exit(1)
    global TUNE
    TUNE = not args.no_tune
    if args.plans:
        core.tune_load(args.plans)
    if args.c4_full:
        print(json.dumps(batch_c4_full(core, workloads, args.c4_full)), flush=True)
        return
    if args.plain_keys:
        print(json.dumps(plain_keys(core, workloads, args.plain_keys)), flush=True)
        return
    if args.c5_full:
        seg, nseg, rate, outp = args.c5_full
        print(json.dumps(serve_c5_full(core, workloads, int(seg), int(nseg), float(rate), outp)), flush=True)
        return

    log(f"[bench] c2: k={k2} n={n2} steps={args.steps} warmup={args.warmup}")
    tune2 = tune_single(core, k2, n2)
    log(f"[bench] c2 plan: {tune2}")
    # Res64 checked after 10^4 squarings from the seed (the count the GMP
    # fixture holds) in an untimed second run, whatever --steps is
    v2 = 10000
    r2 = time_squarings(core, k2, n2, args.steps, args.warmup, verify_at=v2)
    us2 = r2["seconds"] / args.steps * 1e6
    g2 = golden_res64(k2, n2, v2)
    dom, roof, dom_bytes, sq_bytes = roofline(r2)
    roof["traffic"] = measured_traffic(dom, "c2_n2p21")
    log(f"[bench] c2: {us2:.2f} us/squaring, max_err={r2['max_err']:.4g}, kernels(ms)={r2['kernel_ms']}")

    extra = {}
    if not args.no_c3:
        (k3, n3), = workloads.c3().candidates
        # four whole 256-squaring chunks timed after one untimed chunk;
        # Res64 checked after 10^3 squarings, the count the GMP fixture holds
        # 512 warm-up squarings: the first chunk is partly in normal form (the
        # seed's exact value is small) and runs as direct launches; the second
        # is a whole one, which captures the chunk graph before the clock starts
        s3, v3, w3 = 1024, 1000, 512
        log(f"[bench] c3: k={k3} n={n3} steps={s3} warmup={w3}")
        tune3 = tune_single(core, k3, n3)
        log(f"[bench] c3 plan: {tune3}")
        r3 = time_squarings(core, k3, n3, s3, w3, verify_at=v3)
        us3 = r3["seconds"] / s3 * 1e6
        g3 = golden_res64(k3, n3, v3)
        dom3, roof3, dom3_bytes, sq3_bytes = roofline(r3)
        roof3["traffic"] = measured_traffic(dom3, "c3_n2p24")
        roofline_us = sq3_bytes / (HBM_PEAK_GBS * 1e9) * 1e6
        extra["n2p24"] = dict(k=k3, n=n3, fft_len=r3["fft_len"], us_per_squaring=round(us3, 3), squarings=s3,
                              res64_after=v3, max_err=r3["max_err"], res64=format(r3["res64"], "016X"),
                              res64_matches_gmp=(None if g3 is None else r3["res64"] == g3),
                              kernel_ms=r3["kernel_ms"], dominant_kernel=dom3, roofline=roof3,
                              kernels_per_squaring=r3["kernels_per_squaring"], bytes_per_squaring=sq3_bytes,
                              hbm_roofline_us=round(roofline_us, 3), x_of_roofline=round(us3 / roofline_us, 2),
                              fp64=fp64_rate(r3["fft_len"], us3), plan_tuning=tune3)
        log(f"[bench] c3: {us3:.2f} us/squaring, max_err={r3['max_err']:.4g}")

    if not args.no_batch:
        log("[bench] c4 batch")
        extra["batch_c4"] = batch_c4(core, workloads)
        log(f"[bench] c4: {extra['batch_c4']}")
    if not args.no_serve:
        log("[bench] c5 serving")
        extra["serve_c5"] = serve_c5(core, workloads)
        log(f"[bench] c5: {extra['serve_c5']}")

    line = {
        "metric": METRIC,
        "value": round(us2, 3),
        "unit": "us/squaring",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(us2 / 1000.0, 6),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded Proth candidate (PCG64 seed 2), residue a^k mod N built on the device",
        "config": {"workload": "c2_n2p21 single test: n=2^21, k odd in [3,2^20), 1 squaring per step",
                   "k": k2, "n": n2, "base": r2["a"], "fft_len": r2["fft_len"], "rows": r2["rows"],
                   "cols": r2["cols"], "bits_per_digit": r2["bits_per_digit"],
                   "kernels_per_squaring": r2["kernels_per_squaring"], "row_radix": r2["row_radix"],
                   "col_radix": r2["col_radix"], "carry_run": r2["run"]},
        "plan_tuning": tune2,
        "res64_after": v2,
        "res64": format(r2["res64"], "016X"),
        "res64_matches_gmp": None if g2 is None else r2["res64"] == g2,
        "max_err": r2["max_err"],
        "kernel_ms": r2["kernel_ms"],
        "dominant_kernel": dom,
        "dominant_kernel_bytes": dom_bytes,
        "bytes_per_squaring": sq_bytes,
        "hbm_roofline_us": round(sq_bytes / (HBM_PEAK_GBS * 1e9) * 1e6, 3),
        "roofline": roof,
        "fp64": fp64_rate(r2["fft_len"], us2),
        "cpu_baseline": cpu,
        **extra,
    }
    print(json.dumps(line), flush=True)
    if args.save_plans:
        core.tune_save(args.save_plans)


if __name__ == "__main__":
    main()
