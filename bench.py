"""Throughput of the 2-D TM forward + adjoint gradient on one MI355X.

Workload (one step = one full gradient): 1024x1024 model (1065x1065 padded,
CPML 20, stencil 2), dx 0.01 m, dt at CFL 0.9, nt 4000, 64 shots along the
surface row at 16-cell spacing (200 MHz Ricker), 256 receivers per shot,
seeded layered eps_r 1-25 with 12 Gaussian anomalies, sigma U[0, 0.01].
Each step computes the L2 misfit (FP64) against data from a perturbed model
(one extra anomaly, computed before timing) and its eps/sigma gradient
through the public ``tide.MaxwellTM`` operator, shot chunk by shot chunk so
that fp32 snapshots fit in HBM.  Useful work counted: 2 sweeps x 1065^2 x 64
x 4000 = 5.807e11 cell-updates per step.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--nt NT]
                    [--compression bf16] [--no-cpu-baseline]

Prints one JSON line (rank 0).
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for _p in (ROOT / "tide-gpr_amd", ROOT):
    if str(_p) not in sys.path:
        sys.path.insert(0, str(_p))

HBM_PEAK_GBS = 8000.0
# Counted bytes per cell-update of the 2-D kernels come from the committed
# rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-calibrated); its fenced
# ``json traffic`` block maps "kernel:steps per launch:snapshot format" to
# bytes per cell-update.  Reported as roofline.traffic when the bench's
# dominant kernel has an entry.
TRAFFIC_SUMMARY = ROOT / "profiles" / "tm2d_1024x1024_64shot_gradient_hbm_counters.md"


def counted_traffic() -> dict:
    """{"kernel:K:fmt": bytes per cell-update} from the counter summary."""
    import json
    import re

    try:
        text = TRAFFIC_SUMMARY.read_text()
    except OSError:
        return {}
    m = re.search(r"```json traffic\n(.*?)```", text, re.S)
    return json.loads(m.group(1)) if m else {}


def log(msg: str) -> None:
    """Progress on stderr; stdout carries only the JSON line."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)
METRIC = "fwd+adj Gcell-updates/s & %HBM roofline (2D TM); p99 served-solve latency"


def parse() -> argparse.Namespace:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nt", type=int, default=4000)
    ap.add_argument("--shots", type=int, default=64)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--compression", default="none", choices=["none", "bf16"])
    ap.add_argument("--chunk", type=int, default=0, help="shots per chunk (0: fit HBM)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the library's default launch shapes")
    ap.add_argument("--cpu-sample-nt", type=int, default=3000)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one 1 GiB device copy (profiler byte-counter calibration)")
    ap.add_argument("--serve-requests", type=int, default=200,
                    help="served-solve trace length at finite arrival rates (0: skip the serving "
                         "measurement)")
    ap.add_argument("--serve-capacity-requests", type=int, default=24,
                    help="served-solve trace length at rate inf (every request at once: capacity)")
    ap.add_argument("--serve-skip-plain", action="store_true",
                    help="serving: time only the served path (policy sweeps)")
    ap.add_argument("--serve-boost", type=int, default=4,
                    help="serving engine: most steps of a small job per tick (x largest)")
    ap.add_argument("--serve-head-weight", type=int, default=1,
                    help="serving engine: step multiplier of the oldest request's jobs")
    ap.add_argument("--serve-max-bypass", type=int, default=8,
                    help="serving: admissions that may overtake a waiting chunk")
    ap.add_argument("--serve-rates", default="4,inf",
                    help="Poisson arrival rates (per second) of the serving sweep; inf = all at once")
    ap.add_argument("--serve-seed", type=int, default=2024)
    ap.add_argument("--serve-tick", type=int, default=64, help="steps per engine tick")
    ap.add_argument("--batched-models", type=int, default=32,
                    help="models of the batched uneven-model measurement (0: skip)")
    ap.add_argument("--batched-seed", type=int, default=3030)
    ap.add_argument("--no-config1", action="store_true",
                    help="skip the small launch-latency-bound forward measurement")
    ap.add_argument("--em3d-nt", type=int, default=2000,
                    help="time steps of the 3-D measurement (0: skip)")
    ap.add_argument("--em3d-shots", type=int, default=8)
    ap.add_argument("--em3d-only", action="store_true",
                    help="run only the 3-D measurement (for profiling it) and print its JSON")
    ap.add_argument("--accuracy-shots", type=int, default=4,
                    help="shots of the fp32-vs-fp64 and BF16 error sample (0: skip)")
    return ap.parse_args()


def _has_avx2() -> bool:
    try:
        return " avx2" in Path("/proc/cpuinfo").read_text()
    except OSError:
        return False


def cpu_baseline(args, setup_kwargs) -> dict:
    """CPU baselines on the host cores, on bounded samples of the headline
    workload (the first shots, full grid, shortened nt):

    * value: the reference's own CPU forward_with_storage + backward built
      with its Release flags (-Ofast -mavx2, src/tide/csrc/CMakeLists.txt)
      from its sources by oracle/Makefile, all host cores (OpenMP);
    * the same at -O2; the reference CPU forward on the 96x96 x 300-step
      case; and this project's CPU restatement with the exact adjoint (the
      reference's CPU backward re-runs the forward kernels on the adjoint
      fields, which is not the exact transpose with CPML).
    """
    import numpy as np

    from oracle import oracle as O
    from tide.maxwell.tm2d import prepare_tm2d

    cores = len(os.sched_getaffinity(0))

    def sample(n_shots, nt):
        kw = dict(setup_kwargs)
        kw["source_amplitude"] = kw["source_amplitude"][:n_shots, :, :nt].cpu()
        kw["source_location"] = kw["source_location"][:n_shots].cpu()
        kw["receiver_location"] = kw["receiver_location"][:n_shots].cpu()
        for k in ("epsilon", "sigma", "mu"):
            kw[k] = kw[k].detach().cpu()
        case = O.case_from_setup(prepare_tm2d(**kw, nt=nt))
        grad_r = np.zeros((nt, n_shots, case.rec.shape[1]), np.float32)
        grad_r[nt // 2 :] = 1.0
        return case, grad_r

    def rate(updates, secs):
        return round(updates / secs / 1e9, 4)

    out: dict = {"unit": "Gcell-updates/s", "cores": cores}
    n_shots, nt = 2, args.cpu_sample_nt
    case, grad_r = sample(n_shots, nt)
    updates = 2.0 * n_shots * case.ny * case.nx * nt
    desc = f"{n_shots} shots x {case.ny}x{case.nx} x nt {nt} ({updates:.3e} cell-updates)"
    if O.ref_available(fast=True) and _has_avx2():
        t0 = time.perf_counter()
        O.ref_tm2d_forward_with_storage_and_backward(case, grad_r, n_threads=cores, fast=True)
        secs = time.perf_counter() - t0
        out.update(value=rate(updates, secs), kind="reference",
                   sample=f"reference CPU forward_with_storage + backward, Release flags "
                          f"(-Ofast -mavx2, OpenMP): {desc}, {secs:.1f} s")
    if O.ref_available():
        t0 = time.perf_counter()
        O.ref_tm2d_forward_with_storage_and_backward(case, grad_r, n_threads=cores)
        secs = time.perf_counter() - t0
        out["reference_O2"] = {"value": rate(updates, secs), "seconds": round(secs, 2),
                               "sample": f"same, -O2: {desc}"}
        if "value" not in out:
            out.update(value=out["reference_O2"]["value"], kind="reference",
                       sample="reference CPU forward_with_storage + backward (-O2): " + desc)
        # the 96x96 (117x117 padded) x 300-step one-shot forward
        import torch

        import tide

        n1, nt1, dt1 = 96, 300, 4e-11
        eps1 = torch.full((n1, n1), 4.0)
        st1 = prepare_tm2d(eps1, torch.zeros_like(eps1), torch.ones_like(eps1), 0.02, dt1,
                           tide.ricker(8e8, nt1, dt1, dtype=torch.float32).reshape(1, 1, -1),
                           torch.tensor([[[20, 48]]]), torch.tensor([[[20, 60]]]), pml_width=10)
        case1 = O.case_from_setup(st1)
        upd1 = float(case1.ny * case1.nx * nt1)
        for fast in ((True, False) if _has_avx2() and O.ref_available(fast=True) else (False,)):
            O.ref_tm2d_forward(case1, n_threads=cores, fast=fast)
            reps = 20
            t0 = time.perf_counter()
            for _ in range(reps):
                O.ref_tm2d_forward(case1, n_threads=cores, fast=fast)
            secs = (time.perf_counter() - t0) / reps
            out["config1_forward_" + ("release" if fast else "O2")] = {
                "value": rate(upd1, secs), "ms": round(secs * 1e3, 3),
                "sample": "reference CPU forward, 96x96 (117x117 padded) x 300 steps, one shot"}
    # this project's CPU restatement: forward with storage + exact adjoint
    r_shots, r_nt = 1, 300
    case_r, grad_r_r = sample(r_shots, r_nt)
    t0 = time.perf_counter()
    _, _, es, cs = O.tm2d_forward(case_r, store="full")
    O.tm2d_backward(case_r, grad_r_r, es, cs)
    secs = time.perf_counter() - t0
    upd_r = 2.0 * r_shots * case_r.ny * case_r.nx * r_nt
    port = {"value": rate(upd_r, secs), "seconds": round(secs, 2),
            "sample": f"CPU restatement (oracle), forward with storage + exact adjoint, -O2 "
                      f"OpenMP: {r_shots} shot x {case_r.ny}x{case_r.nx} x nt {r_nt}"}
    out["port_exact_adjoint"] = port
    if "value" not in out:
        out.update(value=port["value"], kind="port", sample=port["sample"])
    return out


def _request_updates(kw) -> float:
    """Useful cell-updates of one served request (forward: one sweep;
    gradient: forward + adjoint)."""
    from tide.serve import padded_shape

    p = kw["problem"]
    ny_p, nx_p = padded_shape(*p["epsilon"].shape, p.get("pml_width", 20))
    shots, _, nt = p["source_amplitude"].shape
    return (2.0 if kw["kind"] == "gradient" else 1.0) * ny_p * nx_p * shots * nt


def _plain_solve(kw, dev, budget: int) -> None:
    """One request through the public operator, one shot chunk at a time
    (chunks sized by the same HBM rule as the server)."""
    import torch

    from tide.maxwell.tm2d import maxwelltm
    from tide.serve import job_bytes, padded_shape, plan_chunks

    p = {k: (v.to(dev) if hasattr(v, "to") and k not in ("source_location", "receiver_location") else v)
         for k, v in kw["problem"].items()}
    amp = p["source_amplitude"]
    shots, n_src, nt = amp.shape
    n_rec = p["receiver_location"].shape[1]
    ny_p, nx_p = padded_shape(*p["epsilon"].shape, p.get("pml_width", 20))
    grad = kw["kind"] == "gradient"
    bf16 = kw.get("storage_compression") in (True, "bf16")
    chunks = plan_chunks(shots, lambda c: job_bytes(ny_p, nx_p, c, nt, n_src, n_rec, gradient=grad,
                                                    bf16=bf16), budget)
    if grad:
        obs = kw["observed"].to(dev)
        e = p["epsilon"].clone().requires_grad_(True)
        s = p["sigma"].clone().requires_grad_(True)
    for s0, s1 in chunks:
        sub = dict(p, source_amplitude=amp[s0:s1], source_location=p["source_location"][s0:s1],
                   receiver_location=p["receiver_location"][s0:s1])
        if not grad:
            with torch.no_grad():
                maxwelltm(**sub)
            continue
        r = maxwelltm(**dict(sub, epsilon=e, sigma=s),
                      storage_compression="bf16" if bf16 else False)[-1]
        loss = 0.5 * ((r.double() - obs[:, s0:s1].double()) ** 2).sum()
        loss.backward()


def serve_bench(args, dev, rate: float) -> dict:
    """Served solves: a seeded Poisson trace of forward (70%) and gradient
    (30%) requests replayed in real time through the continuous-batching
    server, then through the plain per-request path (one request at a time,
    first come first served, on the same arrival times).  The same requests
    at every rate (only the arrival times scale); rate inf puts every
    arrival at 0, so solves/s is each path's capacity."""
    import torch

    from tide.serve import TM2DServer, latency_stats, served_trace

    n_req = args.serve_requests if math.isfinite(rate) else args.serve_capacity_requests
    trace = served_trace(n_req, rate, seed=args.serve_seed)
    work = sum(_request_updates(kw) for _, kw in trace)
    log(f"serving trace: {len(trace)} requests, {work:.3e} cell-updates")
    torch.cuda.empty_cache()
    server = TM2DServer(dev, steps_per_tick=args.serve_tick, tick_boost=args.serve_boost,
                        head_weight=args.serve_head_weight, max_bypass=args.serve_max_bypass)
    budget = server.budget
    # first contact with the kernels and allocator outside the measurement
    server.serve([(0.0, dict(trace[0][1]))])
    tickets = server.serve(trace)
    torch.cuda.synchronize()
    served = latency_stats(tickets)
    served["gcell_per_s"] = work / served["span_s"] / 1e9
    log(f"served: {served} in {server.ticks} ticks")
    server.close()
    del server
    torch.cuda.empty_cache()
    r3 = lambda d: {k: (round(v, 4) if isinstance(v, float) else v) for k, v in d.items()}  # noqa: E731
    if args.serve_skip_plain:
        return {"rate_per_s": rate if math.isfinite(rate) else "inf", "boost": args.serve_boost,
                "head_weight": args.serve_head_weight, "max_bypass": args.serve_max_bypass,
                "requests": len(trace), "served": r3(served)}
    # The server allocated its pool before the trace; the plain path gets the
    # same: one untimed pass over the requests leaves its buffers in torch's
    # caching allocator, so its timed solves do not pay first-time device
    # allocation either.
    for _, kw in trace:
        _plain_solve(kw, dev, budget)
    torch.cuda.synchronize()

    lat, t_free, first = [], 0.0, trace[0][0]
    for arrival, kw in trace:
        t0 = time.perf_counter()
        _plain_solve(kw, dev, budget)
        torch.cuda.synchronize()
        svc = time.perf_counter() - t0
        start = max(arrival, t_free)
        t_free = start + svc
        lat.append(t_free - arrival)
    lat.sort()

    def pct(q: float) -> float:  # the interpolation of tide.serve.latency_stats
        pos = q * (len(lat) - 1)
        lo = int(math.floor(pos))
        hi = min(lo + 1, len(lat) - 1)
        return lat[lo] + (lat[hi] - lat[lo]) * (pos - lo)

    plain = {"solves": len(lat), "solves_per_s": len(lat) / (t_free - first), "p50_s": pct(0.5),
             "p99_s": pct(0.99), "mean_s": sum(lat) / len(lat), "span_s": t_free - first,
             "gcell_per_s": work / (t_free - first) / 1e9}
    log(f"plain: {plain}")
    return {
        "rate_per_s": rate if math.isfinite(rate) else "inf",
        "workload": f"{len(trace)} Poisson arrivals at {rate}/s (seed {args.serve_seed}), "
                    "70% forward / 30% forward+gradient, sides log-uniform 128-1536, nt 500-6000, "
                    "1-32 shots, 64 receivers, fp32 snapshots",
        "cell_updates": work,
        "hbm_budget_bytes": budget,
        "engine": {"tick_steps": args.serve_tick, "boost": args.serve_boost,
                   "head_weight": args.serve_head_weight, "max_bypass": args.serve_max_bypass},
        "served": r3(served),
        "plain_one_at_a_time": r3(plain),
        "solves_per_s_ratio": round(served["solves_per_s"] / plain["solves_per_s"], 3),
        "p99_ratio_plain_over_served": round(plain["p99_s"] / served["p99_s"], 3),
    }


def batched_bench(args, dev) -> dict:
    """Batched uneven models (the 256-model distribution, fewer models):
    sides log-uniform 128-1024 per axis, nt 500-4000, 1-16 shots, seeded
    layered + anomaly models, fp32, forward + adjoint gradient each.  The
    models are packed in arrival order into batched plans that fit the HBM
    budget (one launch per step over every model of a plan, graph-replayed
    time loops).  Building a plan (preparation, point tables, graph capture)
    is timed apart from its gradients: an inversion builds it once and then
    runs one gradient per iteration (the second run).  The plain path solves
    the same models one at a time through the public operator."""
    import torch

    from tide.batch import TM2DBatch
    from tide.serve import job_bytes, padded_shape, served_trace

    trace = served_trace(args.batched_models, math.inf, seed=args.batched_seed,
                         side_range=(128, 1024), nt_range=(500, 4000), shot_range=(1, 16),
                         gradient_fraction=1.0, n_receivers=64, device=dev)
    reqs = [kw for _, kw in trace]
    work = sum(_request_updates(kw) for kw in reqs)
    torch.cuda.empty_cache()
    free, _ = torch.cuda.mem_get_info(dev)
    budget = int(0.55 * free)

    def nbytes(kw):
        p = kw["problem"]
        ny_p, nx_p = padded_shape(*p["epsilon"].shape, p.get("pml_width", 20))
        shots, n_src, nt = p["source_amplitude"].shape
        return job_bytes(ny_p, nx_p, shots, nt, n_src, p["receiver_location"].shape[1], gradient=True)

    groups, cur, used = [], [], 0
    for kw in reqs:
        b = nbytes(kw)
        if cur and used + b > budget:
            groups.append(cur)
            cur, used = [], 0
        cur.append(kw)
        used += b
    groups.append(cur)
    log(f"batched: {len(reqs)} models, {work:.3e} cell-updates, plans of {[len(g) for g in groups]}")

    torch.cuda.synchronize()
    t_build = t_first = t_again = 0.0
    for g in groups:
        t0 = time.perf_counter()
        plan = TM2DBatch([kw["problem"] for kw in g])
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        plan.gradient([kw["observed"] for kw in g])
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        plan.gradient([kw["observed"] for kw in g])  # the next iteration of an inversion
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        t_build, t_first, t_again = t_build + t1 - t0, t_first + t2 - t1, t_again + t3 - t2
        plan.close()
        del plan
    t_batched = t_build + t_first
    log(f"batched plans: build {t_build:.2f} s, first gradient {t_first:.2f} s, next {t_again:.2f} s")

    # The plans' build time includes allocating their buffers; the plain
    # path is timed in steady state (an untimed pass first leaves its
    # buffers in torch's caching allocator), as an inversion's later
    # iterations run.
    torch.cuda.empty_cache()
    for kw in reqs:
        _plain_solve(kw, dev, budget)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for kw in reqs:
        _plain_solve(kw, dev, budget)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    log(f"plain one at a time: {t_plain:.2f} s")
    n = len(reqs)
    return {
        "workload": f"{n} models (seed {args.batched_seed}): sides log-uniform 128-1024, nt 500-4000, "
                    "1-16 shots, 64 receivers, fp32, forward + adjoint gradient each",
        "cell_updates": work,
        "plans": [len(g) for g in groups],
        "batched": {"build_s": round(t_build, 3), "first_gradient_s": round(t_first, 3),
                    "next_gradient_s": round(t_again, 3),
                    "solves_per_s": round(n / t_again, 3),
                    "solves_per_s_incl_build": round(n / t_batched, 3),
                    "gcell_per_s": round(work / t_again / 1e9, 2),
                    "hbm_frac": round(work / t_again * 32 / 1e9 / HBM_PEAK_GBS, 4)},
        "plain_one_at_a_time": {"seconds": round(t_plain, 3), "solves_per_s": round(n / t_plain, 3),
                                "gcell_per_s": round(work / t_plain / 1e9, 2)},
        "solves_per_s_ratio": round(t_plain / t_again, 3),
        "solves_per_s_ratio_incl_build": round(t_plain / t_batched, 3),
    }


def config1_bench(dev, reps: int = 20) -> dict:
    """The reference's CPU-runnable case on the GPU: 96x96 (117x117 padded,
    CPML 10), eps_r 4, one shot, nt 300, forward only -- launch-latency
    bound.  Plain operator calls against a one-problem batched plan whose
    time loop replays a captured HIP graph."""
    import torch

    import tide
    from tide.batch import TM2DBatch
    from tide.maxwell.tm2d import maxwelltm

    n, nt, dt, dx = 96, 300, 4e-11, 0.02
    kw = dict(epsilon=torch.full((n, n), 4.0, device=dev), sigma=torch.zeros(n, n, device=dev),
              mu=torch.ones(n, n, device=dev), grid_spacing=dx, dt=dt,
              source_amplitude=tide.ricker(8e8, nt, dt, dtype=torch.float32, device=dev).reshape(1, 1, -1),
              source_location=torch.tensor([[[20, 48]]]), receiver_location=torch.tensor([[[20, 60]]]),
              pml_width=10)
    updates = 117 * 117 * nt

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    with torch.no_grad():
        t_plain = timed(lambda: maxwelltm(**kw))
    plan = TM2DBatch([kw], gradients=False)

    def run_plan():
        plan.reset_state()
        plan.forward()

    t_graph = timed(run_plan)
    plan.close()
    # where the plain call's time goes: the host set-up (padding,
    # coefficients, profiles, point indices) and the native time loop alone
    # (150 two-step launches on a prepared setup)
    from tide.maxwell.tm2d import _run_forward, prepare_tm2d

    args = (kw["epsilon"], kw["sigma"], kw["mu"], dx, dt, kw["source_amplitude"],
            kw["source_location"], kw["receiver_location"])
    t_setup = timed(lambda: prepare_tm2d(*args, pml_width=10))
    st = prepare_tm2d(*args, pml_width=10)
    r = torch.zeros(nt, 1, 1, device=dev)
    t_loop = timed(lambda: _run_forward(st, st.ca, st.cb, st.cq, st.f, st.fields, r, interval=1,
                                        fmt=1))
    return {
        "workload": "96x96 (117x117 padded, CPML 10), eps_r 4, 1 shot, nt 300, forward, fp32",
        "plain_ms": round(t_plain * 1e3, 3), "plain_us_per_step": round(t_plain / nt * 1e6, 2),
        "plain_setup_ms": round(t_setup * 1e3, 3),
        "plain_native_loop_ms": round(t_loop * 1e3, 3),
        "graph_ms": round(t_graph * 1e3, 3), "graph_us_per_step": round(t_graph / nt * 1e6, 2),
        "graph_gcell_per_s": round(updates / t_graph / 1e9, 3),
    }


def em3d_bench(dev, nt: int, shots: int) -> dict:
    """3-D Maxwell forward + adjoint gradient on the 256^3 workload (281^3
    padded, CPML 12, dx 0.02 m, dt at CFL 0.9): seeded 8-layer eps_r 1-25
    with 20 ellipsoidal anomalies, sigma U[0, 0.01], mu_r 1, x-directed
    dipole sources (300 MHz Ricker) near the top face, 32 x 32 receivers (Ex)
    on it; fp32 fields and snapshots.  Through the public operator: the
    observed data (a model with one more anomaly) is one timed forward; the
    gradient of 0.5 |d - d_obs|^2 is one timed ``op.linearize`` call, which
    checkpoints the state every ~sqrt(nt) steps when the snapshots do not
    fit HBM and recomputes each segment's snapshots in the backward.  Useful
    work = the forward and adjoint sweeps; the recompute factor is the
    forward sweeps run per useful forward sweep."""
    import torch

    import tide
    from tide.synthetic import layered_model_3d

    n, pml, dx = 256, 12, 0.02
    dt = 0.9 * dx / (3**0.5 * tide.C0)
    eps, sig, mu = layered_model_3d(n, n, n, seed=5005, device=dev)
    zz = torch.arange(n, dtype=torch.float64, device=dev)
    bump = 3.0 * torch.exp(-0.5 * (((zz[:, None, None] - 60) / 20) ** 2
                                   + ((zz[None, :, None] - 128) / 30) ** 2
                                   + ((zz[None, None, :] - 128) / 30) ** 2))
    eps_obs = (eps.double() + bump).to(eps.dtype)
    del bump
    amp = tide.ricker(3e8, nt, dt, dtype=torch.float32, device=dev).reshape(1, 1, -1).repeat(shots, 1, 1)
    src = torch.stack([torch.tensor([[2, 64 + 32 * (k % 4), 64 + 32 * (k // 4)]]) for k in range(shots)])
    yy, xx = torch.meshgrid(torch.arange(32) * 8 + 2, torch.arange(32) * 8 + 2, indexing="ij")
    rec = torch.stack([torch.full_like(yy, 2), yy, xx], -1).reshape(1, -1, 3).repeat(shots, 1, 1)
    disc = tide.Discretization(spacing=dx, dt=dt, boundary=tide.CPML(pml))

    def operator(n_steps):
        exp = tide.Experiment(tide.Acquisition(src.to(dev), rec.to(dev)), amp[..., :n_steps],
                              source_component="ex", receiver_component="ex")
        return tide.Maxwell3D(disc, exp)

    updates = float((n + 2 * pml + 1) ** 3 * shots * nt)
    # warm-up on a few steps (kernel loading, allocator pools)
    small = operator(8)
    with torch.no_grad():
        obs_small = small(tide.EMModel(eps_obs, sig, mu)).receiver_data
    with small.linearize(tide.EMModel(eps, sig, mu)) as lin:
        tide.workflow.ReceiverObjective(obs_small).value_and_gradient(lin)
    del small, obs_small
    op = operator(nt)
    torch.cuda.synchronize()
    log(f"3-D: forward of {shots} shots x {nt} steps")
    t0 = time.perf_counter()
    with torch.no_grad():
        observed = op(tide.EMModel(eps_obs, sig, mu)).receiver_data
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    del eps_obs
    log(f"3-D: forward {t_fwd:.2f} s; gradient")
    objective = tide.workflow.ReceiverObjective(observed)
    t0 = time.perf_counter()
    with op.linearize(tide.EMModel(eps, sig, mu)) as lin:
        misfit, grad = objective.value_and_gradient(lin)
        plan = dict(getattr(lin, "last_plan", {}))
    torch.cuda.synchronize()
    t_grad = time.perf_counter() - t0
    ok = bool(torch.isfinite(grad.epsilon).all() and torch.isfinite(grad.sigma).all())
    log(f"3-D: gradient {t_grad:.2f} s, plan {plan}")
    seg = int(plan.get("segment") or 0)
    policy = plan.get("policy", "?")
    # checkpointing runs the forward twice (checkpoints, then each segment's
    # snapshots); full storage once
    recompute = 2.0 if policy == "checkpoint" else 1.0
    gbytes = 48 + 24
    return {
        "workload": f"3D Maxwell 256^3 (281^3 padded, CPML 12), 8-layer eps_r 1-25 + 20 ellipsoidal "
                    f"anomalies, {shots} shots x 1024 receivers, nt {nt}, fp32, fp32 snapshots",
        "forward": {"seconds": round(t_fwd, 3), "gcell_per_s": round(updates / t_fwd / 1e9, 2),
                    "hbm_frac_48B": round(updates / t_fwd * 48 / 1e9 / HBM_PEAK_GBS, 4)},
        "gradient": {"seconds": round(t_grad, 3),
                     "gcell_per_s": round(2 * updates / t_grad / 1e9, 2),
                     "policy": policy, "segment_steps": seg,
                     "segments": (nt + seg - 1) // seg if seg else 0,
                     "shot_chunks": plan.get("shot_chunks"),
                     "recompute_factor": recompute,
                     "finite": ok, "misfit": float(misfit),
                     "note": f"useful work = forward + adjoint sweeps ({gbytes} algorithmic B per "
                             "forward+adjoint cell-update pair at one sweep each); the recompute "
                             "sweep is overhead"},
    }


def accuracy_bench(dev, eps, eps_obs, sig, mu, dx, dt, pml, amp, src, rec, shots: int) -> dict:
    """Error of the fp32 gradient step on a bounded sample of the headline
    workload (the first ``shots`` shots, full grid and nt): receivers and
    eps / sigma gradients of the same least-squares objective against an
    FP64 run (the same HIP kernels in double precision, which match the
    FP64 oracle to 1e-12 in the parity tests), and the extra gradient error
    of BF16-compressed snapshots."""
    import torch

    import tide

    disc = tide.Discretization(spacing=dx, dt=dt, boundary=tide.CPML(pml))

    def run(dtype, compression=False):
        c = lambda t: t.to(dtype)  # noqa: E731
        exp = tide.Experiment(tide.Acquisition(src[:shots], rec[:shots]), c(amp[:shots]))
        op = tide.MaxwellTM(disc, exp, storage=tide.StorageOptions(compression=compression))
        with torch.no_grad():
            obs = op(tide.EMModel(c(eps_obs), c(sig), c(mu))).receiver_data
        with op.linearize(tide.EMModel(c(eps), c(sig), c(mu))) as lin:
            value, g = tide.workflow.ReceiverObjective(obs).value_and_gradient(lin)
            r = lin.primal.receiver_data
        out = [x.detach().double().cpu() for x in (r, g.epsilon, g.sigma)]
        del op, lin, g
        torch.cuda.empty_cache()
        return out + [float(value)]

    def rel(a, b):  # relative L2; the absolute one when the reference is zero (short runs)
        d, n = float((a - b).norm()), float(b.norm())
        return d / n if n > 0 else d

    def cos(a, b):
        n = float(a.norm() * b.norm())
        return float((a * b).sum()) / n if n > 0 else 1.0

    r64, e64, s64, v64 = run(torch.float64)
    r32, e32, s32, v32 = run(torch.float32)
    rb, eb, sb, vb = run(torch.float32, "bf16")
    return {
        "sample": f"first {shots} shots of the workload, full grid and nt",
        "fp32_vs_fp64": {"receivers_rel_l2": rel(r32, r64),
                         "misfit_rel": abs(v32 - v64) / abs(v64) if v64 != 0 else abs(v32 - v64),
                         "grad_eps_rel_l2": rel(e32, e64), "grad_sigma_rel_l2": rel(s32, s64)},
        "bf16_snapshots_vs_fp32": {"grad_eps_rel_l2": rel(eb, e32), "grad_sigma_rel_l2": rel(sb, s32),
                                   "grad_eps_cosine": cos(eb, e32)},
        "bf16_snapshots_vs_fp64": {"grad_eps_rel_l2": rel(eb, e64), "grad_sigma_rel_l2": rel(sb, s64)},
        "fp64_reference": "the same HIP kernels in double precision (FP64 fields, FP64 gradient "
                          "accumulation); parity-tested against the FP64 oracle to 1e-12",
    }


def main() -> None:
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU; multi-GPU runs are not part of this build "
              "(shots shard across GPUs with one FP64 gradient all-reduce)", file=sys.stderr)
        sys.exit(2)

    import torch

    import tide
    from tide.maxwell.tm2d import KernelTimer, set_kernel_timer
    from tide.synthetic import add_anomaly, cfl_dt, layered_model, surface_acquisition

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.em3d_only:
        print(json.dumps({"em3d": em3d_bench(dev, args.em3d_nt, args.em3d_shots)}), flush=True)
        return
    n = args.size
    pml, dx = 20, 0.01
    dt = cfl_dt(dx, 2)
    nt, n_shots = args.nt, args.shots
    eps, sig, mu = layered_model(n, n, seed=1234, device=dev)
    # the perturbation sits where its reflection reaches the surface within nt
    eps_obs = add_anomaly(eps, n * 0.12, n * 0.45, 30.0, 2.0)
    src, rec = surface_acquisition(n_shots, 256, device=dev)
    amp = tide.ricker(2e8, nt, dt, dtype=torch.float32, device=dev).reshape(1, 1, -1).repeat(n_shots, 1, 1)
    ny_p, nx_p = n + 2 * pml + 1, n + 2 * pml + 1

    snap_bytes = 2 if args.compression == "bf16" else 4
    from tide.autotune import _install, current
    from tide.maxwell.tm2d import device_snapshot_budget, plan_shot_chunks

    disc = tide.Discretization(spacing=dx, dt=dt, boundary=tide.CPML(pml))
    per_shot = 2 * nt * ny_p * nx_p * snap_bytes
    # --chunk: a device byte limit that holds exactly that many shots'
    # snapshots; by default the operator sizes its chunks from free HBM.
    limit = args.chunk * per_shot if args.chunk else None
    storage = tide.StorageOptions(compression=args.compression == "bf16", bytes_limit_device=limit)
    op = tide.MaxwellTM(disc, tide.Experiment(tide.Acquisition(src, rec), amp), storage=storage)
    with torch.no_grad():
        observed = op(tide.EMModel(eps_obs, sig, mu)).receiver_data
    torch.cuda.synchronize()
    log(f"observed data ready ({n_shots} shots, one operator call)")
    # the chunks the operator will run (same rule as inside the call)
    budget = device_snapshot_budget(dev, limit, 21 * ny_p * nx_p * 4, n_shots)
    default_shapes = current()  # the library's defaults, restored for the other workloads
    plan = plan_shot_chunks(n_shots, per_shot, budget, group=max(1, default_shapes[1][1]))
    sizes = plan.sizes
    tuning = None
    if not args.no_autotune:
        from tide.autotune import tune_tm2d

        t = tune_tm2d(ny_p, nx_p, sizes, device=dev, max_store_bytes=64 << 30)
        log(f"autotuned on chunks {sizes}: forward {t['forward']} adjoint {t['adjoint']} "
            f"{t['chosen_us_per_step']} us per step (tuned in {t.get('seconds', 0)} s)")
        tuning = t
        _install(t["forward"], t["adjoint"])
    else:
        fwd, adj = current()
        tuning = {"forward": fwd, "adjoint": adj, "cached": False}
    n_chunks = len(sizes)

    objective = tide.workflow.ReceiverObjective(observed)
    model = tide.EMModel(eps, sig, mu)
    chunks_run = []

    def step():
        # the public API: J = 0.5 |d(m) - d_obs|^2 and dJ/d(eps, sigma) over all
        # shots of one operator; the operator runs them in HBM-sized chunks
        # (forward with snapshots, residual, exact adjoint per chunk), one
        # FP64 gradient accumulator across chunks
        with op.linearize(model) as lin:
            misfit, grad = objective.value_and_gradient(lin)
            chunks_run[:] = getattr(lin, "last_plan", {}).get("shot_chunks", [])
        return misfit

    for k in range(args.warmup):
        step()
        torch.cuda.synchronize()
        log(f"warmup step {k + 1}/{args.warmup} done")
    from tide import _native

    timer = KernelTimer()
    set_kernel_timer(timer)
    _native.set_tm2d_loop_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        misfit = step()
        log(f"timed step {k + 1}/{args.steps} enqueued")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    log(f"timed region {elapsed:.2f} s")
    set_kernel_timer(None)
    # HIP events around the kernel launches of every time loop (no set-up,
    # gradient fold or copies), on the stream the kernels run on
    loop_ms = {"forward_store": _native.take_tm2d_loop_timing(0),
               "adjoint": _native.take_tm2d_loop_timing(1)}
    _native.set_tm2d_loop_timing(False)

    if args.calibrate_copy:
        a = torch.empty(1 << 28, device=dev, dtype=torch.float32)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    updates_per_step = 2.0 * ny_p * nx_p * n_shots * nt
    gcells = updates_per_step * args.steps / elapsed / 1e9
    summ = timer.summary()
    b_snap = 2 * snap_bytes
    b_alg = {"forward_store": 24 + b_snap, "adjoint": 24 + b_snap, "forward": 24}
    kernels = {}
    # the fp32 K-step adjoint held to four waves per SIMD (this workload is
    # fp32) is its own kernel
    adj_w4 = _native.get_tm2d_adjoint_waves() == 4 and 64 * _native.get_tm2d_tuning()[3] <= 256
    for name, d in summ.items():
        spl = int(d["steps_per_launch"])
        k_ms, k_launches, _ = loop_ms.get(name, (d["ms"], d["launches"], 0))
        achieved = b_alg[name] * d["cell_updates"] / (k_ms * 1e-3) / 1e9
        stem = "adj" if name == "adjoint" else "fwd"
        kname = f"{stem}_step" if spl == 1 else (f"{stem}2_step" if spl == 2 else f"{stem}k_step")
        if stem == "adj" and spl >= 3 and adj_w4:
            kname = "adjk_step_w4"
        kernels[name] = {
            "kernel": kname,
            "steps_per_launch": spl,
            "launches": k_launches,
            "avg_launch_us": round(k_ms / max(k_launches, 1) * 1e3, 2),
            "call_avg_launch_us": round(d["ms"] / d["launches"] * 1e3, 2),
            "bytes_per_cell_update": b_alg[name],
            "achieved_GBs": round(achieved, 1),
            "gcell_per_s": round(d["cell_updates"] / (k_ms * 1e-3) / 1e9, 2),
        }
    dom = max(summ, key=lambda k: summ[k]["ms"])
    dom_k = kernels[dom]
    shots_per_launch = n_shots / n_chunks
    fmt = "bf16" if snap_bytes == 2 else "fp32"
    measured = counted_traffic().get(f"{dom_k['kernel']}:{dom_k['steps_per_launch']}:{fmt}")
    cells_per_launch = ny_p * nx_p * shots_per_launch * dom_k["steps_per_launch"]
    traffic = None if measured is None else round(measured * cells_per_launch)
    roofline = {
        "bound": "hbm",
        "achieved": dom_k["achieved_GBs"],
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(dom_k["achieved_GBs"] / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "traffic_unit": "bytes per launch: rocprofv3 FETCH_SIZE + WRITE_SIZE (L2-to-fabric, "
                        "MALL hits included), copy-calibrated, from " + TRAFFIC_SUMMARY.name,
        "algorithmic_bytes_per_launch": round(b_alg[dom] * cells_per_launch),
        "kernel": dom_k["kernel"],
        "kernels": kernels,
    }
    setup_kwargs = dict(
        epsilon=eps, sigma=sig, mu=mu, grid_spacing=dx, dt=dt, source_amplitude=amp,
        source_location=src, receiver_location=rec, pml_width=pml,
    )
    log("timing the CPU baseline" if not args.no_cpu_baseline else "CPU baseline skipped")
    cpu = None if args.no_cpu_baseline else cpu_baseline(args, setup_kwargs)
    del op, observed, model, setup_kwargs
    torch.cuda.empty_cache()
    accuracy = None
    if args.accuracy_shots > 0:
        log(f"accuracy sample: {args.accuracy_shots} shots in fp64, fp32 and BF16 snapshots")
        accuracy = accuracy_bench(dev, eps, eps_obs, sig, mu, dx, dt, pml, amp, src, rec,
                                  args.accuracy_shots)
        log(f"accuracy: {accuracy}")
    # the shapes tuned above fit this workload's grid and shot chunks; the
    # mixed-size workloads below run with the library's defaults
    _install(*default_shapes)
    serving = None
    if args.serve_requests > 0:
        serving = [serve_bench(args, dev, float(r)) for r in args.serve_rates.split(",")]
    batched = batched_bench(args, dev) if args.batched_models > 0 else None
    config1 = None if args.no_config1 else config1_bench(dev)
    if config1 is not None:
        log(f"config 1: {config1}")
    em3d = em3d_bench(dev, args.em3d_nt, args.em3d_shots) if args.em3d_nt > 0 else None
    if em3d is not None:
        log(f"3-D: {em3d}")
    out = {
        "metric": METRIC,
        "value": round(gcells, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": f"2D TM fwd+adj gradient, {n}x{n} model ({ny_p}x{nx_p} padded, CPML {pml}), "
                        f"{n_shots} shots x 256 receivers, nt {nt}, stencil 2, "
                        f"{'bf16' if snap_bytes == 2 else 'fp32'} snapshots",
            "shot_chunks": chunks_run or sizes,
            "tuned_us_per_step": tuning.get("chosen_us_per_step"),
            "cell_updates_per_step": updates_per_step,
            "hbm_frac_whole_job": round(gcells * 1e9 * (24 + b_snap) / 1e9 / HBM_PEAK_GBS, 4),
            "misfit": float(misfit),
            "launch_shape": {
                "fields": ["rows", "group", "shot_loop", "steps_per_launch", "tile_split",
                           "border_rows"],
                "forward": list(tuning["forward"]),
                "adjoint": list(tuning["adjoint"]),
            },
            "seed": 1234,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        # the metric's second part: served-solve latency at the first finite
        # arrival rate of the sweep (details in "serving")
        "served_latency": next(({"rate_per_s": s["rate_per_s"], "requests": s["served"]["solves"],
                                 "p50_s": s["served"]["p50_s"], "p99_s": s["served"]["p99_s"],
                                 "plain_p50_s": s["plain_one_at_a_time"]["p50_s"],
                                 "plain_p99_s": s["plain_one_at_a_time"]["p99_s"]}
                                for s in (serving or [])
                                if s.get("rate_per_s") != "inf" and "plain_one_at_a_time" in s), None),
        "serving": serving,
        "batched": batched,
        "config1": config1,
        "em3d": em3d,
        "accuracy": accuracy,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
