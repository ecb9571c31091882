"""Throughput of the FP64 ETO-HEOM RK4 propagation on one MI355X.

A bench "step" is one whole propagation of the workload (every RK4 step of it,
V(t) built and P(t) read out on the device), with the hierarchy state resident
in HBM.  The default workload is the heptamer: ground + 7 singles + 21 doubles
(N=29), 2 ETO modes/site (d=14), depth L=4 (3060 ADOs), 2000 RK4 steps of
0.5 fs, 3 Gaussian pulses; `value` is ADO-updates/s (one ADO advanced one full
RK4 step).  Other workloads: --workload dimer | grid2d | chain10 | chain_grid | trimer | trimer_grad.

Printed as one JSON line with the roofline of the dominant kernel (the fused
RHS+RK4 stage kernel, timed per launch with HIP events on its own stream) and
the CPU oracle timed on this host over a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "eto-heom_amd"))

HBM_PEAK_GBS = 8000.0


def workload(name, nsteps=None):
    from etoheom_amd import configs
    if nsteps:   # shortened time grid, for counter runs only
        items, desc = workload(name)
        for p in items:
            p.t_end = p.t_start + p.step_size * (nsteps - 1)
        return items, desc + f" (shortened to {nsteps} RK4 steps)"
    if name == "heptamer":
        return [configs.heptamer()], "heptamer N=29 (ground+7+21), d=14, L=4, 3060 ADOs, 2000 RK4 steps dt=0.5 fs"
    if name == "dimer":
        return [configs.dimer()], "dimer N=4, d=4, L=6, 210 ADOs, 2000 RK4 steps dt=0.5 fs"
    if name == "trimer":
        return [configs.aggregate(3, 5)], "trimer N=7, d=6, L=5, 462 ADOs, 2000 RK4 steps"
    if name == "chain10":
        return [configs.chain()], "10-site chain N=56, d=20, L=3, 1771 ADOs, 3000 RK4 steps"
    if name == "chain_grid":
        return (configs.chain_grid(n_t1=4),
                "10-site chain N=56, L=3, 1771 ADOs, 3000 RK4 steps: 16 propagations (t1 = 0..30 fs x t2 = "
                "0/100/200/400 fs) of the 64 x 4 grid in one batch")
    if name == "trimer_grad":
        return None, ("trimer N=7 (ground+3+3), d=6, L=5, 462 ADOs: 16x16 (t1, t3) rephasing grid at t2=100 fs, "
                      "48 propagations, loss gradient by the adjoint (checkpoint every 32 steps)")
    if name == "grid2d":
        return configs.grid2d(), "512 propagations N=4/7/11, L=3..6, static disorder, 2000 RK4 steps each"
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(p, rho, seconds=20.0, max_threads=16):
    """The oracle (a scalar FP64 port of the reference CPU path, with its dense
    one-hot GEMMs and hashed neighbour lookups) on this host's cores, one
    independent propagation per thread, as the reference farms one process per
    (tau, T) point.  It starts from the GPU's final hierarchy state, so every ADO is
    dense as in the bulk of a propagation (the first steps, before the pulses, have
    an all-zero hierarchy that the reference's zero-skipping GEMM passes over), and
    runs whole RK4 steps after the last pulse.  A one-thread pass sizes the sample
    (about seconds/3), then every thread runs that many steps at once."""
    import threading
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    q = O.Problem(H=p.H, X=p.X, alpha=p.alpha, alpha_t=p.alpha_t, gamma=p.gamma, L=p.L,
                  site_offset=p.site_offset, pulses=p.pulses, t_start=p.t_end, t_end=p.t_end + 10000.0,
                  step_size=p.step_size, print_step=p.print_step)
    M = q.num_ado()
    steps, elapsed = 1, 0.0
    while True:
        t0 = time.perf_counter()
        O.twod(q, nv=(1, 1, 1), rho_init=rho, max_steps=steps)
        elapsed = time.perf_counter() - t0
        if elapsed > seconds / 3 or steps >= 4096:
            break
        steps *= 4 if elapsed < seconds / 30 else 2
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count() or 1
    nthr = max(1, min(max_threads, ncpu))
    # the C oracle keeps no global state and ctypes releases the GIL during the call
    threads = [threading.Thread(target=O.twod, args=(q,), kwargs=dict(nv=(1, 1, 1), rho_init=rho, max_steps=steps))
               for _ in range(nthr)]
    t0 = time.perf_counter()
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    wall = time.perf_counter() - t0
    return {"value": nthr * M * steps / wall, "unit": "ADO-updates/s", "cores": nthr, "kind": "port",
            "single_thread_value": M * steps / elapsed,
            "sample": f"{nthr} threads x {steps} RK4 steps of the same instance ({M} ADOs) from its dense final "
                      f"state, {wall:.1f} s wall (one thread alone: {elapsed:.1f} s)"}


def per_propagation(ctx, items, k):
    """The plain per-launch GPU path on k of the workload's propagations (evenly
    spaced over the list): each one alone, its kernels launched one by one from
    the host, step after step.  Set-up (hierarchy, tables, uploads) is outside the
    timed region, as it is for the batched path."""
    import numpy as np
    import etoheom_amd as E
    idx = sorted(set(np.linspace(0, len(items) - 1, k).astype(int).tolist()))
    bs = [E.Batch(ctx, [{"problem": items[i], "H": items[i].H, "X": items[i].X, "nv": (1, 1, 1)}]) for i in idx]
    bs[0].run_eager()   # warm-up
    bs[0].sync()
    t0 = time.perf_counter()
    for b in bs:
        b.run_eager()
        b.sync()
    wall = time.perf_counter() - t0
    for b in bs:
        b.close()
    return {"propagations": len(idx), "seconds": wall, "propagations_per_s": len(idx) / wall,
            "sample": f"{len(idx)} of {len(items)} propagations, evenly spaced, one at a time, per-launch (no graph)"}


def traffic_of(workload):
    """HBM bytes per stage launch from the committed rocprofv3 --pmc summary, if any."""
    f = os.path.join(ROOT, "profiles", f"{workload}_traffic.json")
    if not os.path.exists(f):
        return None, None
    d = json.load(open(f))
    return d["traffic_bytes"], d["source"]


def bench_gradient(a):
    """Forward + adjoint of the trimer 2DES-grid loss; reports the cost ratio
    (forward + backward) / forward and the equivalent central-FD cost."""
    import numpy as np
    import torch
    import etoheom_amd as E
    from etoheom_amd import spectra as SP
    ctx = E.Context(0)
    spec = SP.grid_spec(ctx)
    with torch.no_grad():
        target = SP.grid_signal(SP.trimer_params(perturb=5), spec)
    import gc
    fw, bw, gfw, gbw, gpf = [], [], [], [], []
    for it in range(a.warmup + a.steps):
        params = SP.trimer_params()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        S = SP.grid_signal(params, spec)
        loss = SP.grid_loss(S, target)
        t1 = time.perf_counter()
        loss.backward()
        t2 = time.perf_counter()
        if it >= a.warmup:
            fw.append(t1 - t0)
            bw.append(t2 - t1)
            gfw.append(spec.timing["forward_ms"] * 1e-3)
            gbw.append(spec.timing["backward_ms"] * 1e-3)
        lval = float(loss.detach())
        del S, loss   # the graph holds the batch (and its trajectory): free it before the next forward
        gc.collect()
        with torch.no_grad():   # one plain forward solve (no adjoint state) of the same grid
            SP.grid_signal(params, spec)
        if it >= a.warmup:
            gpf.append(spec.timing["forward_ms"] * 1e-3)
    n_params = sum(p.numel() for p in params.values())
    f, b = float(np.mean(fw)), float(np.mean(bw))
    gf, gb, gp = float(np.mean(gfw)), float(np.mean(gbw)), float(np.mean(gpf))
    # ADO-updates of the forward: 48 propagations x 462 ADOs x their steps
    steps = sum(int(np.floor(np.float32(np.float32(te) / np.float32(spec.step_size)))) + 1 for te in spec.t_end)
    ups = 462 * steps
    rec = {"metric": "FP64 ADO-updates/s + 2DES grid points/s at 1e-10 err; adjoint-grad time vs FD",
           "value": ups / gf, "unit": "ADO-updates/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": (f + b) * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": workload("trimer_grad")[1], "device": ctx.name,
                      "grid_points_per_s": 16 / gf,
                      "adjoint": {"forward_gpu_s": gf, "backward_gpu_s": gb, "gradient_over_forward": (gf + gb) / gf,
                                  "plain_forward_gpu_s": gp, "gradient_over_plain_forward": (gf + gb) / gp,
                                  "forward_wall_s": f, "backward_wall_s": b,
                                  "gradient_over_forward_wall": (f + b) / f,
                                  "n_params": n_params, "central_fd_s": 2 * n_params * gf,
                                  "fd_over_adjoint": 2 * n_params * gf / (gf + gb), "loss": lval,
                                  "note": "GPU: HIP events around the propagation and the backward pass; "
                                          "forward = the gradient run's forward (it also keeps the adjoint's "
                                          "trajectory), plain forward = the same grid with no gradient wanted; "
                                          "wall: incl. host set-up of each batch (tables, graph capture, "
                                          "allocation)"}}}
    print(json.dumps(rec))
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="heptamer")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--nsteps", type=int, default=0, help="shorten the propagation (counter runs only)")
    ap.add_argument("--autotune", type=int, default=0, help="autotuner steps per launch candidate (0: off)")
    ap.add_argument("--per-propagation", type=int, default=0,
                    help="also time this many of the workload's propagations one at a time on the plain "
                         "per-launch path (no graph, no batching)")
    ap.add_argument("--copy-calibrate", action="store_true",
                    help="also run a plain 2 GiB read + 2 GiB write elementwise copy (counter scale)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}); independent propagations shard "
              "across GPUs with no communication but the final gather of P(t).", file=sys.stderr)
        return 2

    import numpy as np
    import torch
    import etoheom_amd as E

    copy_gbs = None
    if a.copy_calibrate:
        # a plain streaming copy of the same precision: the HBM rate this device reaches
        x = torch.ones(1 << 28, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.mul(x, 1.0, out=y)
        e0.record()
        for _ in range(3):
            torch.mul(x, 1.0, out=y)
        e1.record()
        torch.cuda.synchronize()
        copy_gbs = 3 * 2 * x.numel() * 8 / (e0.elapsed_time(e1) * 1e-3) / 1e9
        del x, y
    if a.workload == "trimer_grad":
        return bench_gradient(a)
    items, desc = workload(a.workload, a.nsteps)
    ctx = E.Context(0)
    # end to end from host buffers: hierarchy and tables built on the host, uploaded,
    # graph captured, propagated, readouts copied back (creation is outside `value`)
    torch.cuda.synchronize()
    t_e2e = time.perf_counter()
    batch = E.Batch(ctx, [{"problem": p, "H": p.H, "X": p.X, "nv": (1, 1, 1)} for p in items])
    batch.run()
    batch.sync()
    _ = [batch.output(i) for i in range(len(items))]
    t_e2e = time.perf_counter() - t_e2e
    tuning = batch.autotune(a.autotune).splitlines() if a.autotune else None
    for _ in range(a.warmup):
        batch.run()
        batch.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        batch.run()
    batch.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = batch.elapsed_ms()
    outs = [batch.output(i) for i in range(len(items))]
    rho_final = batch.rho(0) if not a.no_cpu else None   # dense end-of-run state for the CPU sample
    finite = all(np.all(np.isfinite(o)) for o in outs)

    # per-launch durations of the stage kernel (HIP events on the batch stream)
    prof = batch.profile(a.profile_steps)
    stage_ms = float(np.mean(prof[1:5]))
    elems = sum(p.num_ado() * p.N * p.N for p in items)
    # the algorithmic model: 256 N^2 B per ADO-update (9 reads + 7 writes of the N x N
    # block per RK4 step) over 4 stage launches; the kernels' own schedule needs 7 + 5
    # (192 N^2 B), reported beside it
    bytes_per_launch = 64.0 * elems
    schedule_bytes = 48.0 * elems
    achieved = bytes_per_launch / (stage_ms * 1e-3) / 1e9

    per_prop = per_propagation(ctx, items, a.per_propagation) if a.per_propagation else None

    traffic, tsrc = traffic_of(a.workload)
    ups = batch.ado_updates * a.steps
    value = ups / wall
    rec = {
        "metric": "FP64 ADO-updates/s + 2DES grid points/s at 1e-10 err; adjoint-grad time vs FD",
        "value": value,
        "unit": "ADO-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": desc, "propagations_per_step": len(items),
                   "propagations_per_s": len(items) * a.steps / wall,
                   "grid_points_per_s": len(items) * a.steps / wall / 3.0,
                   "gpu_event_ms_last_run": gpu_ms, "device": ctx.name, "finite": finite,
                   "groups": batch.tuning_report().splitlines(),
                   "end_to_end_from_host": {"seconds": t_e2e, "ado_updates_per_s": batch.ado_updates / t_e2e,
                                            "note": "first run incl. host hierarchy/table build, upload, graph "
                                                    "capture and readout copy-back; before autotuning"}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": tsrc,
                     "kernel": "stage (fused RHS + RK4 stage)", "stage_ms": [float(x) for x in prof[1:5]],
                     "prologue_ms": float(prof[0]), "bytes_per_launch": bytes_per_launch,
                     "schedule_bytes_per_launch": schedule_bytes,
                     "schedule_achieved": schedule_bytes / (stage_ms * 1e-3) / 1e9,
                     # the counters' bytes per launch (L2 misses: HBM or the last-level cache)
                     # over this run's launch time: the rate to quote when the state fits on chip
                     "traffic_achieved": traffic / (stage_ms * 1e-3) / 1e9 if traffic else None},
    }
    if copy_gbs:
        rec["roofline"]["copy_gbs"] = copy_gbs
    if tuning:
        rec["config"]["autotune_steps"] = a.autotune
    if per_prop:
        per_prop["batched_over_per_propagation"] = rec["config"]["propagations_per_s"] / per_prop["propagations_per_s"]
        rec["config"]["per_propagation_path"] = per_prop
    if not a.no_cpu:
        rec["cpu_baseline"] = cpu_baseline(items[0], rho_final)
    print(json.dumps(rec))
    return 0


if __name__ == "__main__":
    sys.exit(main())
