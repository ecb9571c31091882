"""Throughput of the FP64 ETO-HEOM RK4 propagation on one MI355X.

The default workload is the largest single-GPU configuration: the 10-site chain
(ground + 10 singles + 45 doubles, N=56), 2 ETO modes/site (d=20), depth L=3
(1771 ADOs), as a batch of 64 t1 x 4 t2 = 256 propagations of 3000 RK4 steps of
0.5 fs (about 91 GB of hierarchy state), each with its three Gaussian pulses.

A bench "step" is the next S RK4 steps of every instance (one replay of the
forward graph captured with S steps; V(t) built and P(t) read out on the
device), with the state resident in HBM: S = 50 for the chain grid, whose 5
warm-up steps run RK4 steps 0-249 from a reset and whose 20 timed steps run
250-1249 (the first pulse windows among them); S = ceil(max_steps / 20) for the
other workloads.  --full times one whole propagation from a reset instead.
`value` is ADO-updates/s (one ADO advanced one full RK4 step).  Other workloads:
--workload heptamer | dimer | grid2d | chain10 | trimer | trimer_grad | chain_grad.

Printed as one JSON line with the roofline of the dominant kernel (the fused
RHS+RK4 stage kernel, timed per launch with HIP events on its stream), the CPU
oracle timed on this host over a bounded sample of the same workload, one
process per core, and for the chain grid the adjoint leg: the trimer grid
gradient (48 propagations, L = 5) against its forward and central differences.
"""
import argparse
import json
import math
import os
import subprocess
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
        return (configs.chain_grid(n_t1=64),
                "10-site chain N=56 (ground+10+45), d=20, L=3, 1771 ADOs, 3000 RK4 steps dt=0.5 fs: the 64 t1 "
                "(0..630 fs) x 4 t2 (0/100/200/400 fs) grid, 256 propagations in one batch")
    if name == "chain_grid16":
        return (configs.chain_grid(n_t1=4),
                "10-site chain N=56, L=3, 1771 ADOs, 3000 RK4 steps: 16 propagations (t1 = 0..30 fs x t2 = "
                "0/100/200/400 fs) of the 64 x 4 grid in one batch")
    if name == "chain_grad":
        return None, ("10-site chain N=56, d=20, L=3, 1771 ADOs, 3000 RK4 steps: 16 propagations (t1 = 0..30 fs x "
                      "t2 = 0/100/200/400 fs) of its grid, loss gradient w.r.t. the 10 site energies and 9 couplings "
                      "by the adjoint (checkpoint every 32 steps, sub-batches sized to the device memory)")
    if name == "trimer_grad":
        return None, ("trimer N=7 (ground+3+3), d=6, L=5, 462 ADOs: 16x16 (t1, t3) rephasing grid at t2=100 fs, "
                      "48 propagations, loss gradient by the adjoint (checkpoint every 32 steps)")
    if name == "grid2d":
        return configs.grid2d(), "512 propagations N=4/7/11, L=3..6, static disorder, 2000 RK4 steps each"
    raise SystemExit(f"unknown workload {name}")


def first_instance(name):
    """The workload's first propagation, built alone (the CPU sample's instance)."""
    from etoheom_amd import configs
    if name == "chain_grid" or name == "chain_grid16":
        return configs.chain(t1=0.0, t2=0.0)
    items, _ = workload(name)
    return items[0]


def cpu_worker(name, seconds, seed):
    """One process of the CPU baseline: the oracle (a scalar FP64 port of the
    reference CPU path, with its dense one-hot GEMMs and hashed neighbour lookups)
    on the workload's first instance from a dense random hierarchy state (the
    reference's GEMM skips zero entries, so the all-zero hierarchy before the
    pulses would understate its cost), after the last pulse.  Waits for a line on
    stdin, then doubles its step count from 1 until a run takes seconds/3, and
    prints that run's steps and time."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    p = first_instance(name)
    q = O.Problem(H=p.H, X=p.X, alpha=p.alpha, alpha_t=p.alpha_t, gamma=p.gamma, L=p.L,
                  site_offset=p.site_offset, pulses=p.pulses, t_start=p.t_end, t_end=p.t_end + 10000.0,
                  step_size=p.step_size, print_step=p.print_step)
    M, N = q.num_ado(), p.N
    rng = np.random.default_rng(seed)
    rho = (rng.standard_normal((M, N, N)) + 1j * rng.standard_normal((M, N, N))) * (1.0 / (N * M))
    print("ready", flush=True)
    sys.stdin.readline()
    steps = 1
    while True:
        t0 = time.perf_counter()
        O.twod(q, nv=(1, 1, 1), rho_init=rho, max_steps=steps)
        elapsed = time.perf_counter() - t0
        if elapsed > seconds / 3 or steps >= 4096:
            break
        steps *= 2
    print(json.dumps({"steps": steps, "seconds": elapsed, "ados": M}), flush=True)
    return 0


def cpu_baseline(name, seconds=20.0, max_procs=16):
    """The oracle on this host's cores: one single-threaded process per core of
    this process's affinity set (at most max_procs), each on its own copy of the
    workload's first instance, as the reference farms one process per (tau, T)
    point (script/cpu-script/submit_jobs.sh:59-65).  All processes start together
    (each waits for a line on stdin); the rate is the sum of their rates."""
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count() or 1
    nproc = max(1, min(max_procs, ncpu))
    procs = [subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-worker", name,
                               "--cpu-seconds", str(seconds), "--cpu-seed", str(r)],
                              stdin=subprocess.PIPE, stdout=subprocess.PIPE, text=True) for r in range(nproc)]
    for pr in procs:
        line = pr.stdout.readline()
        if line.strip() != "ready":
            raise RuntimeError(f"CPU worker failed to start: {line!r}")
    t0 = time.perf_counter()
    for pr in procs:
        pr.stdin.write("go\n")
        pr.stdin.flush()
    res = []
    for pr in procs:
        out, _ = pr.communicate()
        if pr.returncode:
            raise RuntimeError(f"CPU worker exited with {pr.returncode}")
        res.append(json.loads(out.strip().splitlines()[-1]))
    wall = time.perf_counter() - t0
    rate = sum(r["ados"] * r["steps"] / r["seconds"] for r in res)
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    steps = sorted({r["steps"] for r in res})
    return {"value": rate, "unit": "ADO-updates/s", "cores": nproc, "kind": "port",
            "single_process_value": res[0]["ados"] * res[0]["steps"] / res[0]["seconds"],
            "cpu_model": model, "affinity_cpus": ncpu,
            "sample": f"{nproc} concurrent single-threaded oracle processes, each {'/'.join(map(str, steps))} RK4 "
                      f"step(s) of the workload's first instance ({res[0]['ados']} ADOs) from a dense random "
                      f"state after the last pulse; {wall:.1f} s wall"}


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
    """Forward + adjoint gradient of a 2DES-grid loss; reports the cost ratio
    (forward + backward) / forward and the equivalent central-FD cost.
    trimer_grad: the trimer's 16 x 16 (t1, t3) rephasing grid (48 propagations), one batch.
    chain_grad: 16 points of the 10-site chain's (t1, t2) grid (N = 56, L = 3, 3000 steps),
    gradient w.r.t. all site energies and couplings; the adjoint state of all 16 exceeds
    the device, so the backward runs in sub-batches (each re-runs its forward with
    checkpoints every 32 steps, then its backward)."""
    import numpy as np
    import torch
    import etoheom_amd as E
    from etoheom_amd import spectra as SP
    ctx = E.Context(0)
    print(f"bench: {a.workload} on {ctx.name}", file=sys.stderr, flush=True)
    if a.workload == "chain_grad":
        spec = SP.chain_grid_spec(ctx, t1s=(0.0, 10.0, 20.0, 30.0), t2s=(0.0, 100.0, 200.0, 400.0), nsteps=3000)
        mk = SP.chain_params
        sig = SP.chain_grid_signal
        target_params = SP.chain_params(perturb=7)
        M, grid_points = 1771, None
    else:
        spec = SP.grid_spec(ctx)
        mk = SP.trimer_params
        sig = SP.grid_signal
        target_params = SP.trimer_params(perturb=5)
        M, grid_points = 462, 16
    spec.progress = lambda m: print(f"bench: {m}", file=sys.stderr, flush=True)
    with torch.no_grad():
        target = sig(target_params, spec)
    print("bench: target grid done", file=sys.stderr, flush=True)
    import gc
    fw, bw, gfw, gbw, gpf, grf = [], [], [], [], [], []
    for it in range(a.warmup + a.steps):
        params = mk()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        S = sig(params, spec)
        loss = SP.grid_loss(S, target)
        t1 = time.perf_counter()
        loss.backward()
        t2 = time.perf_counter()
        if it >= a.warmup:
            fw.append(t1 - t0)
            bw.append(t2 - t1)
            gfw.append(spec.timing["forward_ms"] * 1e-3)
            gbw.append(spec.timing["backward_ms"] * 1e-3)
            grf.append(spec.timing.get("recompute_forward_ms", 0.0) * 1e-3)
            subs = spec.timing["sub_batches"]
        lval = float(loss.detach())
        del S, loss   # the graph holds the batch (and its trajectory): free it before the next forward
        gc.collect()
        if a.workload == "chain_grad":
            if it >= a.warmup:   # the gradient's forward was already the plain one (no adjoint state)
                gpf.append(gfw[-1])
        else:
            with torch.no_grad():   # one plain forward solve (no adjoint state) of the same grid
                sig(params, spec)
            if it >= a.warmup:
                gpf.append(spec.timing["forward_ms"] * 1e-3)
    n_params = sum(p.numel() for p in params.values())
    f, b = float(np.mean(fw)), float(np.mean(bw))
    gf, gb, gp, gr = float(np.mean(gfw)), float(np.mean(gbw)), float(np.mean(gpf)), float(np.mean(grf))
    steps = sum(int(np.floor(np.float32(np.float32(te) / np.float32(spec.step_size)))) + 1 for te in spec.t_end)
    ups = M * steps
    cfg = {"workload": workload(a.workload)[1], "device": ctx.name,
           "propagations_per_s": len(spec.nv) / gf,
           "adjoint": {"forward_gpu_s": gf, "backward_gpu_s": gb, "gradient_over_forward": (gf + gb) / gf,
                       "plain_forward_gpu_s": gp, "gradient_over_plain_forward": (gf + gb) / gp,
                       "backward_recompute_forward_gpu_s": gr, "sub_batches": subs,
                       "forward_wall_s": f, "backward_wall_s": b, "gradient_over_forward_wall": (f + b) / f,
                       "n_params": n_params, "central_fd_s": 2 * n_params * gf,
                       "fd_over_adjoint": 2 * n_params * gf / (gf + gb), "loss": lval,
                       "note": "GPU: HIP events around the propagation and the backward pass; "
                               "forward = the gradient run's forward (with one batch it also keeps the "
                               "adjoint's state; in sub-batches it is a plain forward and the backward "
                               "re-runs each sub-batch's forward with checkpoints), plain forward = the "
                               "same grid with no gradient wanted; wall: incl. host set-up of each batch "
                               "(tables, graph capture, allocation)"}}
    if grid_points:
        cfg["grid_points_per_s"] = grid_points / gf
    rec = {"metric": "FP64 ADO-updates/s + 2DES grid points/s at 1e-10 err; adjoint-grad time vs FD",
           "value": ups / gf, "unit": "ADO-updates/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": (f + b) * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic", "config": cfg}
    print(json.dumps(rec))
    return 0


def complex_v_share(items, first, last):
    """Share of (instance, RK4 step) pairs in steps [first, last) whose V(t) is complex
    (some pulse on: the commutator then needs twice the MFMA work), from the same
    float time grid and window rule the library uses."""
    import numpy as np
    on = tot = 0
    for p in items:
        f = np.float32
        n = int(np.floor(float(f(f(p.t_end) - f(p.t_start)) / f(p.step_size)))) + 1
        t = f(p.t_start)
        for i in range(min(last, n)):
            if i >= first:
                tot += 1
                if any(abs(float(t) - pl[1]) <= 2 * pl[2] for pl in p.pulses):
                    on += 1
            t = f(t + f(p.step_size))
    return on / max(tot, 1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="chain_grid")
    ap.add_argument("--chunk", type=int, default=0,
                    help="RK4 steps per bench step (0: 50 for chain_grid, else a twentieth of the propagation)")
    ap.add_argument("--full", action="store_true",
                    help="time one whole propagation of the batch (from a reset) instead of --steps bench steps")
    ap.add_argument("--profile-only", action="store_true",
                    help="only the per-launch timing pass (--profile-steps eager steps from a reset); for "
                         "rocprofv3 runs whose kernel averages must match the roofline figure")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint-leg", action="store_true",
                    help="skip the trimer grid gradient the default workload reports beside its value")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--cpu-worker", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--cpu-seed", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--nsteps", type=int, default=0, help="shorten the propagation (counter runs only)")
    ap.add_argument("--autotune", type=int, default=0, help="autotuner steps per launch candidate (0: off)")
    ap.add_argument("--per-propagation", type=int, default=0,
                    help="also time this many of the workload's propagations one at a time on the plain "
                         "per-launch path (no graph, no batching)")
    ap.add_argument("--copy-calibrate", action="store_true",
                    help="also run a plain 2 GiB read + 2 GiB write elementwise copy (counter scale)")
    a = ap.parse_args()
    if a.cpu_worker:
        return cpu_worker(a.cpu_worker, a.cpu_seconds, a.cpu_seed)
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
    if a.workload in ("trimer_grad", "chain_grad"):
        return bench_gradient(a)
    items, desc = workload(a.workload, a.nsteps)
    print(f"bench: {desc}", file=sys.stderr, flush=True)
    ctx = E.Context(0)
    t_setup = time.perf_counter()
    batch = E.Batch(ctx, [{"problem": p, "H": p.H, "X": p.X, "nv": (1, 1, 1)} for p in items])
    max_steps = batch.max_steps
    S = a.chunk or (min(50, max_steps) if a.workload == "chain_grid" else -(-max_steps // 20))
    if a.full:
        S = -(-max_steps // math.ceil(max_steps / S))   # equal replays covering the whole propagation
    batch.set_chunk(S)
    t_setup = time.perf_counter() - t_setup
    tuning = batch.autotune(a.autotune).splitlines() if a.autotune else None
    if a.profile_only:
        prof = batch.profile(a.profile_steps)
        elems = sum(p.num_ado() * p.N * p.N for p in items)
        stage_ms = float(np.mean(prof[1:5]))
        print(json.dumps({"workload": desc, "profile_steps": a.profile_steps, "prologue_ms": float(prof[0]),
                          "stage_ms": [float(x) for x in prof[1:5]], "bytes_per_launch": 64.0 * elems,
                          "achieved_gbs": 64.0 * elems / (stage_ms * 1e-3) / 1e9, "copy_gbs": copy_gbs,
                          "groups": batch.tuning_report().splitlines()}))
        batch.close()
        return 0
    per_inst = [(p.num_ado(), E._nsteps(p.system()[0])) for p in items]

    pos = [max_steps]

    def bench_step():
        """The next S RK4 steps of every instance (a reset first when the propagation is done)."""
        if pos[0] >= max_steps:
            batch.start()
            pos[0] = 0
        batch.advance(S)
        u = sum(M * max(0, min(pos[0] + S, n) - pos[0]) for M, n in per_inst)
        pos[0] += S
        return u

    steps = math.ceil(max_steps / S) if a.full else a.steps
    warm = 0 if a.full else a.warmup
    for _ in range(warm):
        bench_step()
        print(f"bench: warm-up step done (RK4 steps to {pos[0]})", file=sys.stderr, flush=True)
    if a.full or warm == 0:
        pos[0] = max_steps   # the timed region starts from a reset
    first = pos[0] % max_steps if pos[0] < max_steps else 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ups = 0
    gpu_ms = []   # HIP events around each timed step's graph replays, on the batch stream
    for k in range(steps):
        ups += bench_step()
        torch.cuda.synchronize()
        gpu_ms.append(batch.elapsed_ms())
        if k % 5 == 4:
            print(f"bench: timed step {k + 1}/{steps} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    last = pos[0]
    done = last >= max_steps
    outs = [batch.output(i) for i in range(len(items))]
    finite = all(np.all(np.isfinite(o)) for o in outs)
    timed_range = (f"RK4 steps {first}..{last - 1} of {max_steps}" if steps * S <= max_steps
                   else f"{steps * S} RK4 steps from step {first}, wrapping round (reset) at {max_steps}")

    # per-launch durations of the stage kernel (HIP events on the batch stream)
    prof = batch.profile(a.profile_steps)
    stage_ms = float(np.mean(prof[1:5]))
    elems = sum(p.num_ado() * p.N * p.N for p in items)
    # the algorithmic model: 256 N^2 B per ADO-update (9 reads + 7 writes of the N x N
    # block per RK4 step) over 4 stage launches; the kernels' own schedule needs 7 + 5
    # (192 N^2 B), reported beside it
    bytes_per_launch = 64.0 * elems
    schedule_bytes = 48.0 * elems
    real_v_achieved = bytes_per_launch / (stage_ms * 1e-3) / 1e9
    # the timed region: the stage kernel's mean launch duration from the HIP events around
    # the timed replays (batches of one N class, whose four stage launches per step run one
    # after another in the graph); a batch of several N classes runs them on parallel
    # branches, and its per-launch figure is the eager pass's
    n_groups = len({c[0] for c in batch.candidates()})
    graph_launch_ms = sum(gpu_ms) / (steps * S * 4) if n_groups == 1 else None
    launch_ms = graph_launch_ms if graph_launch_ms else stage_ms
    achieved = bytes_per_launch / (launch_ms * 1e-3) / 1e9

    per_prop = per_propagation(ctx, items, a.per_propagation) if a.per_propagation else None

    traffic, tsrc = traffic_of(a.workload)
    value = ups / wall
    full_ups = sum(M * n for M, n in per_inst)
    rec = {
        "metric": "FP64 ADO-updates/s + 2DES grid points/s at 1e-10 err; adjoint-grad time vs FD",
        "value": value,
        "unit": "ADO-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "ms_per_step": wall / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": desc, "propagations": len(items),
                   "bench_step": f"{S} RK4 steps of every propagation (one replay of the forward graph)",
                   "timed": timed_range + (" (whole propagation)" if first == 0 and done and steps * S == max_steps
                                           else ""),
                   "complex_v_share_timed": complex_v_share(items, first, min(last, max_steps)),
                   "complex_v_share_whole": complex_v_share(items, 0, max_steps),
                   "propagations_per_s": value / (full_ups / len(items)),
                   **({"grid_points_per_s": value / (full_ups / len(items)),
                       "grid_points_note": "the workload's grid has one propagation (the p111 ordering) per "
                                           "(t1, t2) point; the adjoint leg's trimer grid runs p111/p110/p101 "
                                           "triplets and counts a point per triplet"}
                      if a.workload.startswith("chain_grid") else {}),
                   "gpu_event_ms_timed": float(sum(gpu_ms)), "device": ctx.name, "finite": finite,
                   "readouts_complete": done, "groups": batch.tuning_report().splitlines(),
                   "setup_s": t_setup,
                   "setup_note": "host hierarchy/table build, upload, allocation and graph capture; outside value"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": tsrc,
                     "kernel": "stage (fused RHS + RK4 stage)",
                     "launch_ms": launch_ms,
                     "launch_ms_note": ("mean stage launch over the timed replays: HIP events around each timed "
                                        "step on the batch stream / (RK4 steps x 4)" if graph_launch_ms else
                                        "mean of the eager pass's stage launches (several N classes)"),
                     "stage_ms": [float(x) for x in prof[1:5]],
                     "stage_ms_note": f"HIP events around each stage launch over {a.profile_steps} eager steps "
                                      "from a reset (before the first pulse: real V)",
                     "real_v_achieved": real_v_achieved, "real_v_frac": real_v_achieved / HBM_PEAK_GBS,
                     "prologue_ms": float(prof[0]), "bytes_per_launch": bytes_per_launch,
                     "schedule_bytes_per_launch": schedule_bytes,
                     "schedule_achieved": schedule_bytes / (launch_ms * 1e-3) / 1e9,
                     # the counters' bytes per launch (L2 misses: HBM or the last-level cache,
                     # measured on the real-V launches) over the timed launch time
                     "traffic_achieved": traffic / (launch_ms * 1e-3) / 1e9 if traffic else None},
    }
    if copy_gbs:
        rec["roofline"]["copy_gbs"] = copy_gbs
    if tuning:
        rec["config"]["autotune_steps"] = a.autotune
    if per_prop:
        per_prop["batched_over_per_propagation"] = rec["config"]["propagations_per_s"] / per_prop["propagations_per_s"]
        rec["config"]["per_propagation_path"] = per_prop
    batch.close()
    if a.workload == "chain_grid" and not a.no_adjoint_leg:
        rec["adjoint_leg"] = adjoint_leg()
    if not a.no_cpu:
        rec["cpu_baseline"] = cpu_baseline(a.workload, a.cpu_seconds)
    print(json.dumps(rec))
    return 0


def adjoint_leg(timeout=600):
    """The metric's adjoint leg: the trimer grid gradient (16 x 16 (t1, t3) rephasing grid,
    48 propagations, L = 5, checkpoints every 32 steps) timed by this script's trimer_grad
    workload in a child process (2 warm-up and 3 timed gradients); its progress goes to
    stderr as it runs."""
    cmd = [sys.executable, os.path.abspath(__file__), "--workload", "trimer_grad", "--warmup", "2", "--steps", "3",
           "--no-cpu"]
    print("bench: adjoint leg (trimer grid gradient) in a child process", file=sys.stderr, flush=True)
    try:
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=None, text=True, timeout=timeout)
    except subprocess.TimeoutExpired:
        return {"error": f"trimer_grad did not finish in {timeout} s"}
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("{")]
    if r.returncode != 0 or not lines:
        return {"error": f"trimer_grad exited with {r.returncode}"}
    d = json.loads(lines[-1])
    leg = {"workload": d["config"]["workload"], "ado_updates_per_s": d["value"]}
    leg.update(d["config"]["adjoint"])
    leg["grid_points_per_s"] = d["config"].get("grid_points_per_s")
    return leg


if __name__ == "__main__":
    sys.exit(main())
