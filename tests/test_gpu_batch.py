"""GPU batch engine: the per-group graph, the plain per-launch path, the launch
candidates and the autotuner, each against the CPU oracle (<= 1e-10 normwise
relative, FP64) on one batch that holds five N classes (N = 4, 7, 11, 29, 37) and
an instance with three modes per bath."""
import os

import numpy as np
import pytest

import etoheom_amd as E
import oracle as O
from etoheom_amd import configs

pytestmark = pytest.mark.gpu
TOL = 1e-10


@pytest.fixture(scope="module")
def ctx():
    return E.Context(0)


def oproblem(p):
    return O.Problem(H=p.H, X=p.X, alpha=p.alpha, alpha_t=p.alpha_t, gamma=p.gamma, L=p.L,
                     site_offset=p.site_offset, pulses=p.pulses, t_start=p.t_start, t_end=p.t_end,
                     step_size=p.step_size, print_step=p.print_step)


def rel(a, b):
    b = np.asarray(b)
    assert b.size and np.max(np.abs(b)) > 0, "reference is identically zero: the check would be vacuous"
    return np.max(np.abs(np.asarray(a) - b)) / np.max(np.abs(b))


def mixed_items():
    items = [configs.aggregate(2, 3, nsteps=120, taus=(10.0, 20.0, 30.0)),      # N = 4
             configs.aggregate(3, 3, nsteps=100, taus=(10.0, 25.0, 35.0)),      # N = 7
             configs.aggregate(4, 2, nsteps=90, taus=(20.0, 10.0, 30.0)),       # N = 11
             configs.heptamer(L=1, nsteps=40),                                  # N = 29
             configs.aggregate(8, 1, nsteps=30, taus=(2.0, 6.0, 10.0))]         # N = 37
    items[3].pulses[:, 1] = [2.0, 6.0, 10.0]
    # three modes per bath (Km = 3): the modes beyond the kernels' two-phase gathers
    km3 = configs.aggregate(2, 2, nsteps=110, taus=(10.0, 20.0, 30.0))
    km3.alpha = np.concatenate([km3.alpha, 0.5 * km3.alpha[:, :1]], axis=1)
    km3.alpha_t = np.concatenate([km3.alpha_t, 0.5 * km3.alpha_t[:, :1]], axis=1)
    km3.gamma = np.concatenate([km3.gamma, 2.0 * km3.gamma[:, :1]], axis=1)
    items.append(km3)
    for p in items:
        p.pulses[:, 2] = 4.0
        p.pulses[:, 0] *= 1e5          # strong field: every ADO carries signal
    return items


@pytest.fixture(scope="module")
def refs():
    items = mixed_items()
    return items, [O.twod(oproblem(p), nv=(1, 1, 1), want_rho=True) for p in items]


def check(b, items, ref):
    for i, p in enumerate(items):
        Pr, rr = ref[i]
        assert np.max(np.abs(rr[1:])) > 1e-12 * np.max(np.abs(rr[0])), "hierarchy not driven"
        assert rel(b.output(i), Pr) < TOL, i
        assert rel(b.rho(i), rr) < TOL, i


def test_grouped_graph_and_per_launch_path_agree(ctx, refs):
    items, ref = refs
    b = E.Batch(ctx, [{"problem": p, "H": p.H, "X": p.X, "nv": (1, 1, 1)} for p in items])
    b.run()
    b.sync()
    check(b, items, ref)
    g_out = [b.output(i) for i in range(len(items))]
    g_rho = [b.rho(i) for i in range(len(items))]
    b.run_eager()
    b.sync()
    for i in range(len(items)):   # same kernels, same per-instance arithmetic: bitwise equal
        assert np.array_equal(b.output(i), g_out[i])
        assert np.array_equal(b.rho(i), g_rho[i])


def test_every_launch_candidate_matches_oracle(ctx, refs):
    items, ref = refs
    b = E.Batch(ctx, [{"problem": p, "H": p.H, "X": p.X, "nv": (1, 1, 1)} for p in items])
    cands = b.candidates()
    kinds = {k for _, _, k, _ in cands}
    # the N = 29 item also admits the pipelined and two-phase kernels (K_m = 2, 2 K N <= 512)
    mfma = ({"stage_mfma32", "stage_mfma32_tp", "stage_mfma_pipe", "stage_mfma64", "stage_sparse",
             "stage_sparse_rows"}
            if os.environ.get("ETOHEOM_NO_MFMA") != "1" else set())
    assert {"stage_small", "stage_generic"} | mfma <= kinds
    assert {t for _, _, k, t in cands if k == "stage_small"} == {256, 512, 1024}
    # every small-N group also offers its base kernel with the other state-store hint
    rep = b.tuning_report().splitlines()
    for g in {g for g, _, k, _ in cands if k == "stage_small"}:
        assert {l.split("stream=")[1][0] for l in rep if int(l.split()[1]) == g} == {"0", "1"}
        # and with the state operand loaded with the input or after the tile's barrier
        assert {l.split("late=")[1][0] for l in rep if int(l.split()[1]) == g} == {"0", "1"}
    for g, c, kern, tile in cands:
        print(f"group {g} candidate {c}: {kern} tile={tile}", flush=True)
        b.select(g, c)
        b.run()
        b.sync()
        check(b, items, ref)
    with pytest.raises(E.EtoHeomError):
        b.select(len({g for g, _, _, _ in cands}), 0)


def test_autotune_picks_one_candidate_per_group_and_stays_correct(ctx, refs):
    items, ref = refs
    b = E.Batch(ctx, [{"problem": p, "H": p.H, "X": p.X, "nv": (1, 1, 1)} for p in items])
    rep = b.autotune(2)
    print(rep, flush=True)
    lines = rep.splitlines()
    groups = {int(l.split()[1]) for l in lines}
    assert len(groups) == 5
    for g in groups:
        sel = [l for l in lines if int(l.split()[1]) == g and l.endswith("*")]
        assert len(sel) == 1
        ms = [float(l.split("ms_per_step=")[1].split()[0]) for l in lines if int(l.split()[1]) == g]
        assert all(m > 0 for m in ms)
    b.run()
    b.sync()
    check(b, items, ref)


def test_store_hint_and_load_placement_variants_are_bitwise_equal(ctx, refs):
    """The small-N base kernel with the non-temporal store hint and/or the state operand
    loaded after the barrier runs the same arithmetic: every readout and final state is
    bitwise that of the plain variant (and within 1e-10 of the oracle)."""
    items, ref = refs
    b = E.Batch(ctx, [{"problem": p, "H": p.H, "X": p.X, "nv": (1, 1, 1)} for p in items])
    rep = b.tuning_report().splitlines()
    small = sorted({int(l.split()[1]) for l in rep if " stage_small " in l and "tile=512" in l})
    assert small

    def pick(g, stream, late):
        for l in rep:
            f = l.split()
            if (int(f[1]) == g and " stage_small " in l and "tile=512" in l and f"stream={stream}" in l
                    and f"late={late}" in l):
                return int(f[3])
        return None

    results = {}
    for stream, late in ((0, 0), (1, 0), (0, 1), (1, 1)):
        chosen = [(g, pick(g, stream, late)) for g in small]
        if any(c is None for _, c in chosen):
            continue
        for g, c in chosen:
            b.select(g, c)
        b.run()
        b.sync()
        results[(stream, late)] = ([b.output(i) for i in range(len(items))], [b.rho(i) for i in range(len(items))])
    assert (0, 0) in results and len(results) >= 3, sorted(results)
    check(b, items, ref)
    base_out, base_rho = results[(0, 0)]
    for key, (outs, rhos) in results.items():
        for i in range(len(items)):
            assert np.array_equal(outs[i], base_out[i]), (key, i)
            assert np.array_equal(rhos[i], base_rho[i]), (key, i)


def test_propagation_in_pieces_equals_one_run(ctx, refs):
    """eh_batch_start + eh_batch_advance over every step (another graph chunk) gives the
    bits of eh_batch_run, and a second start resets the states."""
    items, ref = refs
    b = E.Batch(ctx, [{"problem": p, "H": p.H, "X": p.X, "nv": (1, 1, 1)} for p in items])
    b.run()
    b.sync()
    out0 = [b.output(i) for i in range(len(items))]
    rho0 = [b.rho(i) for i in range(len(items))]
    b.set_chunk(7)
    with pytest.raises(E.EtoHeomError):
        b.advance(7)           # not started
    for rep in range(2):
        b.start()
        with pytest.raises(E.EtoHeomError):
            b.advance(5)       # not a multiple of the chunk
        done = 0
        while done < b.max_steps:   # the last piece runs past the longest instance's end
            b.advance(14)
            done += 14
        b.sync()
        for i in range(len(items)):
            assert np.array_equal(b.output(i), out0[i]), (rep, i)
            assert np.array_equal(b.rho(i), rho0[i]), (rep, i)
    check(b, items, ref)
    b.close()


def test_reads_before_a_run_and_misshaped_gradients_are_refused(ctx, refs):
    items, _ = refs
    b = E.Batch(ctx, [{"problem": p, "H": p.H, "X": p.X, "nv": (1, 1, 1)} for p in items[:2]], adjoint=16)
    with pytest.raises(E.EtoHeomError) as ei:
        b.output(0)
    assert ei.value.code == -1
    with pytest.raises(E.EtoHeomError):
        b.rho(0)
    b.run()
    b.sync()
    n0, n1 = len(b.output(0)), len(b.output(1))
    with pytest.raises(ValueError):
        b.backward([np.ones(n0)])                        # one instance missing
    with pytest.raises(ValueError):
        b.backward([np.ones(n0), np.ones(n1 - 1)])       # one readout short
    # the C ABI checks the length itself
    flat = np.ones(n0 + n1 - 1, dtype=np.complex128)
    assert E.lib().eh_batch_backward(b._h, flat.ctypes.data, flat.size) == -3
    b.backward([np.ones(n0), np.ones(n1)])
    b.close()
