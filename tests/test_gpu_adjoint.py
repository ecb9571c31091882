"""GPU adjoint gradients against central finite differences of the CPU oracle.

Loss: L = Re sum_t conj(w_t) P_t for fixed random weights w, so dL/dP = w in the
convention g = dL/dRe P + i dL/dIm P; each checked parameter theta is perturbed
by +-h in the oracle (relative steps in REL_STEP, Richardson-extrapolated with
h/2) and dL/dtheta from the adjoint must agree to <= 1e-6 relative (|FD| scale).
"""
import numpy as np
import pytest

import etoheom_amd as E
import oracle as O
from etoheom_amd import configs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return E.Context(0)


def oproblem(p, **over):
    kw = dict(H=p.H, X=p.X, alpha=p.alpha, alpha_t=p.alpha_t, gamma=p.gamma, L=p.L, site_offset=p.site_offset,
              pulses=p.pulses, t_start=p.t_start, t_end=p.t_end, step_size=p.step_size, print_step=p.print_step)
    kw.update(over)
    return O.Problem(**kw)


def loss(q, w):
    P, _ = O.twod(q, nv=(1, 1, 1))
    return float(np.real(np.sum(np.conj(w) * P)))


# relative FD steps: 1e-4 for H, whose entries multiply t in the phase (a plain
# central difference then errs by ~(h t)^2 ~ 1e-5, hence the Richardson step);
# 1e-3 for X and gamma; 1e-1 for alpha, alpha~ (|alpha| ~ 3e-4, so smaller steps
# leave the difference dominated by rounding, eps |L| / h ~ 1e-6 of the derivative,
# while the Richardson value is already converged to ~1e-8 at this step)
REL_STEP = {"H": 1e-4, "X": 1e-3, "alpha": 1e-1, "alpha_t": 1e-1, "gamma": 1e-3}


def fd(p, w, name, idx, part, rel=None):
    """Central FD of the oracle loss wrt Re/Im of one entry of array `name`,
    Richardson-extrapolated from steps h and h/2 (relative step REL_STEP[name] or `rel`)."""
    base = getattr(p, name).astype(np.complex128)
    v = base[idx]
    mag = abs(v.real if part == "re" else v.imag)
    # relative to the entry, or to the array's scale for (near-)zero entries
    h0 = (rel or REL_STEP[name]) * max(mag, 0.5 * np.max(np.abs(base)))

    def central(h):
        step = h if part == "re" else 1j * h
        a, b = base.copy(), base.copy()
        a[idx] += step
        b[idx] -= step
        return (loss(oproblem(p, **{name: a}), w) - loss(oproblem(p, **{name: b}), w)) / (2 * h)

    return (4 * central(h0 / 2) - central(h0)) / 3


def adj_value(g, name, idx, part):
    v = g[name][idx]
    return v.real if part == "re" else v.imag


CHECKS = [("H", (1, 1), "re"), ("H", (1, 2), "re"), ("H", (3, 3), "re"), ("X", (0, 0, 1), "re"),
          ("X", (1, 0, 2), "re"), ("X", (2, 1, 3), "re"), ("X", (3, 1, 0), "re"), ("gamma", (0, 0), "re"),
          ("gamma", (1, 1), "im"), ("alpha", (0, 1), "re"), ("alpha", (1, 0), "im"), ("alpha_t", (0, 0), "im")]


def small_dimer():
    p = configs.dimer(L=2, nsteps=160, t1=10.0, t2=10.0)
    p.pulses[:, 1] = [20.0, 30.0, 40.0]
    p.pulses[:, 2] = 8.0
    p.pulses[:, 0] *= 3e5        # a strong field so that every term matters
    p.gamma = p.gamma * 4.0
    p.alpha = p.alpha * 8.0
    p.alpha_t = p.alpha_t * 8.0
    return p


@pytest.mark.parametrize("C,mode", [(160, "trajectory"), (7, "trajectory"), (7, "replay")])
def test_adjoint_matches_oracle_fd_dimer(ctx, C, mode, monkeypatch):
    """Both backward modes: every stage input kept from the forward (trajectory), or
    checkpoints every C steps with a replay per segment."""
    if mode == "replay":
        monkeypatch.setenv("ETOHEOM_ADJ_REPLAY", "1")
    p = small_dimer()
    P0, _ = O.twod(oproblem(p), nv=(1, 1, 1))
    rng = np.random.default_rng(1)
    w = rng.normal(size=len(P0)) + 1j * rng.normal(size=len(P0))
    b = E.Batch(ctx, [{"problem": p, "H": p.H, "X": p.X, "nv": (1, 1, 1)}], adjoint=C)
    b.run()
    b.sync()
    P = b.output(0)
    assert np.max(np.abs(P - P0)) / np.max(np.abs(P0)) < 1e-10
    b.backward([w])
    g = b.grad(0)
    for name, idx, part in CHECKS:
        ref = fd(p, w, name, idx, part)
        got = adj_value(g, name, idx, part)
        scale = max(abs(ref), 1e-3 * np.max(np.abs(P0)) * np.max(np.abs(w)))
        assert abs(got - ref) / scale < 1e-6, (name, idx, part, got, ref)


def test_adjoint_with_streamed_store_candidate(ctx):
    """The small-N candidate with non-temporal state stores changes cache hints only: the
    forward readouts are bitwise those of the default candidate, and the gradients agree
    to rounding (their reductions add in LDS with atomics, in no fixed order)."""
    p = small_dimer()
    P0, _ = O.twod(oproblem(p), nv=(1, 1, 1))
    rng = np.random.default_rng(5)
    w = rng.normal(size=len(P0)) + 1j * rng.normal(size=len(P0))
    res = []
    for streamed in ("stream=0", "stream=1"):
        b = E.Batch(ctx, [{"problem": p, "H": p.H, "X": p.X, "nv": (1, 1, 1)}])
        line = [l for l in b.tuning_report().splitlines()
                if "stage_small" in l and "tile=512" in l and streamed in l][0]
        b.select(int(line.split()[1]), int(line.split()[3]))
        b.enable_adjoint(16)
        b.run()
        b.sync()
        b.backward([w])
        res.append((b.output(0), b.grad(0)))
    assert np.max(np.abs(res[0][0] - P0)) / np.max(np.abs(P0)) < 1e-10
    assert np.array_equal(res[0][0], res[1][0])
    for name in res[0][1]:
        a, b2 = res[0][1][name], res[1][1][name]
        assert np.max(np.abs(a - b2)) <= 1e-12 * np.max(np.abs(a)), name


def test_adjoint_batch_of_two_sizes(ctx):
    """Per-instance offsets: a dimer and a trimer in one batch, gradients of each vs FD."""
    p1 = small_dimer()
    p2 = configs.aggregate(3, 2, nsteps=120, taus=(15.0, 25.0, 35.0))
    p2.pulses[:, 2] = 8.0
    p2.pulses[:, 0] *= 3e5
    items = [p1, p2]
    ws = []
    rng = np.random.default_rng(2)
    for p in items:
        P0, _ = O.twod(oproblem(p), nv=(1, 1, 1))
        ws.append(rng.normal(size=len(P0)) + 1j * rng.normal(size=len(P0)))
    b = E.Batch(ctx, [{"problem": p, "H": p.H, "X": p.X, "nv": (1, 1, 1)} for p in items], adjoint=16)
    b.run()
    b.sync()
    b.backward(ws)
    for i, p in enumerate(items):
        g = b.grad(i)
        for name, idx, part in [("H", (2, 2), "re"), ("X", (0, 0, 2), "re"), ("gamma", (1, 0), "re"),
                                ("alpha", (0, 0), "re")]:
            ref = fd(p, ws[i], name, idx, part)
            got = adj_value(g, name, idx, part)
            assert abs(got - ref) / max(abs(ref), 1e-12) < 1e-6, (i, name, idx, got, ref)


def test_adjoint_generic_large_n(ctx):
    """N = 29 goes through the generic adjoint kernel (MFMA forward)."""
    p = configs.heptamer(L=1, nsteps=40)
    p.pulses[:, 1] = [2.0, 6.0, 10.0]
    p.pulses[:, 2] = 3.0
    p.pulses[:, 0] *= 3e5
    P0, _ = O.twod(oproblem(p), nv=(1, 1, 1))
    rng = np.random.default_rng(3)
    w = rng.normal(size=len(P0)) + 1j * rng.normal(size=len(P0))
    b = E.Batch(ctx, [{"problem": p, "H": p.H, "X": p.X, "nv": (1, 1, 1)}], adjoint=8)
    b.run()
    b.sync()
    b.backward([w])
    g = b.grad(0)
    for name, idx, part in [("H", (1, 1), "re"), ("H", (2, 9), "re"), ("X", (0, 0, 3), "re"),
                            ("gamma", (2, 1), "re"), ("alpha", (4, 0), "re")]:
        ref = fd(p, w, name, idx, part)
        got = adj_value(g, name, idx, part)
        assert abs(got - ref) / max(abs(ref), 1e-12) < 1e-6, (name, idx, got, ref)


@pytest.mark.parametrize("N,nsteps", [(67, 24), (100, 24)])
def test_adjoint_above_n64(ctx, N, nsteps):
    """A random Hermitian system with three baths of two modes, L = 1, above N = 64: the group
    runs the wide forward kernel and the generic adjoint kernel. At N = 67 its LDS holds
    lambda and rho (2 N^2 + 2 K N complex: 147 KB); at N = 100 (320 KB) it reads them in
    place."""
    base = configs.chain(L=1, nsteps=nsteps)
    rng = np.random.default_rng(N)
    A = rng.normal(size=(N, N)) * 0.02
    H = np.diag(rng.uniform(2.2, 2.45, N)) + (A + A.T) / 2
    H[0, :] = 0.0
    H[:, 0] = 0.0
    X = np.zeros((4, N, N), complex)
    for q in range(4):
        v = rng.normal(size=N - 1)
        X[q, 0, 1:] = v
        X[q, 1:, 0] = v
    pul = base.pulses.copy()
    pul[:, 1] = [2.0, 5.0, 8.0]
    pul[:, 2] = 3.0
    pul[:, 0] *= 3e5
    p = E.Problem(H=H.astype(complex), X=X, alpha=base.alpha[:3], alpha_t=base.alpha_t[:3], gamma=base.gamma[:3],
                  L=1, site_offset=base.site_offset, pulses=pul, t_start=base.t_start, t_end=base.t_end,
                  step_size=base.step_size, print_step=base.print_step)
    P0, _ = O.twod(oproblem(p), nv=(1, 1, 1))
    assert len(P0) > 0
    rng = np.random.default_rng(5)
    w = rng.normal(size=len(P0)) + 1j * rng.normal(size=len(P0))
    b = E.Batch(ctx, [{"problem": p, "H": p.H, "X": p.X, "nv": (1, 1, 1)}], adjoint=8)
    assert "stage_mfma_wide" in {k for _, _, k, _ in b.candidates()}
    b.run()
    b.sync()
    assert np.max(np.abs(b.output(0) - P0)) / np.max(np.abs(P0)) < 1e-10
    b.backward([w])
    g = b.grad(0)
    # gamma's step 3e-3: at N = 100 the oracle's rounding moves its 1e-3 difference by ~1e-6
    # of the derivative (steps 3e-3 and 1e-2 agree to 4e-8)
    for name, idx, part in [("H", (1, 1), "re"), ("H", (2, 40), "re"), ("X", (0, 0, 5), "re"),
                            ("gamma", (1, 1), "re"), ("alpha", (2, 0), "re")]:
        ref = fd(p, w, name, idx, part, rel=3e-3 if name == "gamma" else None)
        got = adj_value(g, name, idx, part)
        assert abs(got - ref) / max(abs(ref), 1e-12) < 1e-6, (name, idx, got, ref)
    b.close()


def test_adjoint_with_pipelined_forward_candidate(ctx):
    """The persistent pipelined N <= 32 kernel has no fused step tail: with it selected, the
    adjoint's forward runs unfused steps (a prologue per step) and the gradient agrees with
    the default candidate's to rounding."""
    p = configs.heptamer(L=1, nsteps=40)
    p.pulses[:, 1] = [2.0, 6.0, 10.0]
    p.pulses[:, 2] = 3.0
    p.pulses[:, 0] *= 3e5
    P0, _ = O.twod(oproblem(p), nv=(1, 1, 1))
    rng = np.random.default_rng(4)
    w = rng.normal(size=len(P0)) + 1j * rng.normal(size=len(P0))
    res = []
    for kern in (None, "stage_mfma_pipe"):
        b = E.Batch(ctx, [{"problem": p, "H": p.H, "X": p.X, "nv": (1, 1, 1)}])
        if kern:
            c = [c for g, c, k, _ in b.candidates() if k == kern]
            if not c:
                pytest.skip("no pipelined candidate (MFMA kernels switched off)")
            b.select(0, c[0])
        b.enable_adjoint(8)
        b.run()
        b.sync()
        assert np.max(np.abs(b.output(0) - P0)) / np.max(np.abs(P0)) < 1e-10, kern
        b.backward([w])
        res.append(b.grad(0))
    for name in res[0]:
        a, c = res[0][name], res[1][name]
        assert np.max(np.abs(a - c)) <= 1e-9 * np.max(np.abs(a)), name


def test_adjoint_chain_n56(ctx):
    """N = 56, the 10-site chain's system (32 < N <= 64: the stage_mfma64 forward and the
    adj_mfma64 adjoint): site energies, a nearest-neighbour coupling, a double-exciton
    entry, a dipole entry and bath-mode parameters (the bath reductions) against FD."""
    p = configs.chain(L=1, nsteps=24)
    p.pulses[:, 1] = [2.0, 5.0, 8.0]
    p.pulses[:, 2] = 3.0
    p.pulses[:, 0] *= 3e5
    assert p.N == 56
    P0, _ = O.twod(oproblem(p), nv=(1, 1, 1))
    rng = np.random.default_rng(4)
    w = rng.normal(size=len(P0)) + 1j * rng.normal(size=len(P0))
    b = E.Batch(ctx, [{"problem": p, "H": p.H, "X": p.X, "nv": (1, 1, 1)}], adjoint=8)
    b.run()
    b.sync()
    assert np.max(np.abs(b.output(0) - P0)) / np.max(np.abs(P0)) < 1e-10
    b.backward([w])
    g = b.grad(0)
    for name, idx, part in [("H", (1, 1), "re"), ("H", (6, 6), "re"), ("H", (1, 2), "re"), ("H", (11, 12), "re"),
                            ("X", (0, 0, 4), "re"), ("gamma", (0, 0), "re"), ("gamma", (3, 1), "im"),
                            ("alpha", (2, 0), "re"), ("alpha_t", (5, 1), "im")]:
        ref = fd(p, w, name, idx, part)
        got = adj_value(g, name, idx, part)
        assert abs(got - ref) / max(abs(ref), 1e-12) < 1e-6, (name, idx, got, ref)


def test_adjoint_chain_mfma_equals_generic(ctx, monkeypatch):
    """The N = 56 chain at L = 2 (231 ADOs) through two pulse windows: every gradient of
    the MFMA adjoint kernel equals the VALU adjoint kernel's (selected by turning the MFMA
    forward off) to 1e-10 of its scale (the two sum in different orders)."""
    p = configs.chain(L=2, nsteps=40)
    p.pulses[:, 1] = [2.0, 5.0, 8.0]
    p.pulses[:, 2] = 3.0
    p.pulses[:, 0] *= 3e5
    rng = np.random.default_rng(5)
    out = []
    for no_mfma in ("0", "1"):
        monkeypatch.setenv("ETOHEOM_NO_MFMA", no_mfma)
        b = E.Batch(ctx, [{"problem": p, "H": p.H, "X": p.X, "nv": (1, 1, 1)}], adjoint=8)
        kinds = {k for _, _, k, _ in b.candidates()}
        assert ("stage_mfma64" in kinds) == (no_mfma == "0")
        b.run()
        b.sync()
        P = b.output(0)
        w = rng.normal(size=len(P)) + 1j * rng.normal(size=len(P)) if not out else out[0][2]
        b.backward([w])
        out.append((b.grad(0), P, w))
        b.close()
    assert np.max(np.abs(out[0][1] - out[1][1])) <= 1e-10 * np.max(np.abs(out[1][1]))
    for name in ("H", "X", "alpha", "alpha_t", "gamma"):
        a, c = out[0][0][name], out[1][0][name]
        assert np.max(np.abs(c)) > 0, name
        assert np.max(np.abs(a - c)) <= 1e-10 * np.max(np.abs(c)), (name, np.max(np.abs(a - c)) / np.max(np.abs(c)))


def test_adjoint_replay_mode_large_n(ctx, monkeypatch):
    """Checkpoints + replay at N = 29: the replayed segments run the per-launch stage
    kernels with the prologue, the forward the fused graph steps."""
    monkeypatch.setenv("ETOHEOM_ADJ_REPLAY", "1")
    p = configs.heptamer(L=1, nsteps=40)
    p.pulses[:, 1] = [2.0, 6.0, 10.0]
    p.pulses[:, 2] = 3.0
    p.pulses[:, 0] *= 3e5
    P0, _ = O.twod(oproblem(p), nv=(1, 1, 1))
    rng = np.random.default_rng(5)
    w = rng.normal(size=len(P0)) + 1j * rng.normal(size=len(P0))
    b = E.Batch(ctx, [{"problem": p, "H": p.H, "X": p.X, "nv": (1, 1, 1)}], adjoint=7)
    b.run()
    b.sync()
    assert np.max(np.abs(b.output(0) - P0)) / np.max(np.abs(P0)) < 1e-10
    b.backward([w])
    g = b.grad(0)
    for name, idx, part in [("H", (3, 3), "re"), ("X", (1, 0, 5), "re"), ("gamma", (1, 0), "re"),
                            ("alpha_t", (2, 1), "im")]:
        ref = fd(p, w, name, idx, part)
        got = adj_value(g, name, idx, part)
        assert abs(got - ref) / max(abs(ref), 1e-12) < 1e-6, (name, idx, got, ref)


def test_adjoint_three_modes_per_bath(ctx):
    """K_m = 3: the modes beyond the two-phase gathers go through gather_rest in the
    adjoint's small-N kernel as well."""
    p = configs.aggregate(2, 2, nsteps=100, taus=(10.0, 20.0, 30.0))
    p.alpha = np.concatenate([p.alpha, 0.5 * p.alpha[:, :1]], axis=1) * 8.0
    p.alpha_t = np.concatenate([p.alpha_t, 0.5 * p.alpha_t[:, :1]], axis=1) * 8.0
    p.gamma = np.concatenate([p.gamma, 2.0 * p.gamma[:, :1]], axis=1)
    p.pulses[:, 2] = 8.0
    p.pulses[:, 0] *= 3e5
    assert p.alpha.shape == (2, 3)
    P0, _ = O.twod(oproblem(p), nv=(1, 1, 1))
    rng = np.random.default_rng(6)
    w = rng.normal(size=len(P0)) + 1j * rng.normal(size=len(P0))
    b = E.Batch(ctx, [{"problem": p, "H": p.H, "X": p.X, "nv": (1, 1, 1)}], adjoint=16)
    b.run()
    b.sync()
    assert np.max(np.abs(b.output(0) - P0)) / np.max(np.abs(P0)) < 1e-10
    b.backward([w])
    g = b.grad(0)
    for name, idx, part in [("H", (1, 1), "re"), ("gamma", (0, 2), "re"), ("alpha", (1, 2), "re"),
                            ("alpha_t", (0, 2), "im"), ("X", (0, 0, 2), "re")]:
        ref = fd(p, w, name, idx, part)
        got = adj_value(g, name, idx, part)
        assert abs(got - ref) / max(abs(ref), 1e-12) < 1e-6, (name, idx, got, ref)
