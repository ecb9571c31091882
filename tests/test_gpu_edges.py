"""GPU parity at the edges of the supported shapes, against the CPU oracle (<= 1e-10
normwise relative, FP64):

* N = 64 (no padding in the 64 x 64 MFMA tiles), with a random Hermitian H and random
  dipoles; N = 60 with a banded H and sparse dipoles on every launch candidate (the
  sparse kernels at eight rows per lane);
* N > 64 on the wide kernel, one or several workgroups per ADO: the 11-site aggregate
  (N = 67), random systems of N = 128 and 256, the 16-site aggregate (N = 137), and N = 67,
  100 and 137 together in one group;
* L = 0: the hierarchy is rho_0 alone (no neighbour links);
* K = 0: no bath at all (d = 0, one ADO), including N = 1;
* instances of different lengths and N classes in one batch, each read out to its own
  end (ragged step counts);
* population-dynamics instances with different site lists beside a 2DES instance in one
  batch (more output sites than N is an error, as in the oracle).
"""
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


def like(p, **over):
    kw = dict(H=p.H, X=p.X, alpha=p.alpha, alpha_t=p.alpha_t, gamma=p.gamma, L=p.L, site_offset=p.site_offset,
              pulses=p.pulses.copy(), t_start=p.t_start, t_end=p.t_end, step_size=p.step_size,
              print_step=p.print_step)
    kw.update(over)
    return E.Problem(**kw)


def strong(p, taus=(2.0, 5.0, 8.0), taup=3.0):
    p.pulses[:, 1] = taus
    p.pulses[:, 2] = taup
    p.pulses[:, 0] *= 1e5
    return p


def check(ctx, p, driven=True):
    P, rho = E.twoD_spectrum_solver(ctx, p, p.H, 1, 1, 1, X=p.X, want_rho=True)
    Pr, rr = O.twod(oproblem(p), nv=(1, 1, 1), want_rho=True)
    if driven and len(rr) > 1:
        assert np.max(np.abs(rr[1:])) > 1e-12 * np.max(np.abs(rr[0]))
    assert rel(P, Pr) < TOL
    assert rel(rho, rr) < TOL


def test_n64_random_hermitian_system(ctx):
    base = configs.chain(L=1, nsteps=24)
    rng = np.random.default_rng(64)
    N = 64
    A = rng.normal(size=(N, N)) * 0.02
    H = np.diag(rng.uniform(2.2, 2.45, N)) + (A + A.T) / 2
    H[0, :] = 0.0
    H[:, 0] = 0.0
    X = np.zeros((4, N, N), complex)
    for q in range(4):
        v = rng.normal(size=N - 1)
        X[q, 0, 1:] = v
        X[q, 1:, 0] = v
    p = strong(like(base, H=H.astype(complex), X=X, alpha=base.alpha[:3], alpha_t=base.alpha_t[:3],
                    gamma=base.gamma[:3]))
    assert p.N == 64
    # a dense H has 64 entries a row: no sparse-commutator candidate
    b = E.Batch(ctx, [{"problem": p, "H": p.H, "X": p.X, "nv": (1, 1, 1)}])
    assert not {"stage_sparse", "stage_sparse_rows"} & {k for _, _, k, _ in b.candidates()}
    b.close()
    check(ctx, p)


def test_n60_banded_system_every_candidate(ctx):
    """N = 60 (eight rows per lane in the row-pattern sparse kernel, whose stage 1 then runs
    the column-list kernel): a banded random H (|a - c| <= 3) and dipoles from the ground
    state to ten states, so V has at most 11 entries a row; three baths, L = 1, 40 steps with
    the pulses' complex V first, then real V.  Every launch candidate against the oracle."""
    base = configs.chain(L=1, nsteps=40)
    rng = np.random.default_rng(60)
    N = 60
    H = np.diag(rng.uniform(2.2, 2.45, N))
    for k in range(1, 4):
        v = rng.normal(size=N - k) * 0.02
        H += np.diag(v, k) + np.diag(v, -k)
    H[0, :] = 0.0
    H[:, 0] = 0.0
    X = np.zeros((4, N, N), complex)
    for q in range(4):
        v = rng.normal(size=10)
        X[q, 0, 1:11] = v
        X[q, 1:11, 0] = v
    p = strong(like(base, H=H.astype(complex), X=X, alpha=base.alpha[:3], alpha_t=base.alpha_t[:3],
                    gamma=base.gamma[:3]))
    assert p.N == 60
    b = E.Batch(ctx, [{"problem": p, "H": p.H, "X": p.X, "nv": (1, 1, 1)}])
    cands = b.candidates()
    assert {"stage_sparse", "stage_sparse_rows"} <= {k for _, _, k, _ in cands}
    Pr, rr = O.twod(oproblem(p), nv=(1, 1, 1), want_rho=True)
    assert np.max(np.abs(rr[1:])) > 1e-12 * np.max(np.abs(rr[0]))
    for g, c, kern, _ in cands:
        print(f"N=60 candidate {c}: {kern}", flush=True)
        b.select(g, c)
        b.run()
        b.sync()
        assert rel(b.output(0), Pr) < TOL, kern
        assert rel(b.rho(0), rr) < TOL, kern
    b.close()


def test_depth_zero_hierarchy(ctx):
    p = strong(configs.dimer(L=0, nsteps=60))
    check(ctx, p, driven=False)


def test_no_bath(ctx):
    p = configs.dimer(L=2, nsteps=60)
    empty = np.zeros((0, 0), complex)
    check(ctx, strong(like(p, alpha=empty, alpha_t=empty, gamma=empty, L=2)), driven=False)
    one = strong(like(p, H=np.array([[0.3 + 0.0j]]), X=np.ones((4, 1, 1), complex), alpha=empty, alpha_t=empty,
                      gamma=empty, L=1))
    P, rho = E.twoD_spectrum_solver(ctx, one, one.H, 1, 1, 1, X=one.X, want_rho=True)
    Pr, rr = O.twod(oproblem(one), nv=(1, 1, 1), want_rho=True)
    assert len(P) == len(Pr) > 0
    assert rel(P, Pr) < TOL
    assert rel(rho, rr) < TOL


def test_ragged_batch_each_instance_to_its_own_end(ctx):
    items = [strong(configs.dimer(L=2, nsteps=37)), strong(configs.aggregate(3, 2, nsteps=301)),
             strong(configs.dimer(L=3, nsteps=1)), strong(configs.aggregate(2, 2, nsteps=129))]
    b = E.Batch(ctx, [{"problem": p, "H": p.H, "X": p.X, "nv": (1, 1, 1)} for p in items])
    b.run()
    b.sync()
    for i, p in enumerate(items):
        Pr, rr = O.twod(oproblem(p), nv=(1, 1, 1), want_rho=True)
        assert len(b.output(i)) == len(Pr), i
        if len(Pr):
            assert rel(b.output(i), Pr) < TOL, i
        assert rel(b.rho(i), rr) < TOL, i


def test_mixed_modes_and_site_lists_in_one_batch(ctx):
    """Population dynamics with different site lists next to a 2DES propagation in one
    batch (the dynamics instances: site offset 0, V = H, readouts Re rho_0[s-1][s-1])."""
    k = E.Param.from_key("tests/golden/key_0_0.key", E.MODE_DYNAMICS)
    H = k.get("Hal")
    base = dict(H=H, alpha=k.get("alpha"), alpha_t=k.get("alpha_t"), gamma=k.get("gamma"), L=k.L)
    d1 = E.Problem(**base, site_offset=0, t_start=0, t_end=400, step_size=2, print_step=10)
    d2 = E.Problem(**base, site_offset=0, t_start=0, t_end=300, step_size=2, print_step=6)
    tw = strong(configs.dimer(L=2, nsteps=80))
    items = [{"problem": d1, "H": H, "mode": E.MODE_DYNAMICS, "sites": [1, 2]},
             {"problem": tw, "H": tw.H, "X": tw.X, "nv": (1, 1, 1)},
             {"problem": d2, "H": H, "mode": E.MODE_DYNAMICS, "sites": [2, 1]}]
    b = E.Batch(ctx, items)
    b.run()
    b.sync()
    for i, sites, p in ((0, (1, 2), d1), (2, (2, 1), d2)):
        q = O.Problem(H=H, alpha=p.alpha, alpha_t=p.alpha_t, gamma=p.gamma, L=p.L, site_offset=0,
                      t_start=p.t_start, t_end=p.t_end, step_size=p.step_size, print_step=p.print_step)
        popr, _ = O.dynamics(q, sites=sites)
        pop = b.output(i)
        assert pop.shape == popr.shape, i
        assert rel(pop, popr) < TOL, i
    Pr, _ = O.twod(oproblem(tw), nv=(1, 1, 1))
    assert rel(b.output(1), Pr) < TOL
    with pytest.raises(Exception):
        E.Batch(ctx, [{"problem": d2, "H": H, "mode": E.MODE_DYNAMICS, "sites": [2, 1, 2]}])


def test_n67_eleven_site_aggregate_every_candidate(ctx):
    """N = 67 (11 sites: ground + 11 singles + 55 doubles, the first size above 64),
    L = 1 (23 ADOs), 40 steps through three strong pulses: the wide MFMA kernel (operands
    and couplings from global memory) and the LDS-staged VALU kernel, each against the
    oracle."""
    p = strong(configs.aggregate(11, 1, nsteps=40))
    assert p.N == 67
    b = E.Batch(ctx, [{"problem": p, "H": p.H, "X": p.X, "nv": (1, 1, 1)}])
    cands = b.candidates()
    kinds = {k for _, _, k, _ in cands}
    assert "stage_generic" in kinds
    if os.environ.get("ETOHEOM_NO_MFMA") != "1":
        assert "stage_mfma_wide" in kinds
    Pr, rr = O.twod(oproblem(p), nv=(1, 1, 1), want_rho=True)
    assert np.max(np.abs(rr[1:])) > 1e-12 * np.max(np.abs(rr[0]))
    for g, c, kern, _ in cands:
        print(f"N=67 candidate {c}: {kern}", flush=True)
        b.select(g, c)
        b.run()
        b.sync()
        assert len(b.output(0)) == len(Pr) > 0
        assert rel(b.output(0), Pr) < TOL, kern
        assert rel(b.rho(0), rr) < TOL, kern
    b.close()


def random_system(N, seed, nsteps, taus=(2.0, 5.0, 8.0)):
    """Random Hermitian H (ground state uncoupled), random ground-to-excited dipoles, three
    baths of two modes on sites 1..3, L = 1 (7 ADOs), strong pulses."""
    base = configs.chain(L=1, nsteps=nsteps)
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(N, N)) * 0.02
    H = np.diag(rng.uniform(2.2, 2.45, N)) + (A + A.T) / 2
    H[0, :] = 0.0
    H[:, 0] = 0.0
    X = np.zeros((4, N, N), complex)
    for q in range(4):
        v = rng.normal(size=N - 1)
        X[q, 0, 1:] = v
        X[q, 1:, 0] = v
    p = strong(like(base, H=H.astype(complex), X=X, alpha=base.alpha[:3], alpha_t=base.alpha_t[:3],
                    gamma=base.gamma[:3]), taus=taus)
    assert p.N == N and p.num_ado() == 7
    return p


def every_candidate(ctx, p, splits):
    """p on every launch candidate against the oracle; the wide kernel must be offered with
    exactly these workgroups per ADO."""
    b = E.Batch(ctx, [{"problem": p, "H": p.H, "X": p.X, "nv": (1, 1, 1)}])
    cands = b.candidates()
    assert sorted(t for _, _, k, t in cands if k == "stage_mfma_wide") == sorted(splits)
    Pr, rr = O.twod(oproblem(p), nv=(1, 1, 1), want_rho=True)
    assert np.max(np.abs(rr[1:])) > 1e-12 * np.max(np.abs(rr[0]))
    for g, c, kern, t in cands:
        print(f"N={p.N} candidate {c}: {kern} {t}", flush=True)
        b.select(g, c)
        b.run()
        b.sync()
        assert len(b.output(0)) == len(Pr) > 0
        assert rel(b.output(0), Pr) < TOL, (kern, t)
        assert rel(b.rho(0), rr) < TOL, (kern, t)
    b.close()


def test_n128_random_hermitian_system(ctx):
    """N = 128: random Hermitian H, three baths, L = 1, 20 steps, on the wide kernel with
    one workgroup per ADO and with its 64 tiles over four; and through the single-propagation
    entry point."""
    p = random_system(128, 128, 20)
    every_candidate(ctx, p, [1, 4])
    check(ctx, p)


def test_n137_sixteen_site_aggregate_every_candidate(ctx):
    """N = 137 (16 sites: ground + 16 singles + 120 doubles), L = 1 (33 ADOs), 8 steps
    with the three strong pulses inside them: the wide kernel with one workgroup per ADO
    and with each ADO's 81 tiles over six."""
    p = strong(configs.aggregate(16, 1, nsteps=8), taus=(0.5, 1.5, 2.5))
    assert p.N == 137 and p.num_ado() == 33
    every_candidate(ctx, p, [1, 6])


def test_mixed_large_n_in_one_group_every_candidate(ctx):
    """N = 67, 100 and 137 in one batch: one N > 64 group, whose spread launch sizes its
    workgroups per ADO for N = 137 (six), so the smaller instances' extra workgroups find no
    tile and exit. Each instance against the oracle on every candidate."""
    ps = [strong(configs.aggregate(11, 1, nsteps=8), taus=(0.5, 1.5, 2.5)),
          random_system(100, 100, 8, taus=(0.5, 1.5, 2.5)),
          strong(configs.aggregate(16, 1, nsteps=8), taus=(0.5, 1.5, 2.5))]
    assert [p.N for p in ps] == [67, 100, 137]
    b = E.Batch(ctx, [{"problem": p, "H": p.H, "X": p.X, "nv": (1, 1, 1)} for p in ps])
    cands = b.candidates()
    assert len({g for g, _, _, _ in cands}) == 1
    assert sorted(t for _, _, k, t in cands if k == "stage_mfma_wide") == [1, 6]
    refs = [O.twod(oproblem(p), nv=(1, 1, 1), want_rho=True) for p in ps]
    for g, c, kern, t in cands:
        print(f"mixed N > 64 candidate {c}: {kern} {t}", flush=True)
        b.select(g, c)
        b.run()
        b.sync()
        for i, (Pr, rr) in enumerate(refs):
            assert rel(b.output(i), Pr) < TOL, (kern, t, ps[i].N)
            assert rel(b.rho(i), rr) < TOL, (kern, t, ps[i].N)
    b.close()


def test_n256_random_hermitian_system(ctx):
    """N = 256 (16 x 16 tiles of 16 x 16 per ADO), 8 steps with the pulses inside them: the
    wide kernel with one workgroup per ADO and with 16 per ADO."""
    every_candidate(ctx, random_system(256, 256, 8, taus=(0.5, 1.5, 2.5)), [1, 16])

