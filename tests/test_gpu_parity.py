"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Tolerance: max normwise relative error <= 1e-10 for P(t), populations and the
final hierarchy state, per propagation (FP64 throughout; the two sides differ
only in summation order).
"""
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


def assert_driven(rho):
    """the hierarchy beyond rho_0 carries signal (otherwise a state comparison pins nothing)"""
    assert np.max(np.abs(rho[1:])) > 1e-12 * np.max(np.abs(rho[0]))


def test_context(ctx):
    print("device:", ctx.name, "CUs:", ctx.n_cu, "LDS/block:", ctx.lds_per_block)
    assert ctx.n_cu > 0


@pytest.mark.parametrize("nv", [(1, 1, 1), (1, 1, 0), (1, 0, 1)])
def test_dimer_2des_matches_oracle(ctx, nv):
    p = configs.dimer(L=6, nsteps=600)
    P, rho = E.twoD_spectrum_solver(ctx, p, p.H, *nv, X=p.X, want_rho=True)
    Pr, rr = O.twod(oproblem(p), nv=nv, want_rho=True)
    assert len(P) == len(Pr) > 100
    assert rel(P, Pr) < TOL
    assert rel(rho, rr) < TOL


def test_dynamics_key_matches_oracle(ctx):
    k = E.Param.from_key("tests/golden/key_0_0.key", E.MODE_DYNAMICS)
    pop = E.dynamics_solver(ctx, k, k.get("Hal"), sites=(1, 2))
    q = O.Problem(H=k.get("Hal"), alpha=k.get("alpha"), alpha_t=k.get("alpha_t"), gamma=k.get("gamma"),
                  L=k.L, site_offset=0, t_start=0, t_end=1000, step_size=2, print_step=10)
    popr, _ = O.dynamics(q, sites=(1, 2))
    assert pop.shape == popr.shape == (2, 101)
    assert rel(pop, popr) < TOL


def test_dynamics_seventy_site_key_matches_oracle(ctx, tmp_path):
    """Population dynamics above N = 64: a 70-site chain key (SIZE 70, sites only, as the
    reference's dynamics keys are: N = 70, site 1 excited at t = 0) with baths on five sites
    (K = 5, L = 1, 11 ADOs), V = H and site offset 0, three site populations over 40 steps
    against the oracle."""
    n = 70
    eps = 12000.0 + 10.0 * np.arange(n)
    J = np.zeros((n, n))
    for i in range(n - 1):
        J[i, i + 1] = J[i + 1, i] = 100.0
    dip = [(0, 1, 1.0, 0.0, 0.0)]
    F = configs.OHMIC_35_100_77
    txt = configs.key_text(n, 5, 1, np.diag(eps) + J, dip, configs.ohmic_modes(), F["lambda_line"],
                           [(1e-4, t, 15.0, 12050.0) for t in (40.0, 90.0, 190.0)], (0.0, 20.0, 0.5, 0.5))
    key = tmp_path / "seventy_dyn.key"
    key.write_text(txt)
    k = E.Param.from_key(str(key), E.MODE_DYNAMICS)
    H = k.get("Hal")
    assert H.shape == (70, 70)
    pop = E.dynamics_solver(ctx, k, H, sites=(1, 2, 3))
    q = O.Problem(H=H, alpha=k.get("alpha"), alpha_t=k.get("alpha_t"), gamma=k.get("gamma"), L=k.L, site_offset=0,
                  t_start=0, t_end=20.0, step_size=0.5, print_step=0.5)
    popr, _ = O.dynamics(q, sites=(1, 2, 3))
    assert pop.shape == popr.shape == (3, 41)
    assert np.min(np.max(np.abs(popr[1:]), axis=1)) > 1e-3
    assert rel(pop, popr) < TOL


@pytest.mark.parametrize("n_sites,L", [(3, 4), (4, 3)])
def test_trimer_tetramer_match_oracle(ctx, n_sites, L):
    p = configs.aggregate(n_sites, L, sigma=100.0, nsteps=300, taus=(20.0, 40.0, 60.0))
    P, rho = E.twoD_spectrum_solver(ctx, p, p.H, 1, 1, 1, X=p.X, want_rho=True)
    Pr, rr = O.twod(oproblem(p), nv=(1, 1, 1), want_rho=True)
    assert_driven(rr)
    assert rel(P, Pr) < TOL
    assert rel(rho, rr) < TOL


def test_heptamer_generic_path_matches_oracle(ctx):
    """N = 29 (the large-N kernel) on a truncated run: L = 2 and 60 steps."""
    p = configs.heptamer(L=2, nsteps=60)
    p.pulses[:, 1] = [5.0, 10.0, 0.0]   # pulses inside the run, read out from t = 0
    p.pulses[:, 2] = 4.0
    p.pulses[:, 0] *= 1e5                # strong field: the hierarchy carries O(1e-3) of the state
    P, rho = E.twoD_spectrum_solver(ctx, p, p.H, 1, 1, 1, X=p.X, want_rho=True)
    Pr, rr = O.twod(oproblem(p), nv=(1, 1, 1), want_rho=True)
    assert_driven(rr)
    assert rel(P, Pr) < TOL
    assert rel(rho, rr) < TOL


def test_batch_of_mixed_sizes_matches_oracle(ctx):
    items = [configs.aggregate(2, 3, nsteps=200, taus=(20.0, 30.0, 40.0)),
             configs.aggregate(3, 2, nsteps=150, taus=(20.0, 35.0, 45.0), sigma=50.0),
             configs.aggregate(4, 2, nsteps=180, taus=(30.0, 20.0, 50.0)),
             configs.dimer(L=4, nsteps=250, t1=20.0, t2=20.0)]
    b = E.Batch(ctx, [{"problem": p, "H": p.H, "X": p.X, "nv": (1, 1, 1)} for p in items])
    b.run()
    b.sync()
    for i, p in enumerate(items):
        Pr, rr = O.twod(oproblem(p), nv=(1, 1, 1), want_rho=True)
        assert rel(b.output(i), Pr) < TOL
        assert rel(b.rho(i), rr) < TOL
    # replaying the same batch gives identical results
    first = [b.output(i) for i in range(len(items))]
    b.run()
    b.sync()
    for i in range(len(items)):
        assert np.array_equal(b.output(i), first[i])


def test_key_driver_matches_sampled_oracle(ctx, tmp_path):
    txt = open("tests/golden/cpu_dimer_ohmic.key-tmpl").read()
    txt = txt.replace("TAU1", "30").replace("TAU2", "50").replace("TAU3", "70").replace("T_END", "120")
    txt = txt.replace("50 114", "2 114")          # two samples keeps the oracle run short
    key = tmp_path / "k.key"
    key.write_text(txt)
    P = E.run_2des_key(ctx, str(key))
    k = E.Param.from_key(str(key))
    H, X = k.samples()
    acc = 0
    for s in range(k.n_sample):
        q = O.Problem(H=H[s], X=X[s], alpha=k.get("alpha"), alpha_t=k.get("alpha_t"), gamma=k.get("gamma"),
                      L=k.L, pulses=np.array([[k.system().pulse[a][c] for c in range(4)] for a in range(3)]),
                      t_start=0, t_end=120, step_size=2, print_step=10)
        p1, _ = O.twod(q, nv=(1, 1, 1))
        p2, _ = O.twod(q, nv=(1, 1, 0))
        p3, _ = O.twod(q, nv=(1, 0, 1))
        acc = acc + (p1 - p2 - p3) / k.n_sample
    assert len(P) == len(acc) == 6
    # the subtraction cancels the first-order signal: compare against the scale of p1
    assert np.max(np.abs(P - acc)) / np.max(np.abs(p1)) < TOL


def test_chain_n56_mfma64_matches_oracle(ctx):
    """N = 56 (64-padded MFMA tiles) on a truncated run: L = 1 and 30 steps."""
    p = configs.chain(L=1, nsteps=30)
    p.pulses[:, 1] = [3.0, 6.0, 0.0]
    p.pulses[:, 2] = 3.0
    p.pulses[:, 0] *= 1e5
    P, rho = E.twoD_spectrum_solver(ctx, p, p.H, 1, 1, 1, X=p.X, want_rho=True)
    Pr, rr = O.twod(oproblem(p), nv=(1, 1, 1), want_rho=True)
    assert_driven(rr)
    assert rel(P, Pr) < TOL
    assert rel(rho, rr) < TOL


def test_chain_n56_every_launch_candidate_matches_oracle(ctx):
    """N = 56 on every launch candidate of its N class (the MFMA kernels with and without
    column copies, the generic one): 24 pulse-window steps with complex V, then 46 with
    V = H, a batch of two instances."""
    ps = []
    for t1 in (0.0, 2.0):
        p = configs.chain(L=1, nsteps=70)
        p.pulses[:, 1] = [3.0, 6.0 + t1, 0.0]
        p.pulses[:, 2] = 3.0
        p.pulses[:, 0] *= 1e5
        ps.append(p)
    refs = [O.twod(oproblem(p), nv=(1, 1, 1), want_rho=True) for p in ps]
    b = E.Batch(ctx, [{"problem": p, "H": p.H, "X": p.X, "nv": (1, 1, 1)} for p in ps])
    cands = b.candidates()
    assert len(cands) >= 3
    # V's pattern has at most 13 entries a row: the sparse kernels are offered and the
    # row-pattern one is the default; stage_mfma64 stays the base (candidate 0) whose tiling
    # the adjoint follows
    rep = b.tuning_report().splitlines()
    assert [k for _, _, k, _ in cands][0] == "stage_mfma64"
    assert any(" stage_sparse_rows " in l and l.endswith("*") for l in rep), rep
    for g, c, kern, _ in cands:
        b.select(g, c)
        b.run()
        b.sync()
        for i, (Pr, rr) in enumerate(refs):
            assert_driven(rr)
            assert rel(b.output(i), Pr) < TOL, (kern, i)
            assert rel(b.rho(i), rr) < TOL, (kern, i)
    b.close()


def test_mixed_sizes_in_the_sparse_group_every_candidate(ctx):
    """One 32 < N <= 64 group holding N = 37, 46 and 56 (8- and 9-site aggregates, the chain;
    L = 1, 40 steps through three strong pulses): every instance lays its rho, couplings and
    pattern lists out in LDS at its own N and K inside the group's allocation.  Every launch
    candidate of the group against the oracle."""
    ps = [configs.aggregate(8, 1, nsteps=40, taus=(2.0, 5.0, 8.0)),
          configs.aggregate(9, 1, nsteps=40, taus=(3.0, 6.0, 9.0)),
          configs.chain(L=1, nsteps=40)]
    ps[2].pulses[:, 1] = [2.5, 5.5, 8.5]
    for p in ps:
        p.pulses[:, 2] = 3.0
        p.pulses[:, 0] *= 1e5
    assert [p.N for p in ps] == [37, 46, 56]
    refs = [O.twod(oproblem(p), nv=(1, 1, 1), want_rho=True) for p in ps]
    b = E.Batch(ctx, [{"problem": p, "H": p.H, "X": p.X, "nv": (1, 1, 1)} for p in ps])
    cands = b.candidates()
    assert len({g for g, _, _, _ in cands}) == 1
    assert {"stage_sparse", "stage_sparse_rows"} <= {k for _, _, k, _ in cands}
    for g, c, kern, _ in cands:
        b.select(g, c)
        b.run()
        b.sync()
        for i, (Pr, rr) in enumerate(refs):
            assert_driven(rr)
            assert rel(b.output(i), Pr) < TOL, (kern, i)
            assert rel(b.rho(i), rr) < TOL, (kern, i)
    b.close()


def test_cli_2des_prints_reference_format(ctx, tmp_path):
    import subprocess
    txt = open("tests/golden/cpu_dimer_ohmic.key-tmpl").read()
    txt = txt.replace("TAU1", "30").replace("TAU2", "50").replace("TAU3", "70").replace("T_END", "120")
    txt = txt.replace("50 114", "3 114")
    key = tmp_path / "k.key"
    key.write_text(txt)
    out = subprocess.run(["eto-heom_amd/bin/etoheom_2des", str(key)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    P = E.run_2des_key(ctx, str(key))
    want = "".join(f"{v.real:.6e} {v.imag:.6e}\n" for v in P)
    assert out.stdout == want and len(P) == 6


def test_cli_2des_eleven_site_key(ctx, tmp_path):
    """The key-file driver above N = 64: an 11-site .key (SIZE 12: ground + 11 singles, and
    the 55 doubles that construct_Hal adds, N = 67), L = 1, 17 steps with three strong pulses
    inside them (E0 = 100 cm^-1: the third-order signal is then 3e-3 of the first-order terms
    it is the difference of, so their 1e-10 agreement stays far below the print's digits).
    etoheom_2des prints p111 - p110 - p101 in the reference's format, and its values match the
    oracle fed the same realised system to the print's 7 digits."""
    import subprocess
    n = 11
    eps = 12000.0 + 100.0 * np.arange(n)
    J = np.zeros((n, n))
    for i in range(n - 1):
        J[i, i + 1] = J[i + 1, i] = 100.0
    rng = np.random.default_rng(11)
    mu = rng.normal(size=(n, 3))
    mu /= np.linalg.norm(mu, axis=1, keepdims=True)
    dip = [(0, i + 1, *mu[i]) for i in range(n)]
    F = configs.OHMIC_35_100_77
    txt = configs.key_text(n + 1, n, 1, configs._sites_H(eps, J), dip, configs.ohmic_modes(), F["lambda_line"],
                           [(100.0, t, 3.0, 12050.0) for t in (0.5, 1.5, 2.5)], (0.0, 8.0, 0.5, 0.5))
    key = tmp_path / "eleven.key"
    key.write_text(txt)
    p, _ = configs.realise(txt, 0)
    assert p.N == 67
    out = subprocess.run(["eto-heom_amd/bin/etoheom_2des", str(key)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    got = np.array([complex(float(a), float(b)) for a, b in (ln.split() for ln in out.stdout.splitlines())])
    q = oproblem(p)
    S = O.twod(q, nv=(1, 1, 1))[0] - O.twod(q, nv=(1, 1, 0))[0] - O.twod(q, nv=(1, 0, 1))[0]
    assert len(got) == len(S) > 0 and np.max(np.abs(S)) > 0
    assert np.max(np.abs(got - S)) <= 2e-6 * np.max(np.abs(S))


def test_cli_dynamics_prints_reference_format(ctx):
    import subprocess
    out = subprocess.run(["eto-heom_amd/bin/etoheom_dynamics", "tests/golden/key_0_0.key"], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.split("\n")
    assert lines[0] == "Now is CPU version" and lines[1] == "Running tests/golden/key_0_0.key"
    assert lines[3] == "site 1 population: " and lines[4] == ""
    k = E.Param.from_key("tests/golden/key_0_0.key", E.MODE_DYNAMICS)
    pop = E.dynamics_solver(ctx, k, k.get("Hal"), sites=(1, 2))
    got1 = [float(x) for x in lines[5:5 + 101]]
    assert np.allclose(got1, pop[0], rtol=1e-5, atol=1e-12)
    assert lines[5] == "1" and lines[5 + 101] == "" and lines[5 + 102] == "site 2 population: "
