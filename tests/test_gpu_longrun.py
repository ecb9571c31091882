"""Long-horizon GPU parity: whole propagations at the physical field amplitude
(E0 = 1e-4 cm^-1) against the CPU oracle, at <= 1e-10 normwise relative error.

* the dimer workload in full (N = 4, L = 6, 210 ADOs, 2000 steps) for the three
  pulse orderings the reference's driver runs (p111, p110, p101,
  etoheom-cpu/src/CPU_2DES.cpp:65-78), oracle computed here;
* fixtures written by tests/golden/make_fixtures.py with the oracle (inputs and
  outputs in the file): a trimer of the adjoint grid (N = 7, L = 5, 2000 steps,
  the whole final hierarchy), a tetramer of the batched 2D grid (N = 11, L = 6,
  3003 ADOs, 2000 steps) and the heptamer benchmark system with its full L = 4
  hierarchy (3060 ADOs) through all three pulse windows and its first 220
  readouts (600 steps); for the two larger ones the final rho_0, the first tier
  and the norm of every ADO; the 10-site chain of the benchmark grid with its full
  L = 3 hierarchy (1771 ADOs, N = 56) over 60 steps with strong, early pulses, on every
  launch candidate of its N class.

At this amplitude rho_0 stays within O(E0^2) of |0><0|, so besides the usual
normwise error over the whole state (which the 1 in rho_0[0][0] dominates) the
ADOs beyond rho_0 are compared on their own scale, and P(t), which is O(E0),
carries the response."""
import os

import numpy as np
import pytest

import etoheom_amd as E
import oracle as O
from etoheom_amd import configs

pytestmark = pytest.mark.gpu
TOL = 1e-10
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def ctx():
    return E.Context(0)


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, (a.shape, b.shape)
    assert b.size and np.max(np.abs(b)) > 0, "reference is identically zero: the check would be vacuous"
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def oproblem(p):
    return O.Problem(H=p.H, X=p.X, alpha=p.alpha, alpha_t=p.alpha_t, gamma=p.gamma, L=p.L,
                     site_offset=p.site_offset, pulses=p.pulses, t_start=p.t_start, t_end=p.t_end,
                     step_size=p.step_size, print_step=p.print_step)


@pytest.mark.parametrize("nv", [(1, 1, 1), (1, 1, 0), (1, 0, 1)])
def test_dimer_whole_run_at_physical_field(ctx, nv):
    p = configs.dimer()   # L = 6, 2000 steps of 0.5 fs, pulses at 40 / 90 / 190 fs, E0 = 1e-4
    assert p.pulses[0, 0] == pytest.approx(1e-4 / 5308.0)
    Pr, rr = O.twod(oproblem(p), nv=nv, want_rho=True)
    P, rho = E.twoD_spectrum_solver(ctx, p, p.H, *nv, X=p.X, want_rho=True)
    assert len(P) == len(Pr) == 1620
    print(f"nv={nv}: P err {rel(P, Pr):.2e} (|P|max {np.max(np.abs(Pr)):.2e}), state err {rel(rho, rr):.2e}, "
          f"ADOs beyond rho_0 {rel(rho[1:], rr[1:]):.2e}")
    assert rel(P, Pr) < TOL
    assert rel(rho, rr) < TOL
    assert rel(rho[1:], rr[1:]) < TOL
    # rho_0 without its (0, 0) entry (1 - O(E0^2), rounded at 1e-16 absolute): the coherences
    # and excited populations the pulses made, on their own scale
    m = np.ones(rr[0].shape, bool)
    m[0, 0] = False
    assert rel(rho[0][m], rr[0][m]) < TOL


def load(name):
    f = np.load(os.path.join(GOLDEN, name + ".npz"))
    t = f["time"]
    p = E.Problem(H=f["H"], X=f["X"], alpha=f["alpha"], alpha_t=f["alpha_t"], gamma=f["gamma"], L=int(f["L"]),
                  site_offset=int(f["site_offset"]), pulses=f["pulses"], t_start=float(t[0]),
                  t_end=float(t[1]), step_size=float(t[2]), print_step=float(t[3]))
    return p, tuple(int(v) for v in f["nv"]), f


def check_fixture(f, P, rho, name):
    assert len(P) == len(f["P"])
    norms = np.sqrt(np.sum(np.abs(rho) ** 2, axis=(1, 2)))
    errs = {"P": rel(P, f["P"]), "ADO norms": rel(norms, f["ado_norms"]),
            "norms beyond rho_0": rel(norms[1:], f["ado_norms"][1:])}
    if "rho" in f:
        errs["state"] = rel(rho, f["rho"])
        errs["state beyond rho_0"] = rel(rho[1:], f["rho"][1:])
    else:
        idx = f["tier01_index"]
        errs["rho_0"] = rel(rho[0], f["rho_tier01"][0])
        errs["first tier"] = rel(rho[idx[1:]], f["rho_tier01"][1:])
    print(name, {k: f"{v:.2e}" for k, v in errs.items()})
    for k, v in errs.items():
        assert v < TOL, (k, v)


@pytest.mark.parametrize("name,npoints,nado", [("trimer_L5_2000", 1620, 462), ("tetramer_L6_2000", 1620, 3003),
                                               ("heptamer_L4_600", 220, 3060), ("chain_L3_60", 28, 1771)])
def test_whole_runs_match_oracle_fixtures(ctx, name, npoints, nado):
    p, nv, f = load(name)
    assert p.num_ado() == nado
    b = E.Batch(ctx, [{"problem": p, "H": p.H, "X": p.X, "nv": nv}])
    b.run()
    b.sync()
    P, rho = b.output(0), b.rho(0)
    b.close()
    assert len(P) == npoints
    check_fixture(f, P, rho, name)


def test_chain_full_hierarchy_every_launch_candidate(ctx):
    """The chain fixture (N = 56, L = 3, 1771 ADOs, 45 steps with complex V then 15 with
    V = H) on every launch candidate of its N class, the sparse-commutator default and the
    dense MFMA kernels among them."""
    p, nv, f = load("chain_L3_60")
    b = E.Batch(ctx, [{"problem": p, "H": p.H, "X": p.X, "nv": nv}])
    cands = b.candidates()
    assert {"stage_sparse_rows", "stage_sparse", "stage_mfma64"} <= {k for _, _, k, _ in cands}
    for g, c, kern, _ in cands:
        b.select(g, c)
        b.run()
        b.sync()
        check_fixture(f, b.output(0), b.rho(0), f"chain_L3_60 {kern}")
    b.close()
