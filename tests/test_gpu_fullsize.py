"""GPU checks at the benchmark sizes, where whole oracle propagations take hours.

* heptamer N=29, L=4 (3060 ADOs, the benchmark configuration): the first steps of a
  driven run against the oracle at the full hierarchy size (<= 1e-10); over the full
  2000 steps, Tr rho_0 = 1 (every term feeding rho_0 is a commutator, so the trace is
  conserved exactly), finite readouts, and the graph-replayed batch bitwise equal to
  the plain per-launch path and to a second copy of the instance in the same batch;
* the 512-propagation grid: sampled instances of the full batch bitwise equal to the
  same instances solved alone (eh_2des_solve), traces conserved;
* the 10-site chain N=56, L=3 (1771 ADOs, 3000 steps): trace conserved, finite; four
  propagations of its (t1, t2) grid in one batch bitwise equal to the same propagations
  solved alone.
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


def oproblem(p):
    return O.Problem(H=p.H, X=p.X, alpha=p.alpha, alpha_t=p.alpha_t, gamma=p.gamma, L=p.L,
                     site_offset=p.site_offset, pulses=p.pulses, t_start=p.t_start, t_end=p.t_end,
                     step_size=p.step_size, print_step=p.print_step)


def item(p):
    return {"problem": p, "H": p.H, "X": p.X, "nv": (1, 1, 1)}


def test_heptamer_full_hierarchy_first_steps_match_oracle(ctx):
    p = configs.heptamer(nsteps=16)                 # L = 4: 3060 ADOs
    assert p.num_ado() == 3060
    p.pulses[:, 1] = [1.0, 3.0, 5.0]
    p.pulses[:, 2] = 2.0
    p.pulses[:, 0] *= 1e5
    P, rho = E.twoD_spectrum_solver(ctx, p, p.H, 1, 1, 1, X=p.X, want_rho=True)
    Pr, rr = O.twod(oproblem(p), nv=(1, 1, 1), want_rho=True)
    assert np.max(np.abs(rr[1:])) > 1e-12 * np.max(np.abs(rr[0]))    # the hierarchy is driven
    assert np.max(np.abs(P - Pr)) / np.max(np.abs(Pr)) < 1e-10
    assert np.max(np.abs(rho - rr)) / np.max(np.abs(rr)) < 1e-10


def test_heptamer_benchmark_run_properties(ctx):
    p = configs.heptamer()                          # the benchmark instance, 2000 steps
    b = E.Batch(ctx, [item(p), item(p)])
    b.run()
    b.sync()
    P0, P1 = b.output(0), b.output(1)
    assert np.all(np.isfinite(P0)) and np.max(np.abs(P0)) > 0
    assert np.array_equal(P0, P1)
    rho = b.rho(0)
    assert abs(np.trace(rho[0]) - 1.0) < 1e-12
    b.run_eager()
    b.sync()
    assert np.array_equal(b.output(0), P0)
    assert np.array_equal(b.rho(0), rho)


def test_grid_batch_equals_instances_solved_alone(ctx):
    items = configs.grid2d()
    assert len(items) == 512
    b = E.Batch(ctx, [item(p) for p in items])
    b.run()
    b.sync()
    for i in (0, 97, 258, 511):                      # dimer, trimer and tetramer instances
        p = items[i]
        P, rho = E.twoD_spectrum_solver(ctx, p, p.H, 1, 1, 1, X=p.X, want_rho=True)
        assert np.array_equal(b.output(i), P), i
        r = b.rho(i)
        assert np.array_equal(r, rho), i
        assert abs(np.trace(r[0]) - 1.0) < 1e-12


def test_chain_n56_full_run_conserves_trace(ctx):
    p = configs.chain()                              # N = 56, L = 3, 3000 steps
    assert p.N == 56 and p.num_ado() == 1771
    P, rho = E.twoD_spectrum_solver(ctx, p, p.H, 1, 1, 1, X=p.X, want_rho=True)
    assert np.all(np.isfinite(P)) and np.max(np.abs(P)) > 0
    assert abs(np.trace(rho[0]) - 1.0) < 1e-12


def test_chain_grid_batch_equals_propagations_solved_alone(ctx):
    items = configs.chain_grid(n_t1=2, t2s=(0.0, 400.0))    # 4 of the grid's propagations, full size
    b = E.Batch(ctx, [item(p) for p in items])
    b.run()
    b.sync()
    for i in (0, 3):
        p = items[i]
        P, rho = E.twoD_spectrum_solver(ctx, p, p.H, 1, 1, 1, X=p.X, want_rho=True)
        assert np.array_equal(b.output(i), P), i
        assert np.array_equal(b.rho(i), rho), i
    assert not np.array_equal(b.output(0), b.output(3))
