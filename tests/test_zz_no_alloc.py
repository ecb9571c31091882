"""The allocation-free entry point eh_2des_solve_noalloc against eh_2des_solve.

Both run the same stage kernels on the same host-built tables (the allocation-free
one launches them one by one over the library's static device scratch array, the
other replays a graph over its own allocations), so P(t) and the final hierarchy
state must agree bit for bit.  Three calls in a row, each smaller than the one
before (the scratch array keeps the previous call's contents), then a propagation
too large for the array (an error, no device work), then a call that still works.
"""
import numpy as np
import pytest

import etoheom_amd as E
from etoheom_amd import configs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return E.Context(0)


def _both(ctx, p, nv):
    print(f"noalloc: N={p.N} L={p.L} nv={nv}", flush=True)
    P0, r0 = E.twoD_spectrum_solver(ctx, p, p.H, *nv, X=p.X, want_rho=True)
    P1, r1 = E.twoD_spectrum_solver_noalloc(ctx, p, p.H, *nv, X=p.X, want_rho=True)
    return (P0, r0), (P1, r1)


def _same(a, b):
    (P0, r0), (P1, r1) = a, b
    assert len(P0) == len(P1) > 10
    assert np.max(np.abs(P0)) > 0 and np.max(np.abs(r0[1:])) > 0   # driven: the comparison pins something
    assert np.array_equal(P0, P1), float(np.max(np.abs(P0 - P1)))
    assert np.array_equal(r0, r1), float(np.max(np.abs(r0 - r1)))


def test_noalloc_equals_allocating_solve_on_shrinking_inputs(ctx):
    # N = 56 (the MFMA kernel with padded V planes), 21 ADOs, 300 steps through the pulses
    _same(*_both(ctx, configs.chain(L=1, nsteps=300, t1=10.0, t2=20.0), (1, 1, 1)))
    # the N = 7 trimer at its full depth L = 5 (462 ADOs), a whole 2000-step run
    _same(*_both(ctx, configs.aggregate(3, 5), (1, 1, 0)))
    # the N = 4 dimer, L = 3, 600 steps: smaller again
    _same(*_both(ctx, configs.dimer(L=3, nsteps=600), (1, 0, 1)))


def test_noalloc_rejects_a_propagation_beyond_its_scratch(ctx):
    p = configs.heptamer(L=4, nsteps=20)   # 3060 ADOs of 29 x 29: about 170 MB of state
    with pytest.raises(E.EtoHeomError) as ei:
        E.twoD_spectrum_solver_noalloc(ctx, p, p.H, 1, 1, 1, X=p.X)
    assert "EH_ERANGE" in str(ei.value) and "scratch" in str(ei.value)
    q = configs.dimer(L=2, nsteps=600)
    _same(*_both(ctx, q, (1, 1, 1)))
