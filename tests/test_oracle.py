"""Pins the CPU oracle (oracle/heom_oracle.c) without any reference run.

The reference C++ path needs GSL 2.8, which this image lacks, and the reference
holds no golden vectors for the propagator, so the oracle is pinned by checks
whose expected values come from elsewhere:

* brute-force enumeration and the closed-form ADO count C(d+L, L);
* an independent dense-Liouvillian build of the UNSCALED hierarchy in
  Kronecker form (tests/heom_dense.py): one RHS on a random state, and whole
  propagations with pulses through the exact RK4 polynomial;
* physics closed forms: the second-order cumulant of a pure-dephasing
  two-level system (exact for a Gaussian bath), and the first-order response
  of a resonant Gaussian pulse (its area, sqrt(pi/2.77) tau_p erf(2 sqrt 2.77));
* invariants: Tr rho_0 conservation and Hermiticity of rho_0 without a field.

Each check is chosen so that a plausible slip (alpha and alpha~ swapped, a
missing sqrt scaling, a wrong sign of the pulse phase, 4 ln 2 for 2.77, an
untransposed dipole, V not frozen over the step) fails at least one of them.
"""
import math

import numpy as np
import pytest

import heom_dense as D
import oracle as O

# ETO modes in the reference key-template style: (a, b, g, w), +/-w pairs
MODES = [(0.00311213, -0.00252971, 0.12340683, 6e-8), (0.00311213, -0.00252971, 0.12340683, -6e-8),
         (-0.00296194, 0.00259667, 0.12950563, 0.02683512), (-0.00296194, 0.00259667, 0.12950563, -0.02683512)]


def bath(K, modes, scale=1.0):
    al = np.array([[complex(a, b) * scale for a, b, g, w in modes]] * K)
    alt = np.array([[complex(a, -b) * scale for a, b, g, w in modes]] * K)
    ga = np.array([[complex(g, w) for a, b, g, w in modes]] * K)
    return al, alt, ga


def dimer_2des():
    """single_size=3 (ground + 2 singles) + 1 double -> N=4, built like construct_Hal."""
    hb = 5308.0
    H = np.zeros((4, 4))
    H[1, 1], H[2, 2], H[1, 2], H[2, 1] = 12000 / hb, 12100 / hb, 100 / hb, 100 / hb
    H[3, 3] = H[1, 1] + H[2, 2]
    X = np.zeros((4, 4, 4), complex)
    for c in range(4):
        X[c][0, 1], X[c][0, 2] = 1.0 + 0.1 * c, 0.5 - 0.2 * c
        X[c][1, 3], X[c][2, 3] = X[c][0, 2], X[c][0, 1]
    return H.astype(complex), X


def test_ado_enumeration_matches_brute_force_and_closed_form():
    for d, L in [(1, 0), (1, 5), (2, 3), (4, 2), (4, 6), (6, 5), (3, 0)]:
        ado = O.build_ado(d, L)
        assert len(ado) == math.comb(d + L, L)
        assert [tuple(int(v) for v in r) for r in ado] == D.ado_list(d, L)
        assert tuple(ado[0]) == (0,) * d
    # up-link count d*C(d+L-1, L-1) (each ADO with |n|<L links to d parents)
    d, L = 4, 6
    ado = O.build_ado(d, L)
    ups = sum(d for r in ado if r.sum() < L)
    assert ups == d * math.comb(d + L - 1, L - 1) == 504


def _scales(ados, alpha, K):
    """c_n = sqrt(prod_p n_p! |alpha_p|^n_p) relating scaled and unscaled ADOs"""
    c = []
    for n in ados:
        v = 1.0
        for p, np_ in enumerate(n):
            j, k = p % K, p // K
            v *= math.factorial(np_) * abs(alpha[j, k]) ** np_
        c.append(math.sqrt(v))
    return np.array(c)


@pytest.mark.parametrize("site_offset", [1, 0])
def test_rhs_equals_dense_unscaled_liouvillian(site_offset):
    rng = np.random.default_rng(5)
    N, K, L = 4, 2, 3
    al, alt, ga = bath(K, MODES[:2] + MODES[2:3])
    alt = alt * (1 + 0.3j)   # decouple alpha~ from conj(alpha) so a swap cannot hide
    V = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    A, ados = D.liouvillian(V, al, alt, ga, L, site_offset)
    M = len(ados)
    yu = rng.normal(size=(M, N, N)) + 1j * rng.normal(size=(M, N, N))
    c = _scales(ados, al, K)
    ys = yu / c[:, None, None]
    p = O.Problem(H=np.zeros((N, N), complex), alpha=al, alpha_t=alt, gamma=ga, L=L,
                  X=np.zeros((4, N, N), complex), site_offset=site_offset)
    ds = O.rhs(p, V, ys)
    du = (A @ yu.reshape(-1)).reshape(M, N, N)
    err = np.max(np.abs(ds * c[:, None, None] - du)) / np.max(np.abs(du))
    assert err < 1e-14


@pytest.mark.parametrize("L", [0, 2])
def test_2des_propagation_equals_dense_rk4_polynomial(L):
    H, X = dimer_2des()
    al, alt, ga = bath(2, MODES[:2], scale=3.0)
    hb = 5308.0
    pul = np.array([[2e3 / hb, 20.0, 8.0, 12050 / hb], [2e3 / hb, 35.0, 8.0, 12050 / hb],
                    [2e3 / hb, 50.0, 8.0, 12050 / hb]])
    kw = dict(t_start=0.0, t_end=80.0, step_size=0.5, print_step=1.0)
    p = O.Problem(H=H, alpha=al, alpha_t=alt, gamma=ga, L=L, X=X, pulses=pul, **kw)
    for nv in [(1, 1, 1), (1, 0, 1)]:
        P, rho = O.twod(p, nv=nv, want_rho=True)
        Pd, r0 = D.propagate_2des(H, X, al, alt, ga, L, pul, kw["t_start"], kw["t_end"],
                                  kw["step_size"], kw["print_step"], nv=nv)
        assert len(P) == len(Pd) == 31   # t in [50, 80] every 1 fs: 31 readouts before the step
        assert np.max(np.abs(P - Pd)) / np.max(np.abs(Pd)) < 1e-11
        assert np.max(np.abs(rho[0] - r0)) / np.max(np.abs(r0)) < 1e-11


def test_trace_of_rho0_is_conserved_with_field_and_bath():
    H, X = dimer_2des()
    al, alt, ga = bath(2, MODES, scale=2.0)
    hb = 5308.0
    pul = np.array([[300 / hb, 20.0, 8.0, 12050 / hb]] * 3)
    pul[1, 1], pul[2, 1] = 30.0, 40.0
    p = O.Problem(H=H, alpha=al, alpha_t=alt, gamma=ga, L=3, X=X, pulses=pul,
                  t_start=0, t_end=100, step_size=0.5, print_step=0.5)
    _, rho = O.twod(p, nv=(1, 1, 1), want_rho=True)
    assert abs(np.trace(rho[0]) - 1.0) < 1e-13 * max(1.0, np.max(np.abs(rho[0])))
    assert abs(rho[0][0, 0] - 1.0) > 1e-4    # the field moved population
    assert np.max(np.abs(rho[1:])) > 1e-6    # the hierarchy was actually driven


def test_rho0_stays_hermitian_without_field():
    """With +/-w paired modes (equal alpha), alpha~ = conj(alpha) makes rho_0 Hermitian."""
    H = np.array([[400, 100], [100, 500]]) / 5308.0
    al, alt, ga = bath(2, MODES, scale=1.0)
    p = O.Problem(H=H.astype(complex), alpha=al, alpha_t=alt, gamma=ga, L=3, site_offset=0,
                  t_start=0, t_end=200, step_size=1, print_step=10)
    M = p.num_ado()
    r0 = np.zeros((M, 2, 2), complex)
    r0[0] = [[0.6, 0.2 + 0.1j], [0.2 - 0.1j, 0.4]]
    pop, rho = O.dynamics(p, sites=(1, 2), rho_init=r0, want_rho=True)
    assert np.max(np.abs(rho[0] - rho[0].conj().T)) < 1e-14
    assert abs(pop[0, -1] + pop[1, -1] - 1.0) < 1e-13
    # swapping alpha~ for alpha breaks Hermiticity: the check is sensitive to it
    p_bad = O.Problem(H=H.astype(complex), alpha=al, alpha_t=al, gamma=ga, L=3, site_offset=0,
                      t_start=0, t_end=200, step_size=1, print_step=10)
    _, rho_bad = O.dynamics(p_bad, sites=(1, 2), rho_init=r0, want_rho=True)
    assert np.max(np.abs(rho_bad[0] - rho_bad[0].conj().T)) > 1e-6


def test_pure_dephasing_matches_second_order_cumulant():
    """H = eps|1><1|, S = |1><1|: rho_10(t) = rho_10(0) exp(-i eps t - g(t)),
    g(t) = sum_p alpha_p/gamma_p^2 (exp(-gamma_p t) + gamma_p t - 1)  (exact for this bath)."""
    eps = 0.1
    al, alt, ga = bath(1, MODES)
    H = np.diag([0, eps]).astype(complex)
    X = np.zeros((4, 2, 2), complex)
    X[3][0, 1] = 1.0                      # Tr(rho X3) = rho_10
    pul = np.array([[0, 0, 1, 0], [0, 0, 1, 0], [0, -1.0, 1, 0]])
    p = O.Problem(H=H, alpha=al, alpha_t=alt, gamma=ga, L=6, X=X, pulses=pul,
                  t_start=0, t_end=200, step_size=0.25, print_step=0.25)
    r0 = np.zeros((p.num_ado(), 2, 2), complex)
    r0[0] = 0.5
    P, _ = O.twod(p, nv=(0, 0, 0), rho_init=r0)
    t = np.arange(len(P)) * 0.25

    def g(alpha):
        return sum(a / gg ** 2 * (np.exp(-gg * t) + gg * t - 1) for a, gg in zip(alpha, ga[0]))

    exact = 0.5 * np.exp(-1j * eps * t - g(al[0]))
    assert np.max(np.abs(P - exact)) / 0.5 < 5e-8
    wrong = 0.5 * np.exp(-1j * eps * t - g(alt[0]))   # alpha~ in place of alpha
    assert np.max(np.abs(P - wrong)) / 0.5 > 1e-2


@pytest.mark.parametrize("q", [0, 1, 2])
def test_resonant_pulse_first_order_area(q):
    """Bath-free two-level system, one weak resonant pulse.  First-order theory
    with V frozen over each step gives, after the pulse,
      rho_coh(t) e^{-+i eps (t - tau)} / (-+i mu) = A phi,
      A = E0 tau_p sqrt(pi/2.77) erf(2 sqrt(2.77)),  phi = step phase factor.
    Pulse 1 uses X[0] with e^{+i w (t - tau)}; pulses 2, 3 use X[q]^T with e^{-i w (t - tau)}."""
    eps, mu, E0, tau, taup, h = 0.1, 0.7, 1e-7, 40.0, 15.0, 1 / 16
    H = np.diag([0, eps]).astype(complex)
    X = np.zeros((4, 2, 2), complex)
    X[q][0, 1] = mu
    if q == 0:
        X[3][1, 0] = 1.0   # P = rho_01
    else:
        X[3][0, 1] = 1.0   # P = rho_10
    pul = np.zeros((3, 4))
    pul[:, 2] = 1.0
    pul[q] = [E0, tau, taup, eps]
    if q != 2:
        pul[2, 1] = 0.0
    nv = [0, 0, 0]
    nv[q] = 1
    p = O.Problem(H=H, alpha=np.array([[1.0 + 0j]]), alpha_t=np.array([[1.0 + 0j]]),
                  gamma=np.array([[0.1 + 0j]]), L=0, X=X, pulses=pul,
                  t_start=0, t_end=128, step_size=h, print_step=h)
    P, _ = O.twod(p, nv=tuple(nv))
    t = np.arange(len(P)) * h + (0.0 if q != 2 else tau)
    A = E0 * taup * math.sqrt(math.pi / 2.77) * math.erf(2 * math.sqrt(2.77))
    if q == 0:
        phi = (1 - np.exp(-1j * eps * h)) / (1j * eps * h)
        got = P * np.exp(-1j * eps * (t - tau)) * 1j / mu
    else:
        phi = (np.exp(1j * eps * h) - 1) / (1j * eps * h)
        got = P * np.exp(1j * eps * (t - tau)) / (1j * mu)
    sel = t > tau + 2 * taup + 1
    assert sel.sum() > 100
    assert np.max(np.abs(got[sel] - A * phi)) / A < 2e-6
    A_4ln2 = E0 * taup * math.sqrt(math.pi / (4 * math.log(2))) * math.erf(2 * math.sqrt(4 * math.log(2)))
    assert abs(A_4ln2 - A) / A > 2e-4   # the check separates the two Gaussian constants


def test_dynamics_mode_equals_dense():
    H = np.array([[400, 100], [100, 500]]) / 5308.0
    al, alt, ga = bath(2, MODES[:2], scale=2.0)
    L = 3
    p = O.Problem(H=H.astype(complex), alpha=al, alpha_t=alt, gamma=ga, L=L, site_offset=0,
                  t_start=0, t_end=60, step_size=2, print_step=10)
    pop, rho = O.dynamics(p, sites=(1, 2), want_rho=True)
    A, ados = D.liouvillian(H.astype(complex), al, alt, ga, L, site_offset=0)
    R = D.rk4_matrix(A, 2.0)
    y = np.zeros(A.shape[0], complex)
    y[0] = 1.0
    want = []
    for i in range(31):
        if i % 5 == 0:
            want.append([y[0].real, y[3].real])
        y = R @ y
    want = np.array(want).T
    assert pop.shape == (2, 7)
    assert np.max(np.abs(pop - want)) < 1e-13
    assert np.max(np.abs(rho[0].reshape(-1) - y[:4])) < 1e-13


def test_field_window_and_transposition():
    H, X = dimer_2des()
    X = X + 1j * 0   # complex dtype
    hb = 5308.0
    pul = np.array([[2e3 / hb, 20.0, 8.0, 12050 / hb], [2e3 / hb, 36.0, 8.0, 12050 / hb],
                    [2e3 / hb, 52.0, 8.0, 12050 / hb]])
    p = O.Problem(H=H, alpha=np.ones((2, 1), complex), alpha_t=np.ones((2, 1), complex),
                  gamma=np.ones((2, 1), complex), L=1, X=X, pulses=pul)
    for t in [3.0, 4.0, 36.0, 52.0, 68.0, 68.5, 100.0]:
        V = O.field_V(p, t, (1, 1, 1))
        Vd = D.V_of_t(H, X, pul, np.float32(t), (1, 1, 1))
        assert np.max(np.abs(V - Vd)) < 1e-15
    assert np.array_equal(O.field_V(p, 3.5, (1, 1, 1)), H)        # outside every window
    assert not np.array_equal(O.field_V(p, 4.0, (1, 1, 1)), H)    # |t - tau| == 2 tau_p is inside
    assert np.array_equal(O.field_V(p, 44.0, (1, 0, 0)), H)       # pulse 1 window is [4, 36]
    assert not np.array_equal(O.field_V(p, 44.0, (0, 1, 0)), H)   # pulse 2 window is [20, 52]
    assert np.array_equal(O.field_V(p, 44.0, (1, 0, 0)), O.field_V(p, 44.0, (0, 0, 0)))


def test_long_run_fixtures_are_this_oracles_output():
    """tests/golden/*.npz (made by make_fixtures.py) hold the oracle's output for their own
    inputs: re-run the trimer fixture's first 500 steps and compare its first readouts bit
    for bit; check every fixture's sizes."""
    import os
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    for name, nado, npts in [("trimer_L5_2000", 462, 1620), ("tetramer_L6_2000", 3003, 1620),
                             ("heptamer_L4_600", 3060, 220)]:
        f = np.load(os.path.join(gold, name + ".npz"))
        t = f["time"]
        q = O.Problem(H=f["H"], X=f["X"], alpha=f["alpha"], alpha_t=f["alpha_t"], gamma=f["gamma"], L=int(f["L"]),
                      site_offset=int(f["site_offset"]), pulses=f["pulses"], t_start=float(t[0]),
                      t_end=float(t[1]), step_size=float(t[2]), print_step=float(t[3]))
        assert q.num_ado() == nado == len(f["ado_norms"]) and len(f["P"]) == npts
        assert np.max(np.abs(f["P"])) > 0
        if name == "trimer_L5_2000":
            P, _ = O.twod(q, nv=tuple(int(v) for v in f["nv"]), max_steps=500)
            assert len(P) == 120 and np.array_equal(P, f["P"][:120])
