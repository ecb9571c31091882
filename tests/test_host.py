"""CPU-side checks of the product library: exports, key parser, hierarchy/CSR,
construct_Hal, dipole assembly and the sample-loop generator.  No GPU calls."""
import math
import re

import numpy as np
import pytest

import etoheom_amd as E
import oracle as O

HB = 5308.0


def test_library_exports_every_declared_symbol():
    hdr = open("include/etoheom.h").read()
    declared = set(re.findall(r"^\s*(?:const char\*|int64_t|int)\s+(eh_\w+)\s*\(", hdr, re.M))
    assert len(declared) >= 30
    assert declared == set(E.EXPORTS)
    lib = E.lib()
    for name in declared:
        assert hasattr(lib, name), name


def test_no_device_reports_error_not_crash():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(E.EtoHeomError) as ei:
        E.Context(0)
    assert ei.value.code == -5


@pytest.mark.parametrize("d,L", [(4, 6), (6, 5), (14, 2), (3, 0), (1, 4)])
def test_hierarchy_matches_oracle_order_and_links(d, L):
    ado, up, dn = E.hierarchy_build(d, L)
    ref = O.build_ado(d, L)
    assert ado.shape == ref.shape == (math.comb(d + L, L), d)
    assert np.array_equal(ado, ref)
    M = len(ado)
    for i in range(M):
        for p in range(d):
            if up[i, p] >= 0:
                assert ado[up[i, p]].sum() == ado[i].sum() + 1 and ado[up[i, p], p] == ado[i, p] + 1
                assert dn[up[i, p], p] == i          # n -> n+e_p -> n
            else:
                assert ado[i].sum() == L
            assert (dn[i, p] >= 0) == (ado[i, p] > 0)


def test_parse_dynamics_key_fixture():
    k = E.Param.from_key("tests/golden/key_0_0.key", E.MODE_DYNAMICS)
    assert (k.sys_size, k.K, k.K_m, k.L, k.M) == (2, 2, 4, 2, math.comb(10, 2))
    lam = 0.00037678975131876413
    H = k.get("Hal")
    assert np.allclose(H.real, np.array([[400, 100], [100, 500]]) / HB + lam * np.eye(2), rtol=0, atol=1e-15)
    a = k.get("alpha")
    assert a.shape == (2, 4) and a[1, 2] == complex(-0.00112609, -0.00015085)
    assert np.array_equal(k.get("alpha_t"), a.conj())
    assert k.get("gamma")[0, 3] == complex(0.01619040, -0.00109453)
    s = k.system()
    assert (s.t_start, s.t_end, s.step_size, s.print_step, s.site_offset) == (0, 1000, 2, 10, 0)


def _template(tau=(30, 50, 70), t_end=120):
    txt = open("tests/golden/cpu_dimer_ohmic.key-tmpl").read()
    for name, v in zip(("TAU1", "TAU2", "TAU3", "T_END"), (*tau, t_end)):
        txt = txt.replace(name, str(v))
    return txt


def test_parse_2des_template(tmp_path):
    f = tmp_path / "k.key"
    f.write_text(_template())
    k = E.Param.from_key(str(f))
    assert (k.single_size, k.sys_size, k.K, k.K_m, k.L, k.n_sample, k.seed) == (3, 4, 2, 6, 2, 50, 114)
    H = k.get("Hal")
    lam = 0.008995192714312095
    want = np.zeros((4, 4))
    want[1:3, 1:3] = np.array([[400, 100], [100, 500]]) / HB
    want[1, 1] += lam
    want[2, 2] += lam
    assert np.allclose(H.real, want, rtol=0, atol=1e-15) and not H.imag.any()
    D = k.get("Disorder")
    assert D[1, 1] == 42.4 / HB and D[2, 2] == 42.4 / HB and D[0, 0] == 0
    # DIPOLE: |0><1| x, |0><2| y; double column (12): X(1,(12)) = X(0,2), X(2,(12)) = X(0,1)
    Xx, Xy = k.get("X_x"), k.get("X_y")
    assert Xx[0, 1] == 1 and Xx[2, 3] == 1 and Xx[1, 3] == 0
    assert Xy[0, 2] == 1 and Xy[1, 3] == 1 and Xy[2, 3] == 0
    s = k.system()
    pul = np.array([[s.pulse[q][c] for c in range(4)] for q in range(3)])
    assert np.allclose(pul[:, 0], 5 / HB) and np.allclose(pul[:, 3], 500 / HB)
    assert list(pul[:, 1]) == [30, 50, 70] and list(pul[:, 2]) == [20, 20, 20]
    assert np.allclose(k.get("angle"), 0.0)


def test_parse_errors(tmp_path):
    with pytest.raises(E.EtoHeomError) as ei:
        E.Param.from_key(str(tmp_path / "missing.key"))
    assert ei.value.code == -2
    bad = _template().replace("PULSE\n3", "PULSE\n2")
    f = tmp_path / "b.key"
    f.write_text(bad)
    with pytest.raises(E.EtoHeomError) as ei:
        E.Param.from_key(str(f))
    assert ei.value.code == -8


def _two_exciton_block(Hs):
    """Independent: the two-exciton block of a Frenkel Hamiltonian (hard-core bosons),
    <ab|H|cd> for a<b, c<d, built from second quantisation on explicit basis states."""
    n = Hs.shape[0]
    pairs = [(a, b) for a in range(n) for b in range(a + 1, n)]
    B = np.zeros((len(pairs), len(pairs)))
    for i, (a, b) in enumerate(pairs):
        for j, (c, d) in enumerate(pairs):
            # sum_{kl} H_kl <ab| b_k^+ b_l |cd>: move one excitation l -> k
            v = 0.0
            for k in range(n):
                for l in range(n):
                    if l not in (c, d):
                        continue
                    rest = {c, d} - {l}
                    if k in rest:
                        continue
                    if rest | {k} == {a, b}:
                        v += Hs[k, l]
            B[i, j] = v
    return B


def test_construct_hal_sort_and_double_block(tmp_path):
    rng = np.random.default_rng(3)
    n = 4
    eps = np.array([12300.0, 12100.0, 12400.0, 12000.0])
    J = rng.uniform(-100, 100, (n, n))
    J = np.triu(J, 1) + np.triu(J, 1).T
    from etoheom_amd import configs
    Hcm = configs._sites_H(eps, J)
    txt = configs.key_text(n + 1, n, 1, Hcm, [(0, i + 1, 1.0, 0.0, 0.0) for i in range(n)],
                           configs.ohmic_modes(), 0.0, [(1, 10, 5, 1)] * 3, (0, 10, 1, 1))
    f = tmp_path / "h.key"
    f.write_text(txt)
    k = E.Param.from_key(str(f))
    H = k.construct_hal()
    N = k.sys_size
    assert N == 5 + 6
    singles = H[1:5, 1:5].real
    assert np.allclose(np.diag(singles), np.sort(eps) / HB)   # ascending sort of the diagonal
    assert np.allclose(singles - np.diag(np.diag(singles)), J / HB)
    B = _two_exciton_block(singles)
    assert np.allclose(H[5:, 5:].real, B, atol=1e-15)
    assert not H[1:5, 5:].any() and not H[0].any()
    # disorder draws enter before the sort
    dH = np.zeros((5, 5))
    dH[1, 1] = 1.0
    H2 = k.construct_hal(dH)
    assert np.isclose(H2[4, 4].real, eps[0] / HB + 1.0)   # site 1 shifted up by 1 sorts last


def test_polar_and_pulse_interaction(tmp_path):
    f = tmp_path / "k.key"
    f.write_text(_template().replace("POLARIZATION\n0 0 0 0", "POLARIZATION\n0 90 0 90"))
    k = E.Param.from_key(str(f))
    pol = k.polar_default()
    assert np.allclose(pol, [[0, 0, 0, 0], [0, 1, 0, 1], [1, 6.123233995736766e-17, 1, 6.123233995736766e-17]])
    c, a = np.array([0.3, -0.5, 0.8]), np.array([0.9, 0.1, -0.2])
    R = E.polar_rotate(c, a, pol)
    cn = c / np.linalg.norm(c)
    an = a - a.dot(cn) * cn
    an /= np.linalg.norm(an)
    rot = np.array([an, np.cross(cn, an), cn])
    assert np.allclose(R, rot @ pol, atol=1e-15)
    assert np.allclose(rot @ rot.T, np.eye(3), atol=1e-14)
    k.set_polar(R)
    X = k.get("X")
    Xc = [k.get(n) for n in ("X_x", "X_y", "X_z")]
    for i in range(4):
        assert np.allclose(X[i], sum(R[j, i] * Xc[j] for j in range(3)), atol=1e-15)


def test_generator_is_mt19937():
    # std::mt19937's defining check value: 10000th output for seed 5489
    assert E.rng_stream(5489, 10000)[-1] == 4123659995
    bg = np.random.MT19937()
    bg._legacy_seeding(114)
    assert np.array_equal(E.rng_stream(114, 2000), bg.random_raw(2000).astype(np.uint32))
    # seed 0 is mapped to 4357
    assert np.array_equal(E.rng_stream(0, 10), E.rng_stream(4357, 10))


def test_samples_are_physical(tmp_path):
    f = tmp_path / "k.key"
    f.write_text(_template())
    k = E.Param.from_key(str(f))
    H, X = k.samples(200)
    assert H.shape == (200, 4, 4) and X.shape == (200, 4, 4, 4)
    d1 = H[:, 1, 1].real - (400 / HB + 0.008995192714312095)
    d2 = H[:, 2, 2].real - (500 / HB + 0.008995192714312095)
    # sorted after the draw: the smaller site energy lands on index 1
    assert np.all(H[:, 1, 1].real <= H[:, 2, 2].real)
    assert np.allclose(H[:, 3, 3], H[:, 1, 1] + H[:, 2, 2])
    sig = 42.4 / HB
    both = np.concatenate([d1, d2])
    assert abs(both.std() - sig * 1.0) < 0.5 * sig   # loose: sorting reshapes the marginals
    # rotated polarization vectors are unit vectors: |X[i][0,1]|^2 + |X[i][0,2]|^2 = 1 for 0-degree pulses
    nrm = np.abs(X[:, 0, 0, 1]) ** 2 + np.abs(X[:, 0, 0, 2]) ** 2
    assert np.all(nrm <= 1 + 1e-12)
    assert np.all(np.abs(X[:, 0]) <= 1 + 1e-12)


def test_torch_system_builder_matches_key_path(tmp_path):
    """The differentiable H/X construction equals the key parser + construct_Hal + dipole path."""
    import torch
    from etoheom_amd import autograd as AG, configs
    rng = np.random.default_rng(11)
    n = 3
    eps = np.array([12400.0, 12150.0, 12300.0])
    J = rng.uniform(-100, 100, (n, n))
    J = np.triu(J, 1) + np.triu(J, 1).T
    mu = rng.normal(size=(n, 3))
    lam = configs.OHMIC_35_100_77["lambda_line"]
    txt = configs.key_text(n + 1, n, 1, configs._sites_H(eps, J), [(0, i + 1, *mu[i]) for i in range(n)],
                           configs.ohmic_modes(), lam, [(1, 10, 5, 1)] * 3, (0, 10, 1, 1),
                           polarization=(0, 90, 0, 90))
    f = tmp_path / "t.key"
    f.write_text(txt)
    k = E.Param.from_key(str(f))
    pol = k.polar_default()
    k.set_polar(pol)
    H, X = AG.build_2des_system(torch.tensor(eps), torch.tensor(J), torch.tensor(mu), lam, pol)
    assert np.allclose(H.numpy(), k.construct_hal(), rtol=0, atol=1e-15)
    assert np.allclose(X.numpy(), k.get("X"), rtol=0, atol=1e-15)


def _chain_item(L, nsteps=50):
    """An N = 56 (10-site) 2DES system with 2 modes per bath at depth L, built without a key file."""
    n = 10
    N = 1 + n + n * (n - 1) // 2
    al = np.full((n, 2), 3.3e-5 - 1.5e-5j)
    p = E.Problem(H=np.diag(np.linspace(0, 1, N)).astype(complex), alpha=al, alpha_t=al.conj(),
                  gamma=np.full((n, 2), 0.023 + 0j), L=L, X=np.zeros((4, N, N), complex), site_offset=1,
                  pulses=np.array([[1e-4 / HB, 40.0, 15.0, 2.27]] * 3), t_start=0.0,
                  t_end=0.5 * (nsteps - 1), step_size=0.5, print_step=0.5)
    return {"problem": p, "H": p.H, "X": p.X, "nv": (1, 1, 1)}


def test_batch_plan_rejects_instances_beyond_32bit_offsets():
    # 10-site chain at L = 7: 888,030 ADOs x 56^2 = 2.78e9 state elements in one instance,
    # beyond the kernels' 32-bit in-instance offsets: refused on the host, before any device work
    with pytest.raises(E.EtoHeomError) as ei:
        E.batch_plan([_chain_item(7)])
    assert ei.value.code == -7 and "2^31" in str(ei.value)
    # L = 6 (230,230 ADOs, 7.2e8 elements) is accepted
    pl = E.batch_plan([_chain_item(6)])
    assert pl["total_ados"] == math.comb(26, 6) and pl["total_elems"] == math.comb(26, 6) * 56 * 56


def test_batch_plan_bytes():
    items = [_chain_item(3, nsteps=3000) for _ in range(4)]
    pl = E.batch_plan(items, checkpoint=32)
    T = 4 * 1771 * 56 * 56
    assert pl["total_elems"] == T and pl["max_steps"] == 3000
    # four state buffers and three coupled-column copies (K N per ADO) dominate the forward
    state = 16 * (4 * T + 3 * 4 * 1771 * 10 * 56)
    assert state <= pl["forward_bytes"] < state * 1.01
    # checkpoints every 32 steps (94 segments + 1), 4 (32 + 1) stage inputs, replay accumulator,
    # lambda and its three work buffers
    adj = 16 * T * (95 + 132 + 1 + 4)
    assert adj <= pl["adjoint_replay_bytes"] < adj * 1.01
    assert pl["adjoint_traj_bytes"] > 16 * T * (4 * 3000 + 1)
    # instance checks: N range, bath sites, missing X4
    bad = _chain_item(1)
    bad["X"] = None
    with pytest.raises(E.EtoHeomError):
        E.batch_plan([bad])


def test_batch_plan_system_size_range():
    """N up to 2048 (the key parser's bound) is accepted: the N > 64 kernels keep nothing
    of rho in LDS.  2049 is refused on the host with EH_ESHAPE, and an instance of 2^31
    state elements or more (32-bit in-instance offsets) with EH_ERANGE."""
    def item(N, L=1):
        al = np.full((3, 2), 3.3e-5 - 1.5e-5j)
        p = E.Problem(H=np.eye(N, dtype=complex), alpha=al, alpha_t=al.conj(), gamma=np.full((3, 2), 0.023 + 0j),
                      L=L, X=np.zeros((4, N, N), complex), site_offset=1,
                      pulses=np.array([[1e-4 / HB, 40.0, 15.0, 2.27]] * 3), t_start=0.0, t_end=5.0,
                      step_size=0.5, print_step=0.5)
        return {"problem": p, "H": p.H, "X": p.X, "nv": (1, 1, 1)}
    for N in (65, 67, 128, 137, 256, 2048):
        pl = E.batch_plan([item(N)])
        assert pl["total_elems"] == 7 * N * N
    with pytest.raises(E.EtoHeomError) as ei:
        E.batch_plan([item(2049)])
    assert ei.value.code == -3 and "1..2048" in str(ei.value)
    # d = 6, L = 7: 1716 ADOs of 2048^2 elements
    with pytest.raises(E.EtoHeomError) as ei:
        E.batch_plan([item(2048, L=7)])
    assert "2^31" in str(ei.value)


def test_float_reciprocal_splits_are_exact():
    """The stage kernels split a coupling-gather entry e by a runtime divisor n as
    int((float(e) + 0.5f) * (1.0f / float(n))) in IEEE float32 (device.hpp entry_split and
    gather_couplings).  Over the ranges they use, that equals e // n: e < 2 K N <= 2^13 for
    n = N (17..64) and n = 2 N (34..128), and the quotient by K (1..64) of values below 2^16.
    numpy's float32 rounds as the device does (correctly rounded division, multiply, add)."""
    f = np.float32
    for divisors, limit in ((range(17, 65), 1 << 13), (range(34, 129, 2), 1 << 13), (range(1, 65), 1 << 16)):
        e = np.arange(limit, dtype=np.int64)
        ef = e.astype(np.float32)
        for n in divisors:
            inv = f(1.0) / f(n)
            q = ((ef + f(0.5)) * inv).astype(np.int64)   # truncation, as the device's conversion
            assert np.array_equal(q, e // n), n
