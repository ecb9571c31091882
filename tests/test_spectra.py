"""The (tau, T) scan and output files of the 2D-spectrum tooling.

CPU: the delays of script/IO-scripts/gen_2d_data.sh:4-61 (rephasing i >= 0:
tau1 = t0, tau2 = t0 + i, tau3 = t0 + i + T; non-rephasing i < 0: tau1 = t0 + |i|,
tau2 = t0, tau3 = t0 + |i| + T; T_END = tau3 + propagate time, first loop first), the
TAU*/T_END substitution, and out_{i}_{T}.out files in CPU_2DES's '%.6e %.6e' format
under the names gen_2d_spectrum.py opens (out_{tau}_{T}.out for tau in
-tau_bound..tau_bound).  The 2D spectrum of gen_2d_spectrum.py:26-66 (spectrum_2d):
against its closed form for single signal values (which fixes the grid layout, the
factor i, the shift and the real part), against scipy's fft2 on grids read back from
out files the way the script reads them, and its gradient.  GPU: run_scan's batch (3
propagations per point, p111 - p110 - p101) equals the key driver on the same keys.
"""
import os

import numpy as np
import pytest

import etoheom_amd as E
from etoheom_amd import spectra as SP

HERE = os.path.dirname(os.path.abspath(__file__))


def test_scan_delays_follow_gen_2d_data():
    pts = SP.scan_taus()                         # t0 = 210, propagate 600, step 10, bound 600, T = (0,)
    assert len(pts) == 61 + 60
    assert pts[0] == (0, 0, 210, 210, 210, 810)
    assert pts[60] == (600, 0, 210, 810, 810, 1410)
    assert pts[61] == (-600, 0, 810, 210, 810, 1410)
    assert pts[-1] == (-10, 0, 220, 210, 220, 820)
    pts = SP.scan_taus(t0=100.0, propagate_time=50.0, tau_step=20, tau_bound=40, T=(0, 30))
    assert [p[:2] for p in pts] == [(0, 0), (20, 0), (40, 0), (0, 30), (20, 30), (40, 30),
                                    (-40, 0), (-20, 0), (-40, 30), (-20, 30)]
    for i, t, a, b, c, e in pts:
        assert c == max(a, b) + t and e == c + 50.0
        assert (a, b) == ((100.0, 100.0 + i) if i >= 0 else (100.0 - i, 100.0))


def test_key_substitution_and_parse(tmp_path):
    tmpl = open(os.path.join(HERE, "golden", "cpu_dimer_ohmic.key-tmpl")).read()
    txt = SP.key_for(tmpl, 210, 230.0, 330, 930.5)
    assert "TAU" not in txt and "T_END" not in txt
    path = tmp_path / "k.key"
    path.write_text(txt)
    k = E.Param.from_key(str(path), E.MODE_2DES)
    s = k.system()
    assert [s.pulse[q][1] for q in range(3)] == [210.0, 230.0, 330.0]
    assert s.t_end == pytest.approx(930.5)


def test_out_files_have_the_reference_format(tmp_path):
    rng = np.random.default_rng(0)
    res = {(0, 0): rng.normal(size=5) + 1j * rng.normal(size=5),
           (-10, 100): 1e-7 * (rng.normal(size=3) + 1j * rng.normal(size=3))}
    SP.write_out_files(res, str(tmp_path))
    assert sorted(os.listdir(tmp_path)) == ["out_-10_100.out", "out_0_0.out"]
    for (i, t), P in res.items():
        lines = open(tmp_path / f"out_{i}_{t}.out").read().splitlines()
        assert len(lines) == len(P)
        for line, v in zip(lines, P):
            assert line == f"{v.real:.6e} {v.imag:.6e}"
        d = np.loadtxt(tmp_path / f"out_{i}_{t}.out", ndmin=2)
        assert np.allclose(d[:, 0] + 1j * d[:, 1], P, rtol=1e-6, atol=0)


def test_spectrum_2d_closed_form():
    """One nonzero signal value c at (delay i0, readout j0) sits at D[j0, i0] = i c, so the
    spectrum is Re(i c exp(-2 pi i (k j0 / n0 + l i0 / n1))) at the shifted frequency
    indices (k, l) = (p - n0 // 2, q - n1 // 2); one value in each set of delays adds."""
    import torch
    n0, n1 = 10, 6
    Sr = torch.zeros((4, 7), dtype=torch.complex128)
    Snr = torch.zeros((5, 12), dtype=torch.complex128)   # readouts past n0 are dropped
    c1, c2 = 0.3 - 1.7j, -0.8 + 0.25j
    Sr[2, 3] = c1
    Snr[4, 9] = c2
    Snr[1, 11] = 5.0   # beyond the grid's rows: dropped
    A = SP.spectrum_2d(Sr, Snr, size=(n0, n1)).numpy()
    p, q = np.meshgrid(np.arange(n0) - n0 // 2, np.arange(n1) - n1 // 2, indexing="ij")
    want = (1j * c1 * np.exp(-2j * np.pi * (p * 3 / n0 + q * 2 / n1))).real + \
        (1j * c2 * np.exp(-2j * np.pi * (p * 9 / n0 + q * 4 / n1))).real
    assert A.shape == (n0, n1)
    assert np.max(np.abs(A - want)) < 1e-14
    with pytest.raises(ValueError):
        SP.spectrum_2d(torch.zeros((n1 + 1, 3), dtype=torch.complex128), size=(n0, n1))


def test_spectrum_2d_matches_fft2_of_the_files_grid(tmp_path):
    """Signals for tau = -600..600 fs (step 10) written as out files, read back column by
    column as gen_2d_spectrum.py:26-61 reads them (real and imaginary columns, times i,
    zero past the last readout, readouts past the grid's rows dropped), then scipy's
    fftshift(fft2(.)).real of the rephasing plus the non-rephasing grid."""
    import torch
    from scipy.fft import fft2, fftshift
    rng = np.random.default_rng(7)
    size = (100, 61)
    res = {}
    for tau in range(-600, 601, 10):
        n = 120 if tau % 20 else 80   # some signals longer, some shorter than the grid's rows
        res[(tau, 0)] = 1e-6 * (rng.normal(size=n) + 1j * rng.normal(size=n))
    SP.write_out_files(res, str(tmp_path))

    def read(tau):
        d = np.loadtxt(tmp_path / f"out_{tau}_0.out", ndmin=2)
        return d[:, 0] + 1j * d[:, 1]

    grids = {}
    for name, taus in (("nr", range(-600, 1, 10)), ("r", range(0, 601, 10))):
        g = np.zeros(size, complex)
        for i, tau in enumerate(taus):
            d = read(tau) * 1j
            for j in range(size[0]):
                g[j][i] = d[j] if j < len(d) else 0
        grids[name] = g
    want = fftshift(fft2(grids["r"])).real + fftshift(fft2(grids["nr"])).real

    def signals(taus):
        sig = [read(t) for t in taus]
        S = np.zeros((len(sig), max(len(x) for x in sig)), complex)
        for i, x in enumerate(sig):
            S[i, : len(x)] = x
        return torch.from_numpy(S)

    A = SP.spectrum_2d(signals(range(0, 601, 10)), signals(range(-600, 1, 10)), size=size).numpy()
    assert np.max(np.abs(A - want)) <= 1e-12 * np.max(np.abs(want))


def test_spectral_loss_gradient():
    """d(spectral loss)/dS through torch.fft equals central differences in the real and
    imaginary part of every signal value."""
    import torch
    rng = np.random.default_rng(3)
    S = torch.tensor(rng.normal(size=(3, 5)) + 1j * rng.normal(size=(3, 5)), requires_grad=True)
    T = SP.spectrum_2d(torch.tensor(rng.normal(size=(3, 5)) + 1j * rng.normal(size=(3, 5))), size=(6, 4))
    loss = SP.spectral_loss(SP.spectrum_2d(S, size=(6, 4)), T)
    loss.backward()
    h = 1e-6
    for i in range(3):
        for j in range(5):
            for part in (1.0, 1j):
                a, b = S.detach().clone(), S.detach().clone()
                a[i, j] += h * part
                b[i, j] -= h * part
                fd = (float(SP.spectral_loss(SP.spectrum_2d(a, size=(6, 4)), T)) -
                      float(SP.spectral_loss(SP.spectrum_2d(b, size=(6, 4)), T))) / (2 * h)
                # torch's gradient of a real loss in a complex input z = x + i y is dL/dx + i dL/dy
                g = S.grad[i, j]
                got = float(g.real) if part == 1.0 else float(g.imag)
                assert abs(got - fd) <= 1e-7 * max(1.0, abs(fd)), (i, j, part, got, fd)


@pytest.mark.gpu
def test_run_scan_matches_key_driver(tmp_path):
    ctx = E.Context(0)
    tmpl = open(os.path.join(HERE, "golden", "cpu_dimer_ohmic.key-tmpl")).read()
    scan = dict(t0=30.0, propagate_time=40.0, tau_step=10, tau_bound=20, T=(0, 20))
    res = SP.run_scan(ctx, tmpl, str(tmp_path / "keys"), **scan)
    assert len(res) == 2 * (3 + 2)
    for (i, t), P in res.items():
        ref = E.run_2des_key(ctx, str(tmp_path / "keys" / f"key_{i}_{t}.key"))
        assert np.max(np.abs(ref)) > 0
        assert np.max(np.abs(P - ref)) <= 1e-12 * np.max(np.abs(ref)), (i, t)
