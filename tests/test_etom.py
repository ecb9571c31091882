"""ETOM bath fitter (CPU).

What pins it:
* the Ohmic TCF quadrature against the closed form obtained from
  coth(x) = 1 + 2 sum_n e^{-2 n x}:  Re C = (1/pi) sum_n c_n (s_n^2 - t^2)/(s_n^2 + t^2)^2,
  s_n = 1/wc + n/kT (c_0 = 1, c_n = 2), and Im C = -(2/pi) wc^3 t / (1 + wc^2 t^2)^2;
* the Debye-Lorentz Matsubara closed form against Fourier-weighted quadrature of the
  Drude density J(w) = 2 lam gamma w / (w^2 + gamma^2);
* the parameter sets the reference's notebooks printed (etom/*_example.ipynb, cell 11):
  they must fit this module's TCF samples as well as they fit the reference's (a wrong
  normalisation, coth argument or sign would leave O(1) residuals);
* recovery of known damped cosines, the BATH block format and the key-file update
  read back through the product's .key parser.
"""
import warnings

import numpy as np
import pytest
from scipy.integrate import quad

import etoheom_amd as E
from etoheom_amd import configs
from etoheom_amd import etom as T

warnings.filterwarnings("ignore", category=Warning, module="scipy")


def test_ohmic_tcf_matches_closed_form():
    wc, kt = 100 / T.HBAR, T.kt_of_kelvin(77.0)
    t = np.array([0.0, 5.0, 40.0, 300.0, 900.0])
    cr, ci = T.tcf(T.ohmic_J(wc), t, kt)
    n = np.arange(1, 2_000_001)
    s0 = 1 / wc
    ref = []
    for tt in t:
        s = s0 + n / kt
        v = (s0 ** 2 - tt ** 2) / (s0 ** 2 + tt ** 2) ** 2 + 2 * np.sum((s ** 2 - tt ** 2) / (s ** 2 + tt ** 2) ** 2)
        v += 2 * kt ** 2 / (n[-1] + 0.5)          # tail of the 1/s^2 series
        ref.append(v / np.pi)
    ref = np.array(ref)
    im = -(2 / np.pi) * wc ** 3 * t / (1 + wc ** 2 * t ** 2) ** 2
    assert np.max(np.abs(cr - ref)) / np.max(np.abs(ref)) < 1e-6
    assert np.max(np.abs(ci - im)) / np.max(np.abs(im)) < 1e-6


def test_debye_lorentz_closed_form_matches_quadrature():
    lam, g = 60 / T.HBAR, 53.08 / T.HBAR
    kt = 0.5 * g
    t = np.array([20.0, 100.0, 400.0])

    def f(w):
        x = w / np.tanh(0.5 * w / kt) if w > 0 else 2 * kt
        return 2 * lam * g / (w ** 2 + g ** 2) * x

    cr = np.array([quad(f, 0, np.inf, weight="cos", wvar=x, limlst=200)[0] / np.pi for x in t])
    ci = np.array([-quad(lambda w: 2 * lam * g * w / (w ** 2 + g ** 2), 0, np.inf, weight="sin", wvar=x)[0] / np.pi
                   for x in t])
    dr, di = T.debye_lorentz_tcf(t, lam, g, 1 / kt, 20000)
    assert np.max(np.abs(cr - dr)) / np.max(np.abs(dr)) < 1e-6
    # the reference writes Im C = +gamma lam e^{-gamma t} for this bath: the opposite
    # sign of the quadrature convention used for the Ohmic densities
    assert np.max(np.abs(ci + di)) / np.max(np.abs(di)) < 1e-6


# cell 11 of the reference notebooks: rows (a/2, b/2, g, w), one per oscillator
GOLDEN = {
    "ohmic": [[0.58957759, 0.47316952, 0.03039521, -0.00000019],
              [-0.58933329, -0.47322380, 0.03038508, -0.00040011],
              [-0.00008124, 0.00006471, 0.14173306, -0.13313245]],
    "superohmic": [[0.02417457, -0.02340886, 0.05494828, 0.02810181],
                   [0.00047894, 0.00069827, 0.04931285, 0.07247539],
                   [-0.02431476, 0.02265864, 0.05595559, -0.02716711]],
    # the BATH block of the reference's dimer example template (example/CPU_dimer_ohmic), the
    # same Ohmic density (wc = 150 cm^-1, 77 K) fitted from another starting point
    "ohmic_example": [[0.00311213, -0.00252971, 0.12340683, 0.00000006],
                      [-0.00296194, 0.00259667, 0.12950563, 0.02683512],
                      [0.00001177, -0.00003282, 0.00754453, 0.0]],
    "debye": [[-0.00001620, 0.00005652, 0.01000000, 0.00000115],
              [-0.00000225, 0.00000000, 0.00036265, 0.00000000],
              [0.00000219, -0.00000000, 0.00036166, 0.00002456]],
}


def _samples(name, n_t=250):
    if name in ("ohmic", "ohmic_example"):
        t = np.linspace(0, 1000, n_t)
        return t, T.tcf(T.ohmic_J(150 / T.HBAR), t, T.kt_of_kelvin(77.0))
    if name == "superohmic":
        wc = 100 / T.HBAR
        t = np.linspace(0, 500, n_t)
        return t, T.tcf(T.superohmic_J(wc), t, 0.5 * wc)
    g, lam = 53.08 / T.HBAR, 60 / T.HBAR
    t = np.linspace(0, 1000, n_t)
    return t, T.debye_lorentz_tcf(t, lam, g, T.HBAR / (0.5 * g), 100)


def _rel_resid(t, y, amp, g, w):
    r = T.damped_cosines(t, amp, g, w) - y
    return float(np.sqrt(np.mean(r ** 2) / np.mean(y ** 2)))


# residuals of the printed parameter sets on the reference's own 1000-point grids
# (8-decimal rounding included) are ohmic 0.023 / 0.12, superohmic 0.020 / 0.18,
# Debye 1.0e-3 / 2.7e-5, the example template's set 0.0067 / 0.33 (real / imaginary);
# bounds leave 30 % for the coarser grid
BOUNDS = {"ohmic": (0.03, 0.16), "superohmic": (0.026, 0.23), "debye": (1.3e-3, 3.5e-5),
          "ohmic_example": (0.009, 0.43)}


@pytest.mark.parametrize("name", ["ohmic", "superohmic", "debye", "ohmic_example"])
def test_reference_fits_describe_our_tcf(name):
    t, (cr, ci) = _samples(name)
    G = np.array(GOLDEN[name])
    a, b, g, w = 2 * G[:, 0], 2 * G[:, 1], G[:, 2], G[:, 3]
    br, bi = BOUNDS[name]
    assert _rel_resid(t, cr, a, g, w) < br
    assert _rel_resid(t, ci, b, g, w) < bi


def test_superohmic_fit_lands_on_the_reference_minimum():
    f, lam = T.fit_superohmic()   # the reference grid: 1000 samples on [0, 500] fs
    G = np.array(GOLDEN["superohmic"])
    order = np.argsort(np.abs(f.w))
    gorder = np.argsort(np.abs(G[:, 3]))
    assert np.allclose(f.g[order], G[gorder, 2], rtol=2e-3)
    assert np.allclose(np.abs(f.w[order]), np.abs(G[gorder, 3]), rtol=2e-3)
    assert np.allclose(f.a[order] / 2, G[gorder, 0], rtol=1e-2, atol=1e-7)
    assert np.allclose(f.b[order] / 2, G[gorder, 1], rtol=1e-2, atol=1e-7)
    assert lam == pytest.approx(6 * (100 / T.HBAR) / np.pi)


def test_fit_recovers_known_damped_cosines():
    t = np.linspace(0, 600, 600)
    a, g, w, b = np.array([3e-4, -1e-4]), np.array([0.02, 0.05]), np.array([0.01, 0.03]), np.array([-2e-4, 5e-5])
    cr, ci = T.damped_cosines(t, a, g, w), T.damped_cosines(t, b, g, w)
    f = T.fit(t, cr, ci, 2, 0.1)
    # damped cosines are even in w; order the recovered oscillators by g
    o = np.argsort(f.g)
    assert np.allclose(f.g[o], g, rtol=1e-7)
    assert np.allclose(np.abs(f.w[o]), w, rtol=1e-7)
    assert np.allclose(f.a[o], a, rtol=1e-7)
    assert np.allclose(f.b[o], b, rtol=1e-7)


def test_ohmic_workload_bath_is_this_fit():
    """The bath of the synthetic workloads is fit_ohmic(wc=100 cm^-1, T=77 K, 1 oscillator,
    eta=0.35, i.e. lambda = 35 cm^-1 / pi); w sits on a flat direction of the fit (~0)."""
    f, lam = T.fit_ohmic(wc_cm=100.0, T=77.0, n_osc=1, eta=0.35)
    F = configs.OHMIC_35_100_77
    assert lam == pytest.approx(F["lambda_line"], rel=1e-12)
    assert f.eta * f.a[0] == pytest.approx(F["a"], rel=1e-3)
    assert f.eta * f.b[0] == pytest.approx(F["b"], rel=1e-3)
    assert f.g[0] == pytest.approx(F["g"], rel=1e-3)
    assert abs(f.w[0]) < 1e-6 and abs(F["w"]) < 1e-6


def test_bath_block_and_key_update_round_trip(tmp_path):
    f = T.EtoFit(a=np.array([2.4e-4, -1.0e-5]), b=np.array([-1.2e-4, 3e-6]), g=np.array([0.023, 0.1]),
                 w=np.array([0.004, 0.02]), eta=0.5)
    lam = 0.0021
    blk = T.bath_block(f, lam)
    lines = blk.splitlines()
    assert lines[0] == "0.0021" and lines[1] == "4"
    assert lines[2] == f"{0.5 * 2.4e-4 / 2:.8f} {0.5 * -1.2e-4 / 2:.8f} {0.023:.8f} {0.004:.8f}"
    assert lines[3].endswith(f"{-0.004:.8f}")
    # a template whose BATH block is a placeholder, then the fitted block read back
    H = np.array([[0, 0, 0], [0, 12000.0, 100.0], [0, 100.0, 12100.0]])
    txt = configs.key_text(3, 2, 2, H, [(0, 1, 1.0, 0.0, 0.0), (0, 2, 0.0, 1.0, 0.0)], [(1.0, 0.0, 0.1, 0.0)],
                           0.0, [(1e-4, 40.0, 15.0, 12050.0)] * 3, (0.0, 10.0, 0.5, 0.5))
    p = tmp_path / "key.key-tmpl"
    p.write_text(txt)
    T.update_key_template(str(p), f, lam)
    k = E.Param.from_key(str(p), E.MODE_2DES)
    al, alt, ga = f.alpha()
    modes = f.modes()
    # every site gets the same modes, in file order, alpha = a/2 + i b/2 (8-decimal text)
    assert np.allclose(k.get("alpha")[0], [complex(round(A, 8), round(B, 8)) for A, B, _, _ in modes], atol=1e-12)
    assert np.allclose(k.get("alpha_t")[1], [complex(round(A, 8), -round(B, 8)) for A, B, _, _ in modes], atol=1e-12)
    assert np.allclose(k.get("gamma")[1], [complex(round(g, 8), round(w, 8)) for _, _, g, w in modes], atol=1e-12)
    assert np.allclose(al, np.conj(alt))
    bad = tmp_path / "bad.key"
    bad.write_text("DIPOLE\n0\nBATH\n")
    with pytest.raises(ValueError):
        T.update_key_template(str(bad), f, lam)
