"""ETOM bath fitter: the BATH block of a .key file from a spectral density.

The propagator's bath is a set of effective thermal oscillators (ETOs): per site,
modes k with alpha_k = a_k/2 + i b_k/2, alpha~_k = a_k/2 - i b_k/2 and
gamma_k = g_k + i w_k, in (+w, -w) pairs.  They come from a fit of the bath
time-correlation function (TCF)

    C(t) = (1/pi) int_0^inf J(w) [coth(w / (2 kT)) cos(w t) - i sin(w t)] dw

by damped cosines, as etom/ohmic_ETOM.py:19-95 does (superohmic_ETOM.py and
debye_lorentz_ETOM.py differ only in J(w), kT and the time grid):

    Re C(t) ~ sum_j a_j cos(w_j t) exp(-g_j t)     all 3N parameters, least squares
    Im C(t) ~ sum_j b_j cos(w_j t) exp(-g_j t)     b_j only, (g_j, w_j) kept

Both fits are Levenberg-Marquardt least squares (scipy.optimize.curve_fit, as the
reference uses) from the reference's starting point: every parameter equal to
`init_guess`.  The BATH block (ohmic_ETOM.py:172-191) is

    <lambda>            reorganisation line as the reference writes it
    <2N>
    a_j/2 b_j/2 g_j  w_j     (8 decimals)
    a_j/2 b_j/2 g_j -w_j

Units: t in fs, w, g, kT in fs^-1, C and a, b in fs^-2 (cm^-1 divide by 5308).
All of this is host-side set-up run once per bath; nothing here is on the
propagation path.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy.integrate import quad
from scipy.optimize import curve_fit

HBAR = 5308.0          # cm^-1 fs
KB = 0.695             # cm^-1 / K


# ----------------------------------------------------------- spectral densities

def ohmic_J(wc):
    """J(w) = w exp(-w/wc) (unit coupling; ohmic_ETOM.py:19-23)."""
    return lambda w: w * np.exp(-w / wc)


def superohmic_J(wc):
    """J(w) = w^3/wc^2 exp(-w/wc) (superohmic_ETOM.py:19-23)."""
    return lambda w: w ** 3 / wc ** 2 * np.exp(-w / wc)


def kt_of_kelvin(T):
    """kT in fs^-1 for a temperature in K (the ohmic fitter's beta*hbar, ohmic_ETOM.py:10,26)."""
    return T * KB / HBAR


# ------------------------------------------------------------------------ TCF

def tcf(J, t, kt, epsabs=1e-10, epsrel=1e-10, limit=10_000_000):
    """C(t) for each t by adaptive quadrature over [0, inf) (ohmic_ETOM.py:25-37).
    Returns (Re C, Im C) arrays."""
    t = np.atleast_1d(np.asarray(t, dtype=np.float64))

    def re(w, tt):
        return J(w) * np.cos(w * tt) / np.tanh(0.5 * w / kt)

    def im(w, tt):
        return -J(w) * np.sin(w * tt)

    cr = np.array([quad(re, 0, np.inf, args=(tt,), epsabs=epsabs, epsrel=epsrel, limit=limit)[0] for tt in t])
    ci = np.array([quad(im, 0, np.inf, args=(tt,), epsabs=epsabs, epsrel=epsrel, limit=limit)[0] for tt in t])
    return cr / np.pi, ci / np.pi


def debye_lorentz_tcf(t, lam, gamma, beta_hbar, n_matsubara=100):
    """Debye-Lorentz TCF in closed form with n Matsubara terms, as
    debye_lorentz_ETOM.py:23-45 evaluates it:
      Re C = gamma lam cot(beta_hbar gamma / 2) e^{-gamma t}
             + sum_k 4 lam gamma / beta_hbar * nu_k / (nu_k^2 - gamma^2) e^{-nu_k t},
      nu_k = 2 pi k / beta_hbar,
      Im C = + gamma lam e^{-gamma t}   (the reference's sign for this bath).
    beta_hbar is the product the reference forms, (1/kT) * 5308 for kT in fs^-1."""
    t = np.atleast_1d(np.asarray(t, dtype=np.float64))
    re = gamma * lam / np.tan(beta_hbar * gamma / 2) * np.exp(-gamma * t)
    for k in range(1, n_matsubara + 1):
        nu = 2 * np.pi * k / beta_hbar
        re = re + 4 * lam * gamma / beta_hbar * nu / (nu ** 2 - gamma ** 2) * np.exp(-nu * t)
    return re, gamma * lam * np.exp(-gamma * t)


# ------------------------------------------------------------------------ fit

def damped_cosines(t, a, g, w):
    """sum_j a_j cos(w_j t) exp(-g_j t), summed in j order."""
    out = 0.0
    for j in range(len(a)):
        out = out + a[j] * np.cos(w[j] * t) * np.exp(-g[j] * t)
    return out


@dataclass
class EtoFit:
    a: np.ndarray      # (N,) real-part amplitudes
    b: np.ndarray      # (N,) imaginary-part amplitudes
    g: np.ndarray      # (N,) damping rates, fs^-1
    w: np.ndarray      # (N,) frequencies, fs^-1
    eta: float = 1.0   # coupling strength the amplitudes are scaled by on output

    def modes(self):
        """[(a/2, b/2, g, w), (a/2, b/2, g, -w)] per oscillator: the BATH mode lines."""
        out = []
        for j in range(len(self.a)):
            A, B = self.eta * self.a[j] / 2, self.eta * self.b[j] / 2
            out.append((A, B, self.g[j], self.w[j]))
            out.append((A, B, self.g[j], -self.w[j]))
        return out

    def alpha(self):
        """alpha, alpha~, gamma (2N,) complex per site in the propagator's convention."""
        m = self.modes()
        al = np.array([complex(A, B) for A, B, _, _ in m])
        return al, np.conj(al), np.array([complex(g, w) for _, _, g, w in m])

    def correlation(self, t):
        """The fitted C(t) (complex, scaled by eta)."""
        t = np.asarray(t, dtype=np.float64)
        return self.eta * (damped_cosines(t, self.a, self.g, self.w) + 1j * damped_cosines(t, self.b, self.g, self.w))


def fit(t, c_re, c_im, n_osc, init_guess, eta=1.0, maxfev=1_000_000_000, imag_bounds=None):
    """Two-stage ETO fit of a sampled TCF (ohmic_ETOM.py:74-95).  The real part
    fixes (a, g, w); the imaginary part then fits b with (g, w) held."""
    t = np.asarray(t, dtype=np.float64)
    N = int(n_osc)

    def f_re(tt, *p):
        p = np.asarray(p)
        return damped_cosines(tt, p[0::3], p[1::3], p[2::3])

    with np.errstate(over="ignore", invalid="ignore"):
        pr, _ = curve_fit(f_re, t, c_re, p0=[init_guess] * (3 * N), maxfev=maxfev)
        g, w = pr[1::3].copy(), pr[2::3].copy()

        def f_im(tt, *p):
            return damped_cosines(tt, np.asarray(p), g, w)

        kw = {} if imag_bounds is None else {"bounds": imag_bounds}
        pi, _ = curve_fit(f_im, t, c_im, p0=[init_guess] * N, maxfev=maxfev, **kw)
    return EtoFit(a=pr[0::3].copy(), b=np.asarray(pi, dtype=np.float64), g=g, w=w, eta=float(eta))


# ----------------------------------------------------- the reference's three baths

def fit_ohmic(wc_cm=150.0, T=77.0, n_osc=3, eta=1.0, init_guess=0.1, n_t=1000, t_max=1000.0):
    """Ohmic bath J(w) = eta w exp(-w/wc) at temperature T (ohmic_ETOM.py defaults).
    Returns (fit, lambda line): lambda = eta wc / pi (fs^-1)."""
    wc = wc_cm / HBAR
    t = np.linspace(0, t_max, n_t)
    cr, ci = tcf(ohmic_J(wc), t, kt_of_kelvin(T))
    return fit(t, cr, ci, n_osc, init_guess, eta), eta * wc / np.pi


def fit_superohmic(wc_cm=100.0, kt_over_wc=0.5, n_osc=3, eta=1.0, init_guess=0.01, n_t=1000, t_max=500.0):
    """Super-Ohmic J(w) = eta w^3/wc^2 exp(-w/wc), kT = kt_over_wc * wc
    (superohmic_ETOM.py defaults).  The lambda line is written as 6 eta wc / pi,
    the value the reference's template update writes."""
    wc = wc_cm / HBAR
    t = np.linspace(0, t_max, n_t)
    cr, ci = tcf(superohmic_J(wc), t, kt_over_wc * wc)
    return fit(t, cr, ci, n_osc, init_guess, eta, imag_bounds=(-np.inf, np.inf)), 6 * eta * wc / np.pi


def fit_debye_lorentz(gamma_cm=53.08, lam_cm=60.0, kt_over_gamma=0.5, n_osc=3, n_matsubara=100,
                      init_guess=0.001, n_t=1000, t_max=1000.0):
    """Debye-Lorentz bath in closed form (debye_lorentz_ETOM.py defaults);
    the lambda line is lam (fs^-1)."""
    gamma, lam = gamma_cm / HBAR, lam_cm / HBAR
    kt = kt_over_gamma * gamma
    t = np.linspace(0, t_max, n_t)
    cr, ci = debye_lorentz_tcf(t, lam, gamma, HBAR / kt, n_matsubara)
    return fit(t, cr, ci, n_osc, init_guess), lam


# ------------------------------------------------------------------ key files

def bath_block(fit_result: EtoFit, lam) -> str:
    """The lines between BATH and DIPOLE, in the reference's format."""
    lines = [f"{lam}", f"{2 * len(fit_result.a)}"]
    lines += [f"{A:.8f} {B:.8f} {g:.8f} {w:.8f}" for A, B, g, w in fit_result.modes()]
    return "\n".join(lines) + "\n"


def update_key_template(path, fit_result: EtoFit, lam):
    """Replace everything between the BATH and DIPOLE keyword lines of a key
    template with the fitted block and a blank line (ohmic_ETOM.py:193-229).
    Raises ValueError when BATH does not come before DIPOLE."""
    with open(path) as f:
        lines = f.readlines()
    bath = next((i for i, l in enumerate(lines) if l.strip() == "BATH"), None)
    dip = next((i for i, l in enumerate(lines) if l.strip() == "DIPOLE"), None)
    if bath is None or dip is None or bath >= dip:
        raise ValueError(f"{path}: BATH must come before DIPOLE")
    out = lines[: bath + 1] + [bath_block(fit_result, lam), "\n"] + lines[dip:]
    with open(path, "w") as f:
        f.writelines(out)
