"""Synthetic workloads: .key files in the reference format, and their realisation.

Each builder writes the reference's keyword file (etoheom-cpu/lib/param.cpp:18-248
reads it) and realises it the way CPU_2DES does for one sample
(CPU_2DES.cpp:50-60: construct_Hal with the DISORDER draws, then the random
lab-frame rotation of the polarizations), through the product library.

Workloads (all FP64, dt = 0.5 fs, RK4):
  dimer        ground + 2 singles + 1 double (N=4), 2 ETO modes/site (d=4), L=6 -> 210 ADOs, 2000 steps
  heptamer     ground + 7 singles + 21 doubles (N=29), d=14, L=4 -> 3060 ADOs, 2000 steps
  grid2d       512 propagations: dimer/trimer/tetramer (N=4/7/11), L in {3..6}, static disorder
               sigma=100 cm^-1, t1 in 0..310 fs x t2 in {0,100,200,400} fs x (rephasing,
               non-rephasing) x 2 disorder realisations
  trimer       ground + 3 singles + 3 doubles (N=7), d=6, L=5 -> 462 ADOs
  chain10      10 sites (N=56), nearest-neighbour J=80 cm^-1, d=20, L=3 -> 1771 ADOs, 3000 steps

The Ohmic bath J(w) = (lambda/wc) w exp(-w/wc), lambda=35 cm^-1, wc=100 cm^-1,
T=77 K, is represented by one damped cosine, the fit
etom.fit_ohmic(wc_cm=100, T=77, n_osc=1, eta=0.35) (two ETO modes, +/-w; w lies on
a flat direction of that fit, ~0); the fitted numbers are kept here so that
building a workload needs no quadrature.
"""
from __future__ import annotations

import os
import tempfile

import numpy as np

from . import MODE_2DES, Param, Problem

HBAR = 5308.0

# etom.fit_ohmic(wc_cm=100, T=77, n_osc=1, eta=0.35): C(t) ~ (a + i b) cos(wt) e^{-gt}, lambda = eta wc / pi
OHMIC_35_100_77 = {
    "lambda_line": 35.0 / HBAR / np.pi,
    "a": 6.736034787240873e-05,
    "b": -3.0590273829394615e-05,
    "g": 0.023177997347717187,
    "w": 8.645362342661063e-08,
}


def ohmic_modes(fit=OHMIC_35_100_77):
    a, b, g, w = fit["a"], fit["b"], fit["g"], fit["w"]
    return [(a / 2, b / 2, g, w), (a / 2, b / 2, g, -w)]


def key_text(single_size, K, L, H_cm, dipoles, modes, lambda_line, pulses, time, polarization=(0, 0, 0, 0),
             n_sample=1, seed=114, sigma_cm=None):
    """A 2DES .key file.  H_cm is single_size x single_size (row/col 0 = ground);
    dipoles are (i, j, x, y, z); pulses are (E0 cm^-1, tau0 fs, taup fs, w0 cm^-1).
    DISORDER goes last: the reference reads sys_size rows for it (param.cpp:61)."""
    out = ["SIZE", str(single_size), "", "HEOM", f"{K} {L}", "", "HAMILTONIAN"]
    for r in H_cm:
        out.append(" ".join(repr(float(v)) for v in r))
    out += ["", "BATH", repr(float(lambda_line)), str(len(modes))]
    out += [" ".join(repr(float(v)) for v in m) for m in modes]
    out += ["", "DIPOLE", str(len(dipoles))]
    out += [f"{int(i)} {int(j)} {float(x)!r} {float(y)!r} {float(z)!r}" for i, j, x, y, z in dipoles]
    out += ["", "POLARIZATION", " ".join(str(v) for v in polarization), "", "PULSE", "3"]
    out += [" ".join(repr(float(v)) for v in p) for p in pulses]
    out += ["TIME", " ".join(repr(float(v)) for v in time), "", "DISORDER", f"{n_sample} {seed}"]
    sig = np.zeros((single_size, single_size)) if sigma_cm is None else np.asarray(sigma_cm)
    for r in sig:
        out.append(" ".join(repr(float(v)) for v in r))
    return "\n".join(out) + "\n"


def realise(text, sample=0):
    """Parse a key text with the product parser and realise sample `sample`
    (construct_Hal + ranrot draws).  Returns (Problem, Param)."""
    fd, path = tempfile.mkstemp(suffix=".key")
    try:
        with os.fdopen(fd, "w") as f:
            f.write(text)
        k = Param.from_key(path, MODE_2DES)
        H, X = k.samples(max(k.n_sample, sample + 1))
        p = Problem(H=H[sample], X=X[sample], alpha=k.get("alpha"), alpha_t=k.get("alpha_t"),
                    gamma=k.get("gamma"), L=k.L, site_offset=1, pulses=_pulses(k),
                    t_start=float(k.system().t_start), t_end=float(k.system().t_end),
                    step_size=float(k.system().step_size), print_step=float(k.system().print_step))
        return p, k
    finally:
        os.unlink(path)


def _pulses(k):
    s = k.system()
    return np.array([[s.pulse[q][c] for c in range(4)] for q in range(3)])


def _three_pulses(tau, E0=1e-4, taup=15.0, w0=12050.0):
    return [(E0, t, taup, w0) for t in tau]


def _sites_H(eps, J):
    n = len(eps)
    H = np.zeros((n + 1, n + 1))
    H[1:, 1:] = J
    H[np.arange(1, n + 1), np.arange(1, n + 1)] = eps
    return H


def dimer(L=6, t1=50.0, t2=100.0, nsteps=2000, dt=0.5, seed=114):
    t0 = 40.0   # 2 taup + 10 fs
    H = _sites_H([12000.0, 12100.0], np.array([[0, 100.0], [100.0, 0]]))
    dip = [(0, 1, 1.0, 0.0, 0.0), (0, 2, 0.5, 0.866, 0.0)]
    txt = key_text(3, 2, L, H, dip, ohmic_modes(), OHMIC_35_100_77["lambda_line"],
                   _three_pulses([t0, t0 + t1, t0 + t1 + t2]), (0.0, dt * (nsteps - 1), dt, dt), seed=seed)
    return realise(txt)[0]


def heptamer(L=4, nsteps=2000, dt=0.5, seed=7):
    rng = np.random.default_rng(seed)
    n = 7
    eps = rng.normal(12400.0, 200.0, n)
    J = np.zeros((n, n))
    iu = np.triu_indices(n, 1)
    J[iu] = rng.uniform(-100.0, 100.0, len(iu[0]))
    J = J + J.T
    mu = rng.normal(size=(n, 3))
    mu /= np.linalg.norm(mu, axis=1, keepdims=True)
    dip = [(0, i + 1, *mu[i]) for i in range(n)]
    t0 = 40.0
    txt = key_text(n + 1, n, L, _sites_H(eps, J), dip, ohmic_modes(), OHMIC_35_100_77["lambda_line"],
                   _three_pulses([t0, t0 + 50.0, t0 + 150.0]), (0.0, dt * (nsteps - 1), dt, dt), seed=seed)
    return realise(txt)[0]


def chain(n_sites=10, L=3, nsteps=3000, dt=0.5, seed=10, t1=50.0, t2=100.0):
    rng = np.random.default_rng(seed)
    eps = rng.normal(12500.0, 150.0, n_sites)
    J = np.zeros((n_sites, n_sites))
    for i in range(n_sites - 1):
        J[i, i + 1] = J[i + 1, i] = 80.0
    mu = rng.normal(size=(n_sites, 3))
    mu /= np.linalg.norm(mu, axis=1, keepdims=True)
    dip = [(0, i + 1, *mu[i]) for i in range(n_sites)]
    t0 = 40.0
    txt = key_text(n_sites + 1, n_sites, L, _sites_H(eps, J), dip, ohmic_modes(), OHMIC_35_100_77["lambda_line"],
                   _three_pulses([t0, t0 + t1, t0 + t1 + t2]), (0.0, dt * (nsteps - 1), dt, dt), seed=seed)
    return realise(txt)[0]


def chain_grid(n_t1=64, t2s=(0.0, 100.0, 200.0, 400.0), nsteps=3000):
    """The chain's batched (t1, t2) grid: t1 = 0, 10, ... fs (n_t1 values) x t2s, one
    propagation each (64 x 4 = 256 for the full grid, about 91 GB of state)."""
    return [chain(t1=10.0 * i, t2=t2, nsteps=nsteps) for t2 in t2s for i in range(n_t1)]


def aggregate(n_sites, L, eps0=12000.0, J=100.0, seed=1, sigma=0.0, taus=(40.0, 90.0, 190.0), nsteps=2000,
              dt=0.5, sample=0):
    """n_sites in a ring-free chain with coupling J, optional static disorder sigma (cm^-1)."""
    rng = np.random.default_rng(seed)
    eps = eps0 + 100.0 * np.arange(n_sites)
    Jm = np.zeros((n_sites, n_sites))
    for i in range(n_sites - 1):
        Jm[i, i + 1] = Jm[i + 1, i] = J
    mu = rng.normal(size=(n_sites, 3))
    mu /= np.linalg.norm(mu, axis=1, keepdims=True)
    dip = [(0, i + 1, *mu[i]) for i in range(n_sites)]
    sig = np.zeros((n_sites + 1, n_sites + 1))
    sig[np.arange(1, n_sites + 1), np.arange(1, n_sites + 1)] = sigma
    txt = key_text(n_sites + 1, n_sites, L, _sites_H(eps, Jm), dip, ohmic_modes(), OHMIC_35_100_77["lambda_line"],
                   _three_pulses(list(taus)), (0.0, dt * (nsteps - 1), dt, dt), n_sample=sample + 1,
                   seed=seed, sigma_cm=sig)
    return realise(txt, sample)[0]


def grid2d(n=512, nsteps=2000, seed=2):
    """The batched 2DES grid: a list of (Problem, nv) items."""
    t0 = 40.0
    items = []
    k = 0
    for real in range(2):
        for t2 in (0.0, 100.0, 200.0, 400.0):
            for order in ("R", "NR"):
                for t1 in np.arange(0.0, 320.0, 10.0):
                    if k >= n:
                        break
                    n_sites = 2 + k % 3
                    L = 3 + (k // 3) % 4
                    taus = (t0, t0 + t1, t0 + t1 + t2) if order == "R" else (t0 + t1, t0, t0 + t1 + t2)
                    p = aggregate(n_sites, L, seed=seed + n_sites, sigma=100.0, taus=taus, nsteps=nsteps,
                                  sample=real)
                    items.append(p)
                    k += 1
    return items
