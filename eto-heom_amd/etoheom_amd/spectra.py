"""2DES grids: the (tau, T) scan of the reference's job scripts, batched on the GPU,
and the differentiable grid loss used for parameter gradients.

* `scan_taus` restates script/IO-scripts/gen_2d_data.sh:4-61: for i in
  [-tau_bound, tau_bound] step tau_step and each population time T,
  i >= 0 (rephasing):      tau1 = t0,       tau2 = t0 + i,   tau3 = t0 + i + T
  i <  0 (non-rephasing):  tau1 = t0 + |i|, tau2 = t0,       tau3 = t0 + |i| + T
  with T_END = tau3 + propagate_time.
* `run_scan` runs every (tau, T) key of a template as one batch (3 propagations
  per point, the p111 - p110 - p101 combination of CPU_2DES.cpp:65-78) and
  `write_out_files` writes `out_{i}_{T}.out` in CPU_2DES's stdout format, the
  files gen_2d_spectrum.py reads (gen-2d-spectrum/gen_2d_spectrum.py:26-61).
* `grid_loss` is the (t1, t3) rephasing-grid L2 loss whose gradient w.r.t.
  site energies, couplings, dipoles and ETO parameters comes from the adjoint.
* `spectrum_2d` forms the 2D spectrum from the signals as
  gen-2d-spectrum/gen_2d_spectrum.py:26-66 does from the out files (before its
  smoothing, interpolation and plot), with torch.fft on the signals' device, and
  `spectral_loss` is the L2 loss on it; both are differentiable, so a spectral
  loss's gradient also comes from the adjoint.
"""
from __future__ import annotations

import os

import numpy as np
import torch

from . import MODE_2DES, Batch, Param
from . import autograd as AG

HBAR = 5308.0
NV3 = ((1, 1, 1), (1, 1, 0), (1, 0, 1))


def scan_taus(t0=210.0, propagate_time=600.0, tau_step=10, tau_bound=600, T=(0,)):
    """[(i, T, tau1, tau2, tau3, t_end)] in the order gen_2d_data.sh writes them."""
    out = []
    for t in T:
        for i in range(0, tau_bound + 1, tau_step):
            out.append((i, t, t0, t0 + i, t0 + i + t, t0 + i + t + propagate_time))
    for t in T:
        for i in range(-tau_bound, 0, tau_step):
            a = abs(i)
            out.append((i, t, t0 + a, t0, t0 + a + t, t0 + a + t + propagate_time))
    return out


def key_for(template: str, tau1, tau2, tau3, t_end) -> str:
    """gen_2d_data.sh's text substitution of TAU1/TAU2/TAU3/T_END."""
    def fmt(v):
        return str(int(v)) if float(v).is_integer() else repr(float(v))
    return (template.replace("TAU1", fmt(tau1)).replace("TAU2", fmt(tau2)).replace("TAU3", fmt(tau3))
            .replace("T_END", fmt(t_end)))


def run_scan(ctx, template: str, workdir: str, **scan):
    """All keys of the scan as one GPU batch; returns {(i, T): P (complex array)}."""
    os.makedirs(workdir, exist_ok=True)
    points = scan_taus(**scan)
    items, keys = [], []
    for (i, t, a, b, c, e) in points:
        path = os.path.join(workdir, f"key_{i}_{t}.key")
        with open(path, "w") as f:
            f.write(key_for(template, a, b, c, e))
        k = Param.from_key(path)
        H, X = k.samples()
        from . import Problem
        s = k.system()
        pul = np.array([[s.pulse[q][c2] for c2 in range(4)] for q in range(3)])
        for sidx in range(k.n_sample):
            p = Problem(H=H[sidx], X=X[sidx], alpha=k.get("alpha"), alpha_t=k.get("alpha_t"), gamma=k.get("gamma"),
                        L=k.L, pulses=pul, t_start=s.t_start, t_end=s.t_end, step_size=s.step_size,
                        print_step=s.print_step)
            for nv in NV3:
                items.append({"problem": p, "H": p.H, "X": p.X, "nv": nv, "mode": MODE_2DES})
        keys.append(((i, t), k.n_sample))
    b = Batch(ctx, items)
    b.run()
    b.sync()
    res, j = {}, 0
    for key, ns in keys:
        acc = None
        for _ in range(ns):
            p1, p2, p3 = b.output(j), b.output(j + 1), b.output(j + 2)
            v = (p1 - p2 - p3) / ns
            acc = v if acc is None else acc + v
            j += 3
        res[key] = acc
    return res


def write_out_files(results, outdir: str):
    """out_{i}_{T}.out in CPU_2DES's format: '%.6e %.6e' per readout."""
    os.makedirs(outdir, exist_ok=True)
    for (i, t), P in results.items():
        with open(os.path.join(outdir, f"out_{i}_{t}.out"), "w") as f:
            for v in P:
                f.write(f"{v.real:.6e} {v.imag:.6e}\n")


# ---------------------------------------------------------------- gradients

def trimer_params(seed=3, perturb=None):
    """Physical parameters of the trimer workload as leaf tensors:
    eps (3) and J (3 pairs) in cm^-1, mu (3, 3), ETO a, b, g, w (1 damped cosine),
    lam = reorganisation scale (1 = 35 cm^-1 Ohmic)."""
    from .configs import OHMIC_35_100_77 as F
    rng = np.random.default_rng(seed)
    eps = np.array([12000.0, 12120.0, 12260.0])
    J = np.array([90.0, 40.0, -70.0])
    mu = rng.normal(size=(3, 3))
    mu /= np.linalg.norm(mu, axis=1, keepdims=True)
    vals = dict(eps=eps, J=J, mu=mu, a=np.array([F["a"]]), b=np.array([F["b"]]), g=np.array([F["g"]]),
                w=np.array([F["w"] + 1e-3]), lam=np.array([1.0]))
    if perturb is not None:
        r = np.random.default_rng(perturb)
        vals["eps"] = vals["eps"] + r.uniform(-50, 50, 3)
        vals["lam"] = vals["lam"] * (1 + r.uniform(-0.2, 0.2))
    return {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in vals.items()}


def _system(params):
    n = params["eps"].shape[0]
    Jm = torch.zeros((n, n), dtype=torch.float64)
    iu = [(i, j) for i in range(n) for j in range(i + 1, n)]
    for k, (i, j) in enumerate(iu):
        Jm[i, j] = params["J"][k]
        Jm[j, i] = params["J"][k]
    lam_line = 35.0 / HBAR / np.pi * params["lam"][0]
    polar = np.array([[0.0] * 4, [0.0] * 4, [1.0] * 4])   # POLARIZATION 0 0 0 0, no rotation
    H, X = AG.build_2des_system(params["eps"], Jm, params["mu"], lam_line, polar)
    s = params["lam"]
    al, alt, ga = AG.eto_modes(params["a"] * s, params["b"] * s, params["g"], params["w"])
    K = n
    return H, X, al.repeat(K, 1), alt.repeat(K, 1), ga.repeat(K, 1)


def grid_spec(ctx, n_t1=16, t1_step=10.0, t2=100.0, n_t3=16, t3_step=10.0, L=5, dt=0.5, checkpoint=32,
              E0_cm=5.0, taup=15.0, w0_cm=12120.0, t0=40.0):
    """Spec of the rephasing (t1, t3) grid: 3 propagations per t1, readouts every t3_step.
    E0 defaults to the reference template's 5 cm^-1: the third-order signal left by
    p111 - p110 - p101 scales as E0^3 and must stay well above the FP64 rounding of
    the first-order terms it is the difference of."""
    pulses, t_end, nv = [], [], []
    for k in range(n_t1):
        t1 = k * t1_step
        taus = (t0, t0 + t1, t0 + t1 + t2)
        for v in NV3:
            pulses.append([[E0_cm / HBAR, tau, taup, w0_cm / HBAR] for tau in taus])
            t_end.append(taus[2] + t3_step * (n_t3 - 1))
            nv.append(v)
    B = len(pulses)
    return AG.Spec(ctx=ctx, L=L, pulses=np.array(pulses), t_start=np.zeros(B), t_end=np.array(t_end),
                   step_size=dt, print_step=t3_step, nv=nv, checkpoint=checkpoint)


def grid_signal(params, spec, propagate=None):
    """S[t1, t3] = p111 - p110 - p101 on the grid (complex tensor)."""
    H, X, al, alt, ga = _system(params)
    B = len(spec.nv)
    P = (propagate or AG.propagate)(H.expand(B, -1, -1), X.expand(B, -1, -1, -1), al.expand(B, -1, -1),
                                    alt.expand(B, -1, -1), ga.expand(B, -1, -1), spec)
    P = P.reshape(B // 3, 3, -1)
    return P[:, 0] - P[:, 1] - P[:, 2]


def grid_loss(S, target):
    return torch.sum(torch.abs(S - target) ** 2) / torch.sum(torch.abs(target) ** 2)


def spectrum_2d(rephasing, non_rephasing=None, size=(1000, 61)):
    """The 2D spectrum as gen-2d-spectrum/gen_2d_spectrum.py:26-66 forms it.

    Each argument holds the signals S[i, j] of one set of delays (i: the delay, in the
    ascending order the script reads the out files; j: the readout).  The script reads
    the signal of delay i into column i of a size[0] x size[1] grid, multiplied by i
    (D[j, i] = i S[i, j], zero past the signal's readouts and the last delay, readouts
    beyond size[0] dropped), and sums Re fftshift(fft2(D)) of the rephasing and the
    non-rephasing grid.  Complex128 tensors in, a float64 tensor out, on the signals'
    device (torch.fft: hipFFT on the GPU); differentiable."""
    def one(S):
        if S.dim() != 2 or S.shape[0] > size[1]:
            raise ValueError(f"spectrum_2d: signals {tuple(S.shape)} do not fit a {size[0]} x {size[1]} grid")
        m = min(S.shape[1], size[0])
        D = torch.zeros(size, dtype=torch.complex128, device=S.device)
        D[:m, : S.shape[0]] = 1j * S[:, :m].transpose(0, 1).to(torch.complex128)
        return torch.fft.fftshift(torch.fft.fft2(D)).real

    A = one(rephasing)
    return A if non_rephasing is None else A + one(non_rephasing.to(A.device))


def spectral_loss(A, target):
    """Relative L2 distance of two 2D spectra (real tensors)."""
    return torch.sum((A - target) ** 2) / torch.sum(target ** 2)


# ---------------------------------------------------------------- the 10-site chain grid

def chain_params(n_sites=10, seed=10, perturb=None):
    """Site energies eps (n, cm^-1) and nearest-neighbour couplings J (n - 1, cm^-1) of
    the chain workload as leaf tensors (the dipoles, bath and polarization stay fixed):
    eps ~ N(12500, 150) from `seed`, J = 80 cm^-1.  perturb: a seed for +/-50 cm^-1
    shifts of the energies and +/-10 % of the couplings (a target system)."""
    rng = np.random.default_rng(seed)
    eps = rng.normal(12500.0, 150.0, n_sites)
    J = np.full(n_sites - 1, 80.0)
    if perturb is not None:
        r = np.random.default_rng(perturb)
        eps = eps + r.uniform(-50, 50, n_sites)
        J = J * (1 + r.uniform(-0.1, 0.1, n_sites - 1))
    return {"eps": torch.tensor(eps, dtype=torch.float64, requires_grad=True),
            "J": torch.tensor(J, dtype=torch.float64, requires_grad=True)}


def _chain_system(params, seed=10):
    from .configs import OHMIC_35_100_77 as F
    n = params["eps"].shape[0]
    rng = np.random.default_rng(seed)
    rng.normal(12500.0, 150.0, n)   # the energies' draws, as configs.chain makes them
    mu = rng.normal(size=(n, 3))
    mu /= np.linalg.norm(mu, axis=1, keepdims=True)
    Jm = torch.zeros((n, n), dtype=torch.float64)
    for i in range(n - 1):
        Jm[i, i + 1] = params["J"][i]
        Jm[i + 1, i] = params["J"][i]
    polar = np.array([[0.0] * 4, [0.0] * 4, [1.0] * 4])
    H, X = AG.build_2des_system(params["eps"], Jm, torch.tensor(mu), F["lambda_line"], polar)
    t = lambda v: torch.tensor([v], dtype=torch.float64)   # noqa: E731
    al, alt, ga = AG.eto_modes(t(F["a"]), t(F["b"]), t(F["g"]), t(F["w"]))
    return H, X, al.repeat(n, 1), alt.repeat(n, 1), ga.repeat(n, 1)


def chain_grid_spec(ctx, t1s=(0.0, 10.0, 20.0, 30.0), t2s=(0.0, 100.0, 200.0, 400.0), nsteps=3000, L=3, dt=0.5,
                    checkpoint=32, E0_cm=1e-4, taup=15.0, w0_cm=12050.0, t0=40.0, max_batch_bytes=None):
    """Spec of propagations of the chain's (t1, t2) grid: one p111 propagation per point,
    `nsteps` steps from t = 0, readouts every dt after the third pulse."""
    pulses, nv = [], []
    for t2 in t2s:
        for t1 in t1s:
            taus = (t0, t0 + t1, t0 + t1 + t2)
            pulses.append([[E0_cm / HBAR, tau, taup, w0_cm / HBAR] for tau in taus])
            nv.append((1, 1, 1))
    B = len(pulses)
    return AG.Spec(ctx=ctx, L=L, pulses=np.array(pulses), t_start=np.zeros(B),
                   t_end=np.full(B, dt * (nsteps - 1)), step_size=dt, print_step=dt, nv=nv, checkpoint=checkpoint,
                   max_batch_bytes=max_batch_bytes)


def chain_grid_signal(params, spec, propagate=None):
    """P[b, t] of the chain grid's propagations (complex tensor)."""
    H, X, al, alt, ga = _chain_system(params)
    B = len(spec.nv)
    return (propagate or AG.propagate)(H.expand(B, -1, -1), X.expand(B, -1, -1, -1), al.expand(B, -1, -1),
                                       alt.expand(B, -1, -1), ga.expand(B, -1, -1), spec)
