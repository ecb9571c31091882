"""torch.autograd front end of the propagator and its discrete adjoint.

`propagate(H, X, alpha, alpha_t, gamma, spec)` runs a batch of 2DES
propagations on the GPU (through libetoheom.so) and returns the readouts
P[b, t] (complex128, zero-padded to the longest instance).  Its backward runs
the hand-written adjoint kernels (checkpointed every `spec.checkpoint` steps)
and returns dL/dH, dL/dX, dL/dalpha, dL/dalpha_t, dL/dgamma in torch's
complex-gradient convention (dL/dRe + i dL/dIm), so any differentiable
construction of those tensors from physical parameters (site energies,
couplings, dipoles, ETO mode rates) chains through torch as usual.

`build_2des_system` is such a construction, written after the reference's
setup path: the lambda shift of the BATH block (param.cpp:189), construct_Hal's
one-exciton sort and double-exciton block (dynamics.cpp:24-49), the dipole
double columns (param.cpp:117-146) and compute_pulse_interaction
(polar.cpp:74-101).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

from . import MODE_2DES, Batch, Context, Problem

HBAR = 5308.0


@dataclass
class Spec:
    """Everything about a batch that is not differentiated."""
    ctx: Context
    L: int
    pulses: np.ndarray                 # (B, 3, 4): E0 [fs^-1], tau0, taup, w0 [fs^-1]
    t_start: np.ndarray                # (B,)
    t_end: np.ndarray                  # (B,)
    step_size: float
    print_step: float
    nv: list = field(default_factory=list)   # (B,) tuples
    checkpoint: int = 32
    site_offset: int = 1
    timing: dict = field(default_factory=dict)   # filled: GPU-event ms of the last forward / backward


class _Propagate(torch.autograd.Function):
    @staticmethod
    def forward(ctx, H, X, alpha, alpha_t, gamma, spec: Spec, want_grad: bool):
        B = H.shape[0]
        items = []
        for i in range(B):
            p = Problem(H=H[i].detach().cpu().numpy().astype(np.complex128),
                        X=X[i].detach().cpu().numpy().astype(np.complex128),
                        alpha=alpha[i].detach().cpu().numpy().astype(np.complex128),
                        alpha_t=alpha_t[i].detach().cpu().numpy().astype(np.complex128),
                        gamma=gamma[i].detach().cpu().numpy().astype(np.complex128),
                        L=spec.L, site_offset=spec.site_offset, pulses=np.asarray(spec.pulses[i]),
                        t_start=float(spec.t_start[i]), t_end=float(spec.t_end[i]),
                        step_size=spec.step_size, print_step=spec.print_step)
            nv = spec.nv[i] if spec.nv else (1, 1, 1)
            items.append({"problem": p, "H": p.H, "X": p.X, "nv": tuple(nv), "mode": MODE_2DES})
        # the adjoint's state (trajectory or checkpoints) only when a gradient is wanted
        need = bool(want_grad)
        b = Batch(spec.ctx, items, adjoint=spec.checkpoint if need else 0)
        b.run()
        b.sync()
        outs = [b.output(i) for i in range(B)]
        lens = [len(o) for o in outs]
        P = np.zeros((B, max(lens) if lens else 0), dtype=np.complex128)
        for i, o in enumerate(outs):
            P[i, : len(o)] = o
        ctx.batch = b if need else None
        ctx.lens = lens
        ctx.spec = spec
        spec.timing["forward_ms"] = b.elapsed_ms()
        spec.timing["forward_adjoint"] = need
        return torch.from_numpy(P)

    @staticmethod
    def backward(ctx, gP):
        b = ctx.batch
        if b is None:
            raise RuntimeError("propagate: the forward ran without gradients (grad mode off or no input "
                               "requiring grad), so there is no adjoint state")
        g = gP.detach().cpu().numpy().astype(np.complex128)
        b.backward([g[i, :n] for i, n in enumerate(ctx.lens)])
        ctx.spec.timing["backward_ms"] = b.elapsed_ms()
        grads = [b.grad(i) for i in range(len(ctx.lens))]
        ctx.batch = None   # release the device state (the trajectory) with the graph

        def stack(name):
            return torch.from_numpy(np.stack([gr[name] for gr in grads]))

        return stack("H"), stack("X"), stack("alpha"), stack("alpha_t"), stack("gamma"), None, None


def propagate(H, X, alpha, alpha_t, gamma, spec: Spec):
    """P[b, t] of a batch of 2DES propagations, differentiable in all five inputs.
    The adjoint's state is kept only when grad mode is on and an input requires grad."""
    want = torch.is_grad_enabled() and any(t.requires_grad for t in (H, X, alpha, alpha_t, gamma))
    return _Propagate.apply(H, X, alpha, alpha_t, gamma, spec, want)


def build_2des_system(eps_cm, J_cm, mu, lambda_line, polar):
    """H (N, N) and X (4, N, N) complex128 tensors of a 2DES system with n sites.

    eps_cm (n,) site energies and J_cm (n, n) symmetric couplings in cm^-1,
    mu (n, 3) transition dipoles, lambda_line the BATH shift (fs^-1), polar (3, 4)
    the lab-frame polarization matrix."""
    n = eps_cm.shape[0]
    ss = n + 1
    N = ss + (ss - 1) * (ss - 2) // 2
    dt = torch.float64
    Jm = J_cm.to(dt) / HBAR
    Hs = torch.zeros((ss, ss), dtype=dt)
    Hs[1:, 1:] = Jm - torch.diag(torch.diagonal(Jm))
    # lambda-shifted diagonal, then construct_Hal's ascending sort (slot 0 holds 0 and takes part)
    diag = torch.cat([torch.zeros(1, dtype=dt), eps_cm.to(dt) / HBAR + lambda_line])
    srt, _ = torch.sort(diag)
    Hs = Hs + torch.diag(torch.cat([torch.zeros(1, dtype=dt), srt[1:]]))
    pairs = [(a, b) for a in range(1, ss) for b in range(a, ss) if a != b]
    H = torch.zeros((N, N), dtype=dt)
    H[:ss, :ss] = Hs
    for i, (a, b) in enumerate(pairs):
        for j, (c, d) in enumerate(pairs):
            if i == j:
                H[ss + i, ss + j] = Hs[a, a] + Hs[b, b]
            else:
                v = 0.0
                if a == c and b != d:
                    v = v + Hs[b, d]
                if a == d and b != c:
                    v = v + Hs[b, c]
                if b == c and a != d:
                    v = v + Hs[a, d]
                if b == d and a != c:
                    v = v + Hs[a, c]
                if torch.is_tensor(v):
                    H[ss + i, ss + j] = v
    # dipoles: X_c[0][i] = mu_i[c]; X_c[i][(ab)] = d_bi X_c[0][a] + d_ai X_c[0][b]
    Xc = torch.zeros((3, N, N), dtype=dt)
    mu = mu.to(dt)
    for i in range(n):
        Xc[:, 0, i + 1] = mu[i]
    for j, (a, b) in enumerate(pairs):
        Xc[:, b, ss + j] = Xc[:, b, ss + j] + mu[a - 1]
        Xc[:, a, ss + j] = Xc[:, a, ss + j] + mu[b - 1]
    pol = torch.as_tensor(polar, dtype=dt)
    X = torch.einsum("cq,cij->qij", pol, Xc)
    return H.to(torch.complex128), X.to(torch.complex128)


def eto_modes(a, b, g, w):
    """alpha, alpha~, gamma (Km,) for ETO pairs (a/2, b/2, g, +-w) from tensors a, b, g, w (n_osc,)."""
    al = torch.complex(a / 2, b / 2)
    alt = torch.complex(a / 2, -b / 2)
    gp = torch.complex(g, w)
    gm = torch.complex(g, -w)
    alpha = torch.stack([al, al], dim=1).reshape(-1)
    alpha_t = torch.stack([alt, alt], dim=1).reshape(-1)
    gamma = torch.stack([gp, gm], dim=1).reshape(-1)
    return alpha, alpha_t, gamma
