"""torch.autograd front end of the propagator and its discrete adjoint.

`propagate(H, X, alpha, alpha_t, gamma, spec)` runs a batch of 2DES
propagations on the GPU (through libetoheom.so) and returns the readouts
P[b, t] (complex128, zero-padded to the longest instance).  Its backward runs
the hand-written adjoint kernels (checkpointed every `spec.checkpoint` steps)
and returns dL/dH, dL/dX, dL/dalpha, dL/dalpha_t, dL/dgamma in torch's
complex-gradient convention (dL/dRe + i dL/dIm), so any differentiable
construction of those tensors from physical parameters (site energies,
couplings, dipoles, ETO mode rates) chains through torch as usual.

When the adjoint's device state (checkpoints and a segment's stage inputs, sized by
the host-side eh_batch_plan) of the whole batch exceeds the device memory
(`Spec.max_batch_bytes`, default 85 % of the free memory), the batch is split:
the forward runs as plain batches without adjoint state, and the backward runs each
sub-batch again with checkpoints and then backwards; every instance's gradient
depends only on its own propagation, so the sub-batches' gradients are the
single batch's (to the rounding of the accumulators' atomic adds).

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

from . import MODE_2DES, Batch, Context, EtoHeomError, Problem, batch_plan

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
    max_batch_bytes: int | None = None   # device bytes one batch may take (None: 85 % of the free memory)
    progress: object = None              # callable(str) told of each finished batch, or None
    timing: dict = field(default_factory=dict)   # filled: GPU-event ms of the last forward / backward


def _items(H, X, alpha, alpha_t, gamma, spec: Spec):
    """Batch items of the B propagations."""
    items = []
    for i in range(H.shape[0]):
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
    return items


def instance_bytes(H, X, alpha, alpha_t, gamma, spec: Spec, adjoint: bool = True):
    """Device bytes of each propagation alone (eh_batch_plan): its forward state plus,
    with adjoint, its checkpoint + replay state at spec.checkpoint."""
    out = []
    for it in _items(H, X, alpha, alpha_t, gamma, spec):
        pl = batch_plan([it], checkpoint=spec.checkpoint)
        out.append(pl["forward_bytes"] + (pl["adjoint_replay_bytes"] if adjoint else 0))
    return out


class _Propagate(torch.autograd.Function):
    @staticmethod
    def forward(ctx, H, X, alpha, alpha_t, gamma, spec: Spec, want_grad: bool):
        B = H.shape[0]
        items = _items(H, X, alpha, alpha_t, gamma, spec)
        # the adjoint's state (trajectory or checkpoints) only when a gradient is wanted
        need = bool(want_grad)
        budget = _budget(spec)
        groups = _groups(items, spec.checkpoint if need else 0, budget)
        outs = [None] * B
        ctx.batch, ctx.groups = None, None
        if need and len(groups) == 1:
            b = Batch(spec.ctx, items, adjoint=spec.checkpoint)
            if spec.progress:
                spec.progress(f"forward batch of {B} propagations with adjoint state: running")
            b.run()
            b.sync()
            outs = [b.output(i) for i in range(B)]
            ctx.batch = b
            spec.timing["forward_ms"] = b.elapsed_ms()
        else:
            # plain forward batches; with a gradient wanted, the backward re-runs `groups`
            fwd = _groups(items, 0, budget) if need else groups
            ms = 0.0
            for grp in fwd:
                b = Batch(spec.ctx, [items[i] for i in grp])
                if spec.progress:
                    spec.progress(f"forward batch of {len(grp)} propagations: running")
                b.run()
                b.sync()
                for k, i in enumerate(grp):
                    outs[i] = b.output(k)
                ms += b.elapsed_ms()
                b.close()
                if spec.progress:
                    spec.progress(f"forward batch of {len(grp)} propagations done")
            spec.timing["forward_ms"] = ms
            if need:
                ctx.groups, ctx.items = groups, items
        lens = [len(o) for o in outs]
        P = np.zeros((B, max(lens) if lens else 0), dtype=np.complex128)
        for i, o in enumerate(outs):
            P[i, : len(o)] = o
        ctx.lens = lens
        ctx.spec = spec
        ctx.need = need
        spec.timing["forward_adjoint"] = need and ctx.batch is not None
        spec.timing["sub_batches"] = len(groups) if need else 0
        return torch.from_numpy(P)

    @staticmethod
    def backward(ctx, gP):
        if not ctx.need:
            raise RuntimeError("propagate: the forward ran without gradients (grad mode off or no input "
                               "requiring grad), so there is no adjoint state")
        g = gP.detach().cpu().numpy().astype(np.complex128)
        b = ctx.batch
        if b is not None:
            if ctx.spec.progress:
                ctx.spec.progress(f"backward of {len(ctx.lens)} propagations: running")
            b.backward([g[i, :n] for i, n in enumerate(ctx.lens)])
            ctx.spec.timing["backward_ms"] = b.elapsed_ms()
            ctx.spec.timing["recompute_forward_ms"] = 0.0
            grads = [b.grad(i) for i in range(len(ctx.lens))]
            ctx.batch = None   # release the device state (the trajectory) with the graph
        else:
            # sub-batches: each one's forward again, with checkpoints, then its backward
            grads = [None] * len(ctx.lens)
            fms = bms = 0.0
            for grp in ctx.groups:
                sb = Batch(ctx.spec.ctx, [ctx.items[i] for i in grp], adjoint=ctx.spec.checkpoint)
                if ctx.spec.progress:
                    ctx.spec.progress(f"gradient sub-batch of {len(grp)} propagations: forward with checkpoints")
                sb.run()
                sb.sync()
                fms += sb.elapsed_ms()
                if ctx.spec.progress:
                    ctx.spec.progress(f"gradient sub-batch of {len(grp)} propagations: backward")
                sb.backward([g[i, : ctx.lens[i]] for i in grp])
                bms += sb.elapsed_ms()
                for k, i in enumerate(grp):
                    grads[i] = sb.grad(k)
                sb.close()
                if ctx.spec.progress:
                    ctx.spec.progress(f"gradient sub-batch of {len(grp)} propagations done")
            ctx.spec.timing["recompute_forward_ms"] = fms
            ctx.spec.timing["backward_ms"] = fms + bms
            ctx.groups = ctx.items = None

        def stack(name):
            return torch.from_numpy(np.stack([gr[name] for gr in grads]))

        return stack("H"), stack("X"), stack("alpha"), stack("alpha_t"), stack("gamma"), None, None


def _budget(spec: Spec) -> int:
    if spec.max_batch_bytes is not None:
        return int(spec.max_batch_bytes)
    free, _ = spec.ctx.mem_info()
    return int(0.85 * free)


def _groups(items, checkpoint: int, budget: int):
    """Consecutive instances packed into batches of at most `budget` device bytes
    (eh_batch_plan of each instance; with checkpoint > 0 including the adjoint's
    checkpoint + replay state)."""
    groups, cur, used = [], [], 0
    for i, it in enumerate(items):
        pl = batch_plan([it], checkpoint=max(1, checkpoint))
        need = pl["forward_bytes"] + (pl["adjoint_replay_bytes"] if checkpoint > 0 else 0)
        if need > budget:
            raise EtoHeomError(-6, f"propagate: instance {i} needs {need} device bytes, more than the "
                                   f"{budget} available to one batch")
        if cur and used + need > budget:
            groups.append(cur)
            cur, used = [], 0
        cur.append(i)
        used += need
    if cur:
        groups.append(cur)
    return groups


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
