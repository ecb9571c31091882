"""The torch.autograd path: gradient of a 2DES (t1, t3) grid loss w.r.t. physical
parameters (site energies, couplings, dipoles, ETO a/b/g/w, reorganisation scale),
adjoint on the GPU, checked against Richardson central FD of the same loss built
on the CPU oracle (relative <= 1e-6)."""
import numpy as np
import pytest
import torch

import etoheom_amd as E
import oracle as O
from etoheom_amd import spectra as SP

pytestmark = pytest.mark.gpu


def oracle_propagate(H, X, al, alt, ga, spec):
    B = H.shape[0]
    outs = []
    for i in range(B):
        q = O.Problem(H=H[i].detach().numpy(), X=X[i].detach().numpy(), alpha=al[i].detach().numpy(),
                      alpha_t=alt[i].detach().numpy(), gamma=ga[i].detach().numpy(), L=spec.L,
                      pulses=spec.pulses[i], t_start=float(spec.t_start[i]), t_end=float(spec.t_end[i]),
                      step_size=spec.step_size, print_step=spec.print_step)
        outs.append(O.twod(q, nv=spec.nv[i])[0])
    n = max(len(o) for o in outs)
    P = np.zeros((B, n), complex)
    for i, o in enumerate(outs):
        P[i, :len(o)] = o
    return torch.from_numpy(P)


def test_grid_gradient_matches_oracle_fd():
    ctx = E.Context(0)
    spec = SP.grid_spec(ctx, n_t1=2, t1_step=10.0, t2=20.0, n_t3=5, t3_step=5.0, L=2, checkpoint=16, taup=8.0,
                        E0_cm=200.0)
    with torch.no_grad():
        target = SP.grid_signal(SP.trimer_params(perturb=5), spec, propagate=oracle_propagate)
    params = SP.trimer_params()
    S = SP.grid_signal(params, spec)
    loss = SP.grid_loss(S, target)
    with torch.no_grad():
        S_cpu = SP.grid_signal(params, spec, propagate=oracle_propagate)
    # S is a difference of first-order signals ~1e3 times larger, hence 1e-9 rather than 1e-10
    assert torch.max(torch.abs(S - S_cpu)) / torch.max(torch.abs(S_cpu)) < 1e-9
    loss.backward()

    def L_cpu(over):
        p = {k: v.detach().clone() for k, v in params.items()}
        p.update(over)
        with torch.no_grad():
            return float(SP.grid_loss(SP.grid_signal(p, spec, propagate=oracle_propagate), target))

    # relative FD steps, Richardson-extrapolated: small where the loss curves on the
    # parameter's own scale (site energies multiply t in a phase), large for the ETO
    # frequency w ~ 1e-3 fs^-1 and amplitude a ~ 7e-5, where the loss varies on a scale
    # of 1/t ~ 1e-2 and |a| and smaller steps leave mostly rounding in the difference
    checks = [("eps", 0, 1e-4), ("J", 1, 1e-3), ("mu", (0, 2), 1e-3), ("g", 0, 1e-3), ("w", 0, 0.3),
              ("lam", 0, 1e-3), ("a", 0, 0.1)]
    for name, idx, rel in checks:
        base = params[name].detach().clone()
        h0 = rel * max(abs(float(base[idx])), 0.5 * float(base.abs().max()))

        def central(h):
            a, b = base.clone(), base.clone()
            a[idx] += h
            b[idx] -= h
            return (L_cpu({name: a}) - L_cpu({name: b})) / (2 * h)

        ref = (4 * central(h0 / 2) - central(h0)) / 3
        got = float(params[name].grad[idx])
        assert abs(got - ref) <= 1e-6 * abs(ref), (name, idx, got, ref)


def test_forward_without_gradient_keeps_no_adjoint_state():
    """Under torch.no_grad the propagation runs the plain forward (no trajectory or
    checkpoints): same signal as the gradient run's forward, and no backward."""
    ctx = E.Context(0)
    spec = SP.grid_spec(ctx, n_t1=2, t1_step=10.0, t2=20.0, n_t3=5, t3_step=5.0, L=2, checkpoint=16, taup=8.0,
                        E0_cm=200.0)
    params = SP.trimer_params()
    S_grad = SP.grid_signal(params, spec)
    assert spec.timing["forward_adjoint"] is True
    with torch.no_grad():
        S_plain = SP.grid_signal(params, spec)
    assert spec.timing["forward_adjoint"] is False
    assert torch.equal(S_grad.detach(), S_plain)
    frozen = {k: v.detach() for k, v in params.items()}
    S_frozen = SP.grid_signal(frozen, spec)                     # no input requires grad
    assert spec.timing["forward_adjoint"] is False and not S_frozen.requires_grad
