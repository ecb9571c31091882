"""Writes the long-horizon oracle fixtures tests/golden/*.npz.

Each fixture holds one propagation's inputs (H, X, bath, pulses, time grid) and
the CPU oracle's outputs for it (oracle/heom_oracle.c, which restates
etoheom-cpu/lib/dynamics.cpp:298-432): every readout P(t), and the final
hierarchy state (in full when small; otherwise rho_0, the first-tier ADOs and
the Frobenius norm of every ADO, which pins each ADO of the hierarchy).  The GPU
tests (tests/test_gpu_longrun.py) load the inputs from the fixture, so they do
not depend on the workload builders.

The oracle takes minutes to an hour per case on one core, so the fixtures are
made here once, not in the tests:

    python tests/golden/make_fixtures.py [name ...]
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "eto-heom_amd")]

import oracle as O  # noqa: E402


def cases():
    from etoheom_amd import configs
    out = {}
    # a trimer instance of the adjoint grid (N = 7, L = 5, 462 ADOs), its whole 2000-step run
    out["trimer_L5_2000"] = (configs.aggregate(3, 5), (1, 1, 1), "full")
    # a tetramer of the batched 2D grid (N = 11, L = 6, 3003 ADOs), static-disorder sample 0,
    # t1 = 50 fs, t2 = 100 fs, its whole 2000-step run
    out["tetramer_L6_2000"] = (configs.aggregate(4, 6, seed=2 + 4, sigma=100.0, taus=(40.0, 90.0, 190.0)),
                               (1, 1, 1), "tiers")
    # the heptamer benchmark system (N = 29, L = 4, 3060 ADOs) at its physical field, through
    # all three pulse windows (the last ends at 220 fs) and the first 220 readouts: 600 steps
    h = configs.heptamer()
    h.t_end = h.t_start + h.step_size * 599
    out["heptamer_L4_600"] = (h, (1, 1, 1), "tiers")
    # the 10-site chain of the benchmark grid (N = 56, L = 3, 1771 ADOs) with its full
    # hierarchy: the three pulses moved to 4 / 10 / 16 fs (3 fs wide, field x1e4, so that
    # every tier carries signal within the run), 60 steps: 45 with a pulse on (complex V)
    # and 15 without, 28 readouts
    c = configs.chain(nsteps=60)
    c.pulses[:, 1] = [4.0, 10.0, 16.0]
    c.pulses[:, 2] = 3.0
    c.pulses[:, 0] *= 1e4
    out["chain_L3_60"] = (c, (1, 1, 1), "tiers")
    return out


def oproblem(p):
    return O.Problem(H=p.H, X=p.X, alpha=p.alpha, alpha_t=p.alpha_t, gamma=p.gamma, L=p.L,
                     site_offset=p.site_offset, pulses=p.pulses, t_start=p.t_start, t_end=p.t_end,
                     step_size=p.step_size, print_step=p.print_step)


def make(name, p, nv, keep):
    t0 = time.time()
    P, rho = O.twod(oproblem(p), nv=nv, want_rho=True)
    d = p.alpha.size
    rec = {"H": p.H, "X": p.X, "alpha": p.alpha, "alpha_t": p.alpha_t, "gamma": p.gamma,
           "L": np.int64(p.L), "site_offset": np.int64(p.site_offset), "pulses": np.asarray(p.pulses),
           "time": np.array([p.t_start, p.t_end, p.step_size, p.print_step], dtype=np.float32),
           "nv": np.array(nv, dtype=np.int64), "P": P,
           "ado_norms": np.sqrt(np.sum(np.abs(rho) ** 2, axis=(1, 2)))}
    if keep == "full":
        rec["rho"] = rho
    else:
        # ADO 0 and the first tier (|n| = 1), with their indices in the DFS order
        ado = O.build_ado(d, p.L)
        idx = np.flatnonzero(ado.sum(axis=1) <= 1) if d else np.array([0])
        rec["tier01_index"] = idx.astype(np.int64)
        rec["rho_tier01"] = rho[idx]
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **rec)
    print(f"{name}: {len(P)} readouts, {rho.shape[0]} ADOs, {time.time() - t0:.0f} s", flush=True)


def main(argv):
    cs = cases()
    for name in (argv or list(cs)):
        p, nv, keep = cs[name]
        make(name, p, nv, keep)


if __name__ == "__main__":
    main(sys.argv[1:])
