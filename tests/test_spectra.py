"""The (tau, T) scan and output files of the 2D-spectrum tooling.

CPU: the delays of script/IO-scripts/gen_2d_data.sh:4-61 (rephasing i >= 0:
tau1 = t0, tau2 = t0 + i, tau3 = t0 + i + T; non-rephasing i < 0: tau1 = t0 + |i|,
tau2 = t0, tau3 = t0 + |i| + T; T_END = tau3 + propagate time, first loop first), the
TAU*/T_END substitution, and out_{i}_{T}.out files in CPU_2DES's '%.6e %.6e' format
under the names gen_2d_spectrum.py opens (out_{tau}_{T}.out for tau in
-tau_bound..tau_bound).  GPU: run_scan's batch (3 propagations per point, p111 -
p110 - p101) equals the key driver on the same keys.
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
