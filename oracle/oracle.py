"""ctypes front end of the CPU oracle (oracle/heom_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, and nowhere else.  The product library never
routes through this module.

All arrays are numpy complex128, row-major, in the reference layout: a matrix
is N*N, a hierarchy state is (M, N, N) in the reference DFS ADO order
(etoheom-cpu/lib/utilize.cpp:3-18).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")


def build() -> str:
    """Compile liboracle.so with the Makefile beside this file."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB


class _Problem(C.Structure):
    _fields_ = [
        ("N", C.c_int), ("K", C.c_int), ("Km", C.c_int), ("L", C.c_int),
        ("site_offset", C.c_int), ("pad_", C.c_int),
        ("H", C.c_void_p), ("X", C.c_void_p),
        ("alpha", C.c_void_p), ("alpha_t", C.c_void_p), ("gamma", C.c_void_p),
        ("pulse", (C.c_double * 4) * 3),
        ("t_start", C.c_float), ("t_end", C.c_float),
        ("step_size", C.c_float), ("print_step", C.c_float),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        l = C.CDLL(_LIB)
        l.oh_num_ado.restype = C.c_long
        l.oh_num_ado.argtypes = [C.c_int, C.c_int]
        l.oh_build_ado.restype = C.c_long
        l.oh_build_ado.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_long]
        l.oh_rhs.argtypes = [C.POINTER(_Problem), C.c_void_p, C.c_void_p, C.c_void_p]
        l.oh_field.argtypes = [C.POINTER(_Problem), C.c_float, C.c_int, C.c_int, C.c_int, C.c_void_p]
        l.oh_2des.argtypes = [C.POINTER(_Problem), C.c_int, C.c_int, C.c_int, C.c_void_p,
                              C.c_long, C.c_void_p, C.c_long, C.POINTER(C.c_long), C.c_void_p]
        l.oh_dynamics.argtypes = [C.POINTER(_Problem), C.c_void_p, C.c_int, C.c_void_p,
                                  C.c_long, C.c_void_p, C.c_long, C.POINTER(C.c_long), C.c_void_p]
        _lib = l
    return _lib


@dataclass
class Problem:
    """Plain-array HEOM problem, the data the reference keeps in `param`
    (etoheom-cpu/lib/param.h:38-82).  alpha/alpha_t/gamma are (K, Km) complex;
    pulses is (3, 4) rows of (E0 [fs^-1], tau0 [fs], taup [fs], w0 [fs^-1])."""
    H: np.ndarray
    alpha: np.ndarray
    alpha_t: np.ndarray
    gamma: np.ndarray
    L: int
    X: np.ndarray | None = None
    site_offset: int = 1
    pulses: np.ndarray = field(default_factory=lambda: np.zeros((3, 4)))
    t_start: float = 0.0
    t_end: float = 0.0
    step_size: float = 1.0
    print_step: float = 1.0

    @property
    def N(self) -> int:
        return self.H.shape[0]

    @property
    def K(self) -> int:
        return self.alpha.shape[0]

    @property
    def Km(self) -> int:
        return self.alpha.shape[1]

    def _c(self):
        keep = []

        def ptr(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.complex128)
            keep.append(a)
            return a.ctypes.data

        s = _Problem()
        s.N, s.K, s.Km, s.L = self.N, self.K, self.Km, self.L
        s.site_offset = self.site_offset
        s.H = ptr(self.H)
        s.X = ptr(self.X)
        s.alpha = ptr(self.alpha)
        s.alpha_t = ptr(self.alpha_t)
        s.gamma = ptr(self.gamma)
        pul = np.asarray(self.pulses, dtype=np.float64).reshape(3, 4)
        for q in range(3):
            for c in range(4):
                s.pulse[q][c] = float(pul[q, c])
        s.t_start, s.t_end = self.t_start, self.t_end
        s.step_size, s.print_step = self.step_size, self.print_step
        return s, keep

    def num_ado(self) -> int:
        return num_ado(self.K * self.Km, self.L)


def num_ado(d: int, L: int) -> int:
    return int(lib().oh_num_ado(d, L))


def build_ado(d: int, L: int) -> np.ndarray:
    M = num_ado(d, L)
    out = np.zeros((M, max(d, 1)), dtype=np.uint8)
    got = lib().oh_build_ado(d, L, out.ctypes.data, M)
    assert got == M
    return out[:, :d]


def rhs(p: Problem, V: np.ndarray, rho: np.ndarray) -> np.ndarray:
    s, keep = p._c()
    V = np.ascontiguousarray(V, dtype=np.complex128)
    rho = np.ascontiguousarray(rho, dtype=np.complex128)
    out = np.zeros_like(rho)
    rc = lib().oh_rhs(C.byref(s), V.ctypes.data, rho.ctypes.data, out.ctypes.data)
    if rc:
        raise RuntimeError(f"oh_rhs failed: {rc}")
    return out


def field_V(p: Problem, t: float, nv=(1, 1, 1)) -> np.ndarray:
    s, keep = p._c()
    V = np.zeros((p.N, p.N), dtype=np.complex128)
    lib().oh_field(C.byref(s), C.c_float(t), nv[0], nv[1], nv[2], V.ctypes.data)
    return V


def _nsteps(p: Problem) -> int:
    nstep = np.float32(np.float32(p.t_end) - np.float32(p.t_start)) / np.float32(p.step_size)
    return int(np.floor(np.float32(nstep))) + 1


def twod(p: Problem, nv=(1, 1, 1), rho_init=None, max_steps: int = -1, want_rho: bool = False):
    """twoD_spectrum_solver: returns (P[n_points] complex, final state or None)."""
    s, keep = p._c()
    cap = _nsteps(p) + 1
    P = np.zeros(cap, dtype=np.complex128)
    n = C.c_long(0)
    M = p.num_ado()
    rho_out = np.zeros((M, p.N, p.N), dtype=np.complex128) if want_rho else None
    ri = None
    if rho_init is not None:
        ri = np.ascontiguousarray(rho_init, dtype=np.complex128)
    rc = lib().oh_2des(C.byref(s), nv[0], nv[1], nv[2],
                       ri.ctypes.data if ri is not None else None, max_steps,
                       P.ctypes.data, cap, C.byref(n),
                       rho_out.ctypes.data if rho_out is not None else None)
    if rc:
        raise RuntimeError(f"oh_2des failed: {rc}")
    return P[: n.value].copy(), rho_out


def dynamics(p: Problem, sites=(1, 2), rho_init=None, max_steps: int = -1, want_rho: bool = False):
    """dynamics_solver: returns (pop[len(sites), n_points], final state or None)."""
    s, keep = p._c()
    cap = _nsteps(p) + 1
    sites_a = np.ascontiguousarray(sites, dtype=np.int32)
    pop = np.zeros((len(sites_a), cap), dtype=np.float64)
    n = C.c_long(0)
    M = p.num_ado()
    rho_out = np.zeros((M, p.N, p.N), dtype=np.complex128) if want_rho else None
    ri = None
    if rho_init is not None:
        ri = np.ascontiguousarray(rho_init, dtype=np.complex128)
    rc = lib().oh_dynamics(C.byref(s), sites_a.ctypes.data, len(sites_a),
                           ri.ctypes.data if ri is not None else None, max_steps,
                           pop.ctypes.data, cap, C.byref(n),
                           rho_out.ctypes.data if rho_out is not None else None)
    if rc:
        raise RuntimeError(f"oh_dynamics failed: {rc}")
    return pop[:, : n.value].copy(), rho_out
