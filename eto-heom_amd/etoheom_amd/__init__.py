"""Python front end of libetoheom.so, the MI355X-native ETO-HEOM propagator.

Mirrors the reference's host interface for the propagation path:

* ``Param.from_key(path, mode)``          param::param_2des / param_dynamics (etoheom-cpu/lib/param.cpp:18-348)
* ``Param.construct_hal(dH)``             construct_Hal (etoheom-cpu/lib/dynamics.cpp:3-51)
* ``Param.set_polar(polar)``              compute_pulse_interaction (etoheom-cpu/lib/polar.cpp:74-101)
* ``twoD_spectrum_solver(ctx, key, H, nv1, nv2, nv3)``   etoheom-cpu/lib/dynamics.cpp:298-432
* ``dynamics_solver(ctx, key, H, sites)``                etoheom-cpu/lib/dynamics.cpp:202-295
* ``run_2des_key(ctx, path)``             main of etoheom-cpu/src/CPU_2DES.cpp
* ``Batch``                               many independent propagations in one graph-captured run

Everything computes on the GPU through the C ABI of include/etoheom.h; there is
no CPU fallback: loading fails loudly when libetoheom.so is missing, and the
propagation calls raise when no HIP device is present.  Errors come back as
``EtoHeomError`` carrying the library's code and message instead of the
reference's exit().
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libetoheom.so")

MODE_2DES = 0
MODE_DYNAMICS = 1

ERRORS = {0: "EH_OK", -1: "EH_EINVAL", -2: "EH_EPARSE", -3: "EH_ESHAPE", -4: "EH_ENAN",
          -5: "EH_EHIP", -6: "EH_ENOMEM", -7: "EH_ERANGE", -8: "EH_EPULSE"}

# every entry point of include/etoheom.h
EXPORTS = [
    "eh_last_error", "eh_problem_from_key", "eh_problem_destroy", "eh_problem_dims", "eh_problem_system",
    "eh_problem_get", "eh_problem_polar_default", "eh_problem_set_polar", "eh_polar_rotate",
    "eh_construct_hal", "eh_problem_samples", "eh_rng_stream", "eh_hierarchy_size", "eh_hierarchy_build",
    "eh_ctx_create", "eh_ctx_destroy", "eh_ctx_info", "eh_2des_solve", "eh_2des_solve_noalloc", "eh_dynamics_solve",
    "eh_2des_key_run", "eh_dynamics_key_run", "eh_batch_create", "eh_batch_destroy", "eh_batch_run",
    "eh_batch_sync", "eh_batch_elapsed_ms", "eh_batch_get_output", "eh_batch_get_rho", "eh_batch_dims",
    "eh_batch_ado_updates", "eh_batch_profile", "eh_batch_enable_adjoint", "eh_batch_backward",
    "eh_batch_get_grad", "eh_batch_run_eager", "eh_batch_autotune", "eh_batch_tuning_report",
    "eh_batch_select", "eh_batch_plan", "eh_ctx_mem_info", "eh_batch_set_chunk", "eh_batch_start",
    "eh_batch_advance",
]


class EtoHeomError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class System(C.Structure):
    """eh_system: the numbers of one propagation problem (param.h:38-82)."""
    _fields_ = [
        ("N", C.c_int), ("K", C.c_int), ("Km", C.c_int), ("L", C.c_int),
        ("site_offset", C.c_int), ("reserved", C.c_int),
        ("alpha", C.c_void_p), ("alpha_t", C.c_void_p), ("gamma", C.c_void_p),
        ("pulse", (C.c_double * 4) * 3),
        ("t_start", C.c_float), ("t_end", C.c_float), ("step_size", C.c_float), ("print_step", C.c_float),
    ]


class Instance(C.Structure):
    _fields_ = [
        ("sys", C.POINTER(System)), ("H", C.c_void_p), ("X4", C.c_void_p),
        ("nv", C.c_int * 3), ("mode", C.c_int), ("sites", C.c_void_p), ("n_sites", C.c_int),
    ]


_lib = None


def lib():
    """Load libetoheom.so (built by __graft_entry__.build / eto-heom_amd/csrc/Makefile)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libetoheom.so not built: {LIB_PATH} is missing (run __graft_entry__.build())")
        l = C.CDLL(LIB_PATH)
        P, I64, VP = C.POINTER, C.c_int64, C.c_void_p
        sig = {
            "eh_last_error": (C.c_char_p, []),
            "eh_problem_from_key": (C.c_int, [C.c_char_p, C.c_int, P(VP)]),
            "eh_problem_destroy": (C.c_int, [VP]),
            "eh_problem_dims": (C.c_int, [VP, VP]),
            "eh_problem_system": (C.c_int, [VP, P(System)]),
            "eh_problem_get": (I64, [VP, C.c_char_p, VP, I64]),
            "eh_problem_polar_default": (C.c_int, [VP, VP]),
            "eh_problem_set_polar": (C.c_int, [VP, VP]),
            "eh_polar_rotate": (C.c_int, [VP, VP, VP]),
            "eh_construct_hal": (C.c_int, [VP, VP, VP]),
            "eh_problem_samples": (C.c_int, [VP, C.c_int, VP, VP]),
            "eh_rng_stream": (C.c_int, [C.c_ulong, C.c_int, VP]),
            "eh_hierarchy_size": (I64, [C.c_int, C.c_int]),
            "eh_hierarchy_build": (C.c_int, [C.c_int, C.c_int, VP, VP, VP]),
            "eh_ctx_create": (C.c_int, [C.c_int, P(VP)]),
            "eh_ctx_destroy": (C.c_int, [VP]),
            "eh_ctx_info": (C.c_int, [VP, C.c_char_p, C.c_int, P(C.c_int), P(C.c_int)]),
            "eh_2des_solve": (C.c_int, [VP, P(System), VP, VP, C.c_int, C.c_int, C.c_int, VP, I64, P(I64), VP]),
            "eh_2des_solve_noalloc": (C.c_int, [VP, P(System), VP, VP, C.c_int, C.c_int, C.c_int, VP, I64,
                                                P(I64), VP]),
            "eh_dynamics_solve": (C.c_int, [VP, P(System), VP, VP, C.c_int, VP, I64, P(I64), VP]),
            "eh_2des_key_run": (C.c_int, [VP, C.c_char_p, VP, I64, P(I64)]),
            "eh_dynamics_key_run": (C.c_int, [VP, C.c_char_p, VP, C.c_int, VP, I64, P(I64)]),
            "eh_batch_create": (C.c_int, [VP, C.c_int, P(Instance), P(VP)]),
            "eh_batch_destroy": (C.c_int, [VP]),
            "eh_batch_run": (C.c_int, [VP]),
            "eh_batch_sync": (C.c_int, [VP]),
            "eh_batch_elapsed_ms": (C.c_int, [VP, P(C.c_double)]),
            "eh_batch_get_output": (C.c_int, [VP, C.c_int, VP, I64, P(I64)]),
            "eh_batch_get_rho": (C.c_int, [VP, C.c_int, VP]),
            "eh_batch_dims": (C.c_int, [VP, VP]),
            "eh_batch_ado_updates": (I64, [VP]),
            "eh_batch_profile": (C.c_int, [VP, C.c_int, VP]),
            "eh_batch_enable_adjoint": (C.c_int, [VP, C.c_int]),
            "eh_batch_backward": (C.c_int, [VP, VP, I64]),
            "eh_batch_plan": (C.c_int, [C.c_int, P(Instance), C.c_int, VP]),
            "eh_ctx_mem_info": (C.c_int, [VP, P(I64), P(I64)]),
            "eh_batch_set_chunk": (C.c_int, [VP, C.c_int]),
            "eh_batch_start": (C.c_int, [VP]),
            "eh_batch_advance": (C.c_int, [VP, C.c_int]),
            "eh_batch_get_grad": (C.c_int, [VP, C.c_int, VP, VP, VP, VP, VP]),
            "eh_batch_run_eager": (C.c_int, [VP]),
            "eh_batch_autotune": (C.c_int, [VP, C.c_int]),
            "eh_batch_select": (C.c_int, [VP, C.c_int, C.c_int]),
            "eh_batch_tuning_report": (I64, [VP, C.c_char_p, I64]),
        }
        for name, (res, args) in sig.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def _check(rc, allow_nan=False):
    if rc == 0 or (allow_nan and rc == -4):
        return rc
    raise EtoHeomError(int(rc), lib().eh_last_error().decode())


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _ptr(a):
    return None if a is None else a.ctypes.data


@dataclass
class Problem:
    """Plain-array problem (the same fields as the oracle's Problem); alpha/alpha_t/gamma
    are (K, Km) complex, pulses (3, 4) = (E0 [fs^-1], tau0, taup, w0 [fs^-1])."""
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
    def N(self):
        return self.H.shape[0]

    def system(self):
        """(eh_system, keep-alive list)"""
        keep = [_c128(self.alpha), _c128(self.alpha_t), _c128(self.gamma)]
        s = System()
        s.N = self.N
        s.K, s.Km = self.alpha.shape
        s.L = self.L
        s.site_offset = self.site_offset
        s.alpha, s.alpha_t, s.gamma = (k.ctypes.data for k in keep)
        pul = np.asarray(self.pulses, dtype=np.float64).reshape(3, 4)
        for q in range(3):
            for c in range(4):
                s.pulse[q][c] = float(pul[q, c])
        s.t_start, s.t_end, s.step_size, s.print_step = self.t_start, self.t_end, self.step_size, self.print_step
        return s, keep

    def num_ado(self):
        return hierarchy_size(self.alpha.size, self.L)


class Param:
    """A parsed .key file (the reference `param`)."""

    def __init__(self, handle, path):
        self._h = C.c_void_p(handle)
        self.path = path
        d = np.zeros(9, dtype=np.int64)
        _check(lib().eh_problem_dims(self._h, d.ctypes.data))
        (self.mode, self.single_size, self.sys_size, self.K, self.K_m, self.L, self.M,
         self.n_sample, self.seed) = (int(v) for v in d)

    @classmethod
    def from_key(cls, path: str, mode: int = MODE_2DES) -> "Param":
        h = C.c_void_p()
        _check(lib().eh_problem_from_key(os.fsencode(path), mode, C.byref(h)))
        return cls(h.value, path)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.eh_problem_destroy(self._h)
            self._h = None

    def get(self, name: str) -> np.ndarray:
        n = lib().eh_problem_get(self._h, name.encode(), None, 0)
        if n < 0:
            _check(int(n))
        out = np.zeros(max(int(n), 1))
        _check(0 if lib().eh_problem_get(self._h, name.encode(), out.ctypes.data, int(n)) >= 0 else -7)
        out = out[: int(n)]
        N = self.sys_size
        if name in ("Hal", "X_x", "X_y", "X_z"):
            return out.view(np.complex128).reshape(N, N)
        if name == "X":
            return out.view(np.complex128).reshape(-1, N, N)
        if name in ("alpha", "alpha_t", "gamma"):
            return out.view(np.complex128).reshape(self.K, self.K_m) if out.size else out.view(np.complex128)
        if name == "Disorder":
            return out.reshape(N, N)
        return out

    def system(self):
        s = System()
        _check(lib().eh_problem_system(self._h, C.byref(s)))
        return s

    def polar_default(self) -> np.ndarray:
        p = np.zeros((3, 4))
        _check(lib().eh_problem_polar_default(self._h, p.ctypes.data))
        return p

    def set_polar(self, polar) -> None:
        p = np.ascontiguousarray(polar, dtype=np.float64)
        _check(lib().eh_problem_set_polar(self._h, p.ctypes.data))

    def construct_hal(self, dH=None) -> np.ndarray:
        N = self.sys_size
        out = np.zeros((N, N), dtype=np.complex128)
        d = None if dH is None else np.ascontiguousarray(dH, dtype=np.float64)
        _check(lib().eh_construct_hal(self._h, _ptr(d), out.ctypes.data))
        return out

    def samples(self, n_sample=None):
        n = self.n_sample if n_sample is None else n_sample
        N = self.sys_size
        H = np.zeros((n, N, N), dtype=np.complex128)
        X = np.zeros((n, 4, N, N), dtype=np.complex128)
        _check(lib().eh_problem_samples(self._h, n, H.ctypes.data, X.ctypes.data))
        return H, X


def polar_rotate(c3, a3, polar) -> np.ndarray:
    p = np.ascontiguousarray(polar, dtype=np.float64).copy()
    c = np.ascontiguousarray(c3, dtype=np.float64)
    a = np.ascontiguousarray(a3, dtype=np.float64)
    _check(lib().eh_polar_rotate(c.ctypes.data, a.ctypes.data, p.ctypes.data))
    return p


def rng_stream(seed: int, count: int) -> np.ndarray:
    out = np.zeros(count, dtype=np.uint32)
    _check(lib().eh_rng_stream(seed, count, out.ctypes.data))
    return out


def hierarchy_size(d: int, L: int) -> int:
    m = lib().eh_hierarchy_size(d, L)
    if m < 0:
        _check(int(m))
    return int(m)


def hierarchy_build(d: int, L: int):
    """(ado[M, d] uint8, up[M, d], dn[M, d]) in the reference DFS order."""
    M = hierarchy_size(d, L)
    ado = np.zeros((M, max(d, 1)), dtype=np.uint8)
    up = np.zeros((M, max(d, 1)), dtype=np.int32)
    dn = np.zeros((M, max(d, 1)), dtype=np.int32)
    _check(lib().eh_hierarchy_build(d, L, ado.ctypes.data, up.ctypes.data, dn.ctypes.data))
    return ado[:, :d], up[:, :d], dn[:, :d]


class Context:
    """A device context (one per host thread)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(lib().eh_ctx_create(device, C.byref(h)))
        self._h = h
        name = C.create_string_buffer(256)
        ncu, lds = C.c_int(0), C.c_int(0)
        _check(lib().eh_ctx_info(self._h, name, 256, C.byref(ncu), C.byref(lds)))
        self.name, self.n_cu, self.lds_per_block = name.value.decode(), ncu.value, lds.value

    def mem_info(self):
        """(free, total) device memory in bytes."""
        fr, tot = C.c_int64(0), C.c_int64(0)
        _check(lib().eh_ctx_mem_info(self._h, C.byref(fr), C.byref(tot)))
        return fr.value, tot.value

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.eh_ctx_destroy(self._h)
            self._h = None

    __del__ = close


def _instances(items):
    """(eh_instance array, keep-alive list) for Batch items."""
    keep = []
    arr = (Instance * len(items))()
    for i, it in enumerate(items):
        s, k = it["problem"].system()
        H = _c128(it["H"])
        X = _c128(it["X"]) if it.get("X") is not None else None
        sites = np.ascontiguousarray(it.get("sites", ()), dtype=np.int32)
        keep += [s, k, H, X, sites]
        arr[i].sys = C.pointer(s)
        arr[i].H = H.ctypes.data
        arr[i].X4 = _ptr(X)
        nv = it.get("nv", (1, 1, 1))
        for q in range(3):
            arr[i].nv[q] = nv[q]
        arr[i].mode = it.get("mode", MODE_2DES)
        arr[i].sites = sites.ctypes.data if sites.size else None
        arr[i].n_sites = int(sites.size)
    return arr, keep


def batch_plan(items, checkpoint: int = 32) -> dict:
    """Host-only checks and device-memory plan of a batch (eh_batch_plan): no device needed."""
    arr, keep = _instances(items)
    o = np.zeros(6, dtype=np.int64)
    _check(lib().eh_batch_plan(len(items), arr, int(checkpoint), o.ctypes.data))
    return {"total_ados": int(o[0]), "total_elems": int(o[1]), "max_steps": int(o[2]),
            "forward_bytes": int(o[3]), "adjoint_replay_bytes": int(o[4]), "adjoint_traj_bytes": int(o[5])}


def _sys_of(key):
    if isinstance(key, Param):
        return key.system(), []
    return key.system()


def _nsteps(s: System) -> int:
    f = np.float32
    nstep = f(f(s.t_end) - f(s.t_start)) / f(s.step_size)
    return int(np.floor(float(nstep))) + 1


def twoD_spectrum_solver(ctx: Context, key, H, nv1: int, nv2: int, nv3: int, X=None, want_rho=False):
    """P(t) of one propagation (complex array); with want_rho also the final state (M, N, N)."""
    s, keep = _sys_of(key)
    if X is None:
        X = key.get("X") if isinstance(key, Param) else key.X
    Hc, Xc = _c128(H), _c128(X)
    cap = _nsteps(s) + 1
    P = np.zeros(cap, dtype=np.complex128)
    n = C.c_int64(0)
    M = hierarchy_size(s.K * s.Km, s.L)
    rho = np.zeros((M, s.N, s.N), dtype=np.complex128) if want_rho else None
    rc = lib().eh_2des_solve(ctx._h, C.byref(s), Hc.ctypes.data, Xc.ctypes.data, nv1, nv2, nv3,
                             P.ctypes.data, cap, C.byref(n), _ptr(rho))
    _check(rc, allow_nan=True)
    return (P[: n.value].copy(), rho) if want_rho else P[: n.value].copy()


def twoD_spectrum_solver_noalloc(ctx: Context, key, H, nv1: int, nv2: int, nv3: int, X=None, want_rho=False):
    """twoD_spectrum_solver through eh_2des_solve_noalloc: no device allocation (the
    library's fixed scratch array; EH_ERANGE for propagations that do not fit it)."""
    s, keep = _sys_of(key)
    if X is None:
        X = key.get("X") if isinstance(key, Param) else key.X
    Hc, Xc = _c128(H), _c128(X)
    cap = _nsteps(s) + 1
    P = np.zeros(cap, dtype=np.complex128)
    n = C.c_int64(0)
    M = hierarchy_size(s.K * s.Km, s.L)
    rho = np.zeros((M, s.N, s.N), dtype=np.complex128) if want_rho else None
    rc = lib().eh_2des_solve_noalloc(ctx._h, C.byref(s), Hc.ctypes.data, Xc.ctypes.data, nv1, nv2, nv3,
                                     P.ctypes.data, cap, C.byref(n), _ptr(rho))
    _check(rc, allow_nan=True)
    return (P[: n.value].copy(), rho) if want_rho else P[: n.value].copy()


def dynamics_solver(ctx: Context, key, H, sites=(1, 2), want_rho=False):
    """Re rho_0(s-1, s-1) per readout for each site: array (len(sites), points)."""
    s, keep = _sys_of(key)
    Hc = _c128(H)
    cap = _nsteps(s) + 1
    st = np.ascontiguousarray(sites, dtype=np.int32)
    pop = np.zeros((len(st), cap))
    n = C.c_int64(0)
    M = hierarchy_size(s.K * s.Km, s.L)
    rho = np.zeros((M, s.N, s.N), dtype=np.complex128) if want_rho else None
    rc = lib().eh_dynamics_solve(ctx._h, C.byref(s), Hc.ctypes.data, st.ctypes.data, len(st),
                                 pop.ctypes.data, cap, C.byref(n), _ptr(rho))
    _check(rc, allow_nan=True)
    return (pop[:, : n.value].copy(), rho) if want_rho else pop[:, : n.value].copy()


def run_2des_key(ctx: Context, path: str) -> np.ndarray:
    """The whole CPU_2DES driver for a key file: averaged p111 - p110 - p101."""
    n = C.c_int64(0)
    _check(lib().eh_2des_key_run(ctx._h, os.fsencode(path), None, 0, C.byref(n)))
    P = np.zeros(max(n.value, 1), dtype=np.complex128)
    _check(lib().eh_2des_key_run(ctx._h, os.fsencode(path), P.ctypes.data, n.value, C.byref(n)), allow_nan=True)
    return P[: n.value]


def run_dynamics_key(ctx: Context, path: str, sites=(1, 2)) -> np.ndarray:
    n = C.c_int64(0)
    st = np.ascontiguousarray(sites, dtype=np.int32)
    _check(lib().eh_dynamics_key_run(ctx._h, os.fsencode(path), st.ctypes.data, len(st), None, 0, C.byref(n)))
    pop = np.zeros((len(st), max(n.value, 1)))
    _check(lib().eh_dynamics_key_run(ctx._h, os.fsencode(path), st.ctypes.data, len(st), pop.ctypes.data,
                                     pop.shape[1], C.byref(n)), allow_nan=True)
    return pop[:, : n.value]


class Batch:
    """Independent propagations run together: items are dicts with keys
    problem (Problem), H, X (2DES), nv, mode, sites (dynamics)."""

    def __init__(self, ctx: Context, items, adjoint: int = 0):
        """adjoint > 0 enables gradients with a checkpoint every `adjoint` steps."""
        self.ctx = ctx
        self.items = items
        arr, self._keep = _instances(items)
        self._arr = arr
        h = C.c_void_p()
        _check(lib().eh_batch_create(ctx._h, len(items), arr, C.byref(h)))
        self._h = h
        d = np.zeros(5, dtype=np.int64)
        _check(lib().eh_batch_dims(self._h, d.ctypes.data))
        self.n_inst, self.total_ados, self.total_elems, self.max_steps, self.chunk = (int(v) for v in d)
        self.ado_updates = int(lib().eh_batch_ado_updates(self._h))
        if adjoint:
            self.enable_adjoint(adjoint)

    def enable_adjoint(self, checkpoint: int):
        """Enable gradients with a checkpoint every `checkpoint` steps (after any
        select / autotune: the backward uses the launch candidates chosen then)."""
        _check(lib().eh_batch_enable_adjoint(self._h, int(checkpoint)))

    def backward(self, dP):
        """dP: list of complex arrays (dL/dRe P + i dL/dIm P), one per instance, each
        shaped like that instance's output."""
        if len(dP) != self.n_inst:
            raise ValueError(f"backward: {len(dP)} gradients for {self.n_inst} instances")
        parts = [np.asarray(g, dtype=np.complex128).ravel() for g in dP]
        for i, g in enumerate(parts):
            n = C.c_int64(0)
            _check(lib().eh_batch_get_output(self._h, i, None, 0, C.byref(n)), allow_nan=True)
            want = n.value * max(1, len(self.items[i].get("sites", ())) if self.items[i].get("mode", MODE_2DES)
                                 != MODE_2DES else 1)
            if g.size != want:
                raise ValueError(f"backward: instance {i} has {want} readouts, its gradient {g.size}")
        flat = np.ascontiguousarray(np.concatenate(parts)) if parts else np.zeros(0, np.complex128)
        _check(lib().eh_batch_backward(self._h, flat.ctypes.data, flat.size))

    def grad(self, i: int):
        """Gradients of instance i: dict H (N,N), X (4,N,N), alpha, alpha_t, gamma (K,Km)."""
        p = self.items[i]["problem"]
        N, (K, Km) = p.N, p.alpha.shape
        out = {"H": np.zeros((N, N), complex), "X": np.zeros((4, N, N), complex),
               "alpha": np.zeros((K, Km), complex), "alpha_t": np.zeros((K, Km), complex),
               "gamma": np.zeros((K, Km), complex)}
        _check(lib().eh_batch_get_grad(self._h, i, out["H"].ctypes.data, out["X"].ctypes.data,
                                       out["alpha"].ctypes.data, out["alpha_t"].ctypes.data,
                                       out["gamma"].ctypes.data))
        return out

    def run(self):
        _check(lib().eh_batch_run(self._h))

    def sync(self):
        _check(lib().eh_batch_sync(self._h))

    def elapsed_ms(self) -> float:
        v = C.c_double(0)
        _check(lib().eh_batch_elapsed_ms(self._h, C.byref(v)))
        return v.value

    def output(self, i: int):
        n = C.c_int64(0)
        _check(lib().eh_batch_get_output(self._h, i, None, 0, C.byref(n)), allow_nan=True)
        it = self.items[i]
        if it.get("mode", MODE_2DES) == MODE_2DES:
            P = np.zeros(max(n.value, 1), dtype=np.complex128)
            _check(lib().eh_batch_get_output(self._h, i, P.ctypes.data, n.value, C.byref(n)), allow_nan=True)
            return P[: n.value]
        ns = len(it.get("sites", ()))
        pop = np.zeros((ns, max(n.value, 1)))
        _check(lib().eh_batch_get_output(self._h, i, pop.ctypes.data, pop.shape[1], C.byref(n)), allow_nan=True)
        return pop[:, : n.value]

    def profile(self, nsteps: int = 20):
        """Mean per-launch durations (ms) of {prologue, stage0..3} from HIP events."""
        out = np.zeros(5)
        _check(lib().eh_batch_profile(self._h, nsteps, out.ctypes.data))
        return out

    def set_chunk(self, chunk: int):
        """Re-capture the forward graph with `chunk` RK4 steps per replay."""
        _check(lib().eh_batch_set_chunk(self._h, int(chunk)))
        self.chunk = int(chunk)

    def start(self):
        """Reset every state; then advance() runs the propagation piece by piece."""
        _check(lib().eh_batch_start(self._h))

    def advance(self, nsteps: int):
        """Run the next nsteps RK4 steps of every instance (a multiple of the chunk)."""
        _check(lib().eh_batch_advance(self._h, int(nsteps)))

    def run_eager(self):
        """The same propagation with every kernel launched one by one from the host (no graph)."""
        _check(lib().eh_batch_run_eager(self._h))

    def autotune(self, steps: int = 5) -> str:
        """Pick the fastest launch candidate per N class; returns the tuning report."""
        _check(lib().eh_batch_autotune(self._h, steps))
        return self.tuning_report()

    def select(self, group: int, cand: int):
        """Launch candidate `cand` for group `group` (as numbered in tuning_report)."""
        _check(lib().eh_batch_select(self._h, group, cand))

    def candidates(self):
        """[(group, cand, kernel, tile)] parsed from the tuning report."""
        out = []
        for line in self.tuning_report().splitlines():
            f = line.split()
            out.append((int(f[1]), int(f[3]), f[6], int(f[7].split("=")[1])))
        return out

    def tuning_report(self) -> str:
        n = lib().eh_batch_tuning_report(self._h, None, 0)
        buf = C.create_string_buffer(int(n) + 1)
        lib().eh_batch_tuning_report(self._h, buf, n + 1)
        return buf.value.decode()

    def rho(self, i: int):
        p = self.items[i]["problem"]
        M = p.num_ado()
        out = np.zeros((M, p.N, p.N), dtype=np.complex128)
        _check(lib().eh_batch_get_rho(self._h, i, out.ctypes.data))
        return out

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.eh_batch_destroy(self._h)
            self._h = None

    __del__ = close
