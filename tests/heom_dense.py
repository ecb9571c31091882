"""Independent dense-Liouvillian formulation of the ETO-HEOM, used to pin the oracle.

This is written from the textbook (Tanimura) hierarchy in its UNSCALED form,
with superoperators built from Kronecker products, not from the oracle's code:

    d rho_n/dt = -i (V x I - I x V^T) rho_n - (sum_p n_p gamma_p) rho_n
                 + sum_p        i (S_p x I - I x S_p^T) rho_{n+e_p}
                 + sum_p  n_p ( i alpha_p S_p x I - i alpha~_p I x S_p^T ) rho_{n-e_p}

(row-major vec: vec(A rho) = (A x I) vec(rho), vec(rho B) = (I x B^T) vec(rho)).
The reference (etoheom-cpu/lib/dynamics.cpp:143-179) uses the scaled ADOs
rho~_n = rho_n / sqrt(prod_p n_p! |alpha_p|^n_p); rho_0 is identical in both.

The RK4 step with V frozen over the step is the exact degree-4 Taylor
polynomial of the step's generator: y <- sum_{m<=4} (hA)^m/m! y.
"""
from __future__ import annotations

import itertools
import math

import numpy as np


def ado_list(d: int, L: int):
    """All n in N^d with |n| <= L, lexicographic (= the reference's DFS order)."""
    return sorted(t for t in itertools.product(range(L + 1), repeat=d) if sum(t) <= L)


def liouvillian(V, alpha, alpha_t, gamma, L, site_offset=1):
    """Dense generator A of the unscaled hierarchy; alpha etc. are (K, Km)."""
    N = V.shape[0]
    K, Km = alpha.shape
    d = K * Km
    ados = ado_list(d, L)
    idx = {a: i for i, a in enumerate(ados)}
    M = len(ados)
    nn = N * N
    I = np.eye(N)
    A = np.zeros((M * nn, M * nn), dtype=np.complex128)
    comm_V = -1j * (np.kron(V, I) - np.kron(I, V.T))
    for i, n in enumerate(ados):
        blk = slice(i * nn, (i + 1) * nn)
        damp = 0j
        for p in range(d):
            j, k = p % K, p // K
            damp += n[p] * gamma[j, k]
        A[blk, blk] += comm_V - damp * np.eye(nn)
        for p in range(d):
            j, k = p % K, p // K
            S = np.zeros((N, N))
            S[j + site_offset, j + site_offset] = 1.0
            up = list(n)
            up[p] += 1
            up = tuple(up)
            if up in idx:
                b = idx[up]
                A[blk, b * nn:(b + 1) * nn] += 1j * (np.kron(S, I) - np.kron(I, S.T))
            if n[p] > 0:
                dn = list(n)
                dn[p] -= 1
                b = idx[tuple(dn)]
                A[blk, b * nn:(b + 1) * nn] += n[p] * (1j * alpha[j, k] * np.kron(S, I)
                                                       - 1j * alpha_t[j, k] * np.kron(I, S.T))
    return A, ados


def rk4_matrix(A, h):
    """The RK4 polynomial sum_{m=0}^{4} (hA)^m / m!"""
    n = A.shape[0]
    out = np.eye(n, dtype=np.complex128)
    term = np.eye(n, dtype=np.complex128)
    for m in range(1, 5):
        term = term @ (h * A) / m
        out = out + term
    return out


def float_time_grid(t_start, t_end, dt):
    """t_i sequence of the reference loop: float32 arithmetic, floor(nstep)+1 steps."""
    f = np.float32
    nstep = f(f(t_end) - f(t_start)) / f(dt)
    n = int(math.floor(float(nstep))) + 1
    ts = []
    t = f(t_start)
    for _ in range(n):
        ts.append(t)
        t = f(t + f(dt))
    return ts


def field_scalar(pulse, t, dagger):
    """E(t) e^{+-i w (t - tau0)}, written from the Gaussian pulse definition."""
    E0, tau0, taup, w0 = pulse
    s = float(t) - tau0
    env = E0 * math.exp(-2.77 * s * s / (taup * taup))
    th = (-w0 if dagger else w0) * s
    return env * complex(math.cos(th), math.sin(th))


def V_of_t(H, X, pulses, t, nv):
    V = H.astype(np.complex128).copy()
    ops = [X[0], X[1].T, X[2].T]
    for q in range(3):
        E0, tau0, taup, w0 = pulses[q]
        if nv[q] == 1 and abs(float(t) - tau0) <= taup * 2:
            V = V - field_scalar(pulses[q], t, q != 0) * ops[q]
    return V


def propagate_2des(H, X, alpha, alpha_t, gamma, L, pulses, t_start, t_end, dt, print_step,
                   nv=(1, 1, 1), rho0=None, max_steps=None, site_offset=1):
    """Dense reference propagation; returns (P list, final rho_0)."""
    N = H.shape[0]
    A0, ados = liouvillian(np.zeros_like(H, dtype=np.complex128), alpha, alpha_t, gamma, L, site_offset)
    M = len(ados)
    nn = N * N
    Icomm = np.eye(N)
    y = np.zeros(M * nn, dtype=np.complex128)
    if rho0 is None:
        y[0] = 1.0
    else:
        y[:nn] = np.asarray(rho0).reshape(-1)
    ts = float_time_grid(t_start, t_end, dt)
    ps = int(np.float32(print_step) / np.float32(dt))
    h = float(np.float32(dt))
    P = []
    cache = {}
    for i, t in enumerate(ts):
        if max_steps is not None and i >= max_steps:
            break
        V = V_of_t(H, X, pulses, t, nv) if X is not None else H.astype(np.complex128)
        if float(t) >= pulses[2][1] and i % ps == 0:
            r0 = y[:nn].reshape(N, N)
            P.append(np.trace(r0 @ X[3]))
        key = V.tobytes()
        if key not in cache:
            A = A0.copy()
            comm = -1j * (np.kron(V, Icomm) - np.kron(Icomm, V.T))
            for b in range(M):
                A[b * nn:(b + 1) * nn, b * nn:(b + 1) * nn] += comm
            cache = {key: rk4_matrix(A, h)}
        y = cache[key] @ y
    return np.array(P), y[:nn].reshape(N, N)
