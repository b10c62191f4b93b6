"""Independent mathematics that pins the oracle (test infrastructure).

Nothing here reuses the oracle's per-amplitude formulas:

* ``dense_qaoa`` builds every gate as a full 2^N x 2^N matrix from Pauli
  strings with ``scipy.linalg.expm`` -- H on every qubit, then per layer
  RZZ(w*gamma) = expm(-i*w*gamma/2 * Z_u Z_v) on every edge and
  RX(beta) = expm(-i*beta/2 * X_q) on every qubit -- the circuit that
  scripts/valiate_ans_scripts/validate_state.py:36-46 builds in Qiskit, in
  Qiskit's little-endian qubit order.  Small N only.
* ``kn_phi`` is the Hamming-weight closed form for the reference's own
  instance (K_N, unit weights, uniform angles): the state stays symmetric
  under qubit permutations, so psi(b) = phi(popcount(b)); evaluated in mpmath.
* ``p1_expect_cut`` is the single-layer expectation of the weighted cut,
  derived in closed form (Heisenberg picture: U_B^+ Z U_B = cos(beta) Z +
  sin(beta) Y, then the cost exponentials on |+>^N), valid for any graph and
  any N.  Checked against ``dense_qaoa`` in tests/test_oracle_pins.py before
  it is used at large N.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.linalg

_I2 = np.eye(2, dtype=np.complex128)
_X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
_Z = np.array([[1, 0], [0, -1]], dtype=np.complex128)
_H = np.array([[1, 1], [1, -1]], dtype=np.complex128) / math.sqrt(2.0)


def _embed(N, ops):
    """Full operator with ops[q] on qubit q (qubit 0 = least significant bit,
    i.e. the rightmost Kronecker factor)."""
    M = np.ones((1, 1), dtype=np.complex128)
    for q in reversed(range(N)):
        M = np.kron(M, ops.get(q, _I2))
    return M


def dense_qaoa(N, edges, gammas, betas):
    assert N <= 10
    psi = np.zeros(1 << N, dtype=np.complex128)
    psi[0] = 1.0
    psi = _embed(N, {q: _H for q in range(N)}) @ psi
    rx_cache = {}
    for g, b in zip(gammas, betas):
        for u, v, w in edges:
            if w == 0:
                continue
            ZZ = _embed(N, {u: _Z, v: _Z})
            psi = scipy.linalg.expm(-1j * (w * g) / 2.0 * ZZ) @ psi
        if b not in rx_cache:
            rx_cache[b] = scipy.linalg.expm(-1j * b / 2.0 * _X)
        RX = rx_cache[b]
        for q in range(N):
            psi = _embed(N, {q: RX}) @ psi
    return psi


def dense_cut_diag(N, edges):
    """Cut(z) for every basis index, from the Z-string form
    Cut = sum_e w_e (1 - Z_u Z_v)/2."""
    d = np.zeros(1 << N)
    for u, v, w in edges:
        zz = np.real(np.diag(_embed(N, {u: _Z, v: _Z})))
        d += w * (1.0 - zz) / 2.0
    return d


def kn_phi(N, gammas, betas, dps=40):
    """phi[h] for K_N with unit weights after len(gammas) layers (mpmath)."""
    import mpmath as mp

    mp.mp.dps = dps
    phi = [mp.mpc(mp.mpf(2) ** (-mp.mpf(N) / 2))] * (N + 1)
    binom = [[mp.binomial(n, k) for k in range(N + 1)] for n in range(N + 1)]
    npairs = N * (N - 1) // 2
    for g, b in zip(gammas, betas):
        g = mp.mpf(float(g))
        b = mp.mpf(float(b))
        phi = [phi[h] * mp.expj(g * (2 * h * (N - h) - npairs) / 2) for h in range(N + 1)]
        c = mp.cos(b / 2)
        s = mp.sin(b / 2)
        # m[d] = c^(N-d) (-i s)^d
        m = [c ** (N - d) * (mp.mpc(0, -1) * s) ** d for d in range(N + 1)]
        new = []
        for hp in range(N + 1):
            acc = mp.mpc(0)
            for h in range(N + 1):
                inner = mp.mpc(0)
                for j in range(max(0, h + hp - N), min(h, hp) + 1):
                    inner += binom[hp][j] * binom[N - hp][h - j] * m[h + hp - 2 * j]
                acc += phi[h] * inner
            new.append(acc)
        phi = new
    return phi


def kn_expect_cut(N, phi):
    import mpmath as mp

    return sum(mp.binomial(N, h) * abs(phi[h]) ** 2 * h * (N - h) for h in range(N + 1))


def p1_expect_cut(N, edges, gamma, beta):
    """<Cut> after one layer, for any weighted graph (float64)."""
    W = np.zeros((N, N))
    for u, v, w in edges:
        W[u, v] = W[v, u] = w
    tot = 0.0
    for u, v, w in edges:
        others = [k for k in range(N) if k != u and k != v]
        wu = W[u, others]
        wv = W[v, others]
        A = np.prod(np.cos(gamma * wu))
        B = np.prod(np.cos(gamma * wv))
        Cm = np.prod(np.cos(gamma * (wu - wv)))
        Cp = np.prod(np.cos(gamma * (wu + wv)))
        zz = 0.5 * math.sin(2 * beta) * math.sin(gamma * w) * (A + B) + 0.5 * math.sin(beta) ** 2 * (Cm - Cp)
        tot += w * (1.0 - zz) / 2.0
    return tot


def axis_swap_blocks(sv, N, out, inn, k):
    """Exchange qubit blocks [out,out+k) and [inn,inn+k) with a numpy axis
    permutation of the (2,)*N tensor (axis N-1-q holds qubit q)."""
    t = sv.reshape((2,) * N)
    perm = list(range(N))
    for i in range(k):
        a, b = N - 1 - (out + i), N - 1 - (inn + i)
        perm[a], perm[b] = perm[b], perm[a]
    return np.ascontiguousarray(np.transpose(t, perm)).reshape(-1)


def logical_from_physical(sv, qmap):
    """Physical order -> logical order with numpy axes (independent of
    oq_to_logical): physical qubit p holds logical qubit qmap[p]."""
    N = len(qmap)
    t = sv.reshape((2,) * N)
    # axis N-1-p of t is physical qubit p; logical qubit L must end at axis N-1-L
    src_axis_for_logical = {int(qmap[p]): N - 1 - p for p in range(N)}
    perm = [src_axis_for_logical[N - 1 - ax] for ax in range(N)]
    return np.ascontiguousarray(np.transpose(t, perm)).reshape(-1)
