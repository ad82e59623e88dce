"""Lattice Hamiltonians of the reference's regression tests, written as integrals (h, eri) of the
ab initio form H = sum h_pq c+_p c_q + 1/2 sum (pq|rs) c+_p c+_r c_s c_q, and small exact solvers
for them that share no code with the engine or the oracle."""
import itertools
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_observables.json")


def reference_observables():
    with open(GOLDEN) as f:
        return json.load(f)


def ladder_hopping(L, t=1.0, tp=1.0):
    """Open two-leg ladder of L rungs, site = 2 * rung + leg: -t along the legs, -t' on the rungs."""
    n = 2 * L
    h = np.zeros((n, n))
    for x in range(L):
        h[2 * x, 2 * x + 1] = h[2 * x + 1, 2 * x] = -tp
        if x + 1 < L:
            for y in range(2):
                h[2 * x + y, 2 * (x + 1) + y] = h[2 * (x + 1) + y, 2 * x + y] = -t
    return h


def square_hopping(Lx, Ly, t=1.0):
    """Open Lx x Ly square lattice, site = x + Lx * y, -t on nearest neighbours."""
    n = Lx * Ly
    h = np.zeros((n, n))
    for y in range(Ly):
        for x in range(Lx):
            i = x + Lx * y
            if x + 1 < Lx:
                h[i, i + 1] = h[i + 1, i] = -t
            if y + 1 < Ly:
                h[i, i + Lx] = h[i + Lx, i] = -t
    return h


def hubbard_integrals(hop, U):
    """(h, eri) of sum hop_pq c+_ps c_qs + U sum_p n_p,up n_p,down: (pp|pp) = U."""
    n = hop.shape[0]
    eri = np.zeros((n, n, n, n))
    for p in range(n):
        eri[p, p, p, p] = U
    return hop.copy(), eri


def exact_hubbard(hop, U, n_up, n_down):
    """Ground state of the Hubbard model in the (n_up, n_down) sector by direct diagonalisation over
    products of up and down occupation strings: energy and the one-body density matrices
    <c+_i,s c_j,s> of both spins."""
    n = hop.shape[0]

    def strings(k):
        return [sum(1 << i for i in c) for c in itertools.combinations(range(n), k)]

    def hop_matrix(st):
        pos = {b: a for a, b in enumerate(st)}
        T = np.zeros((len(st), len(st)))
        ops = {}
        for a, b in enumerate(st):
            for i in range(n):
                for j in range(n):
                    if not (b >> j) & 1:
                        continue
                    b1 = b & ~(1 << j)
                    if (b1 >> i) & 1:
                        continue
                    sign = (-1) ** (bin(b1 & ((1 << i) - 1)).count("1") + bin(b & ((1 << j) - 1)).count("1"))
                    b2 = b1 | (1 << i)
                    ops.setdefault((i, j), []).append((pos[b2], a, sign))
                    if hop[i, j] != 0.0:
                        T[pos[b2], a] += hop[i, j] * sign
        return T, ops

    su, sd = strings(n_up), strings(n_down)
    Tu, ou = hop_matrix(su)
    Td, od = hop_matrix(sd)
    nu = np.array([[(b >> i) & 1 for i in range(n)] for b in su], float)
    nd = np.array([[(b >> i) & 1 for i in range(n)] for b in sd], float)
    H = np.kron(Tu, np.eye(len(sd))) + np.kron(np.eye(len(su)), Td) + U * np.diag((nu @ nd.T).ravel())
    w, v = np.linalg.eigh(H)
    psi = v[:, 0].reshape(len(su), len(sd))
    gu, gd = np.zeros((n, n)), np.zeros((n, n))
    for (i, j), lst in ou.items():
        gu[i, j] = sum(sg * float(psi[r] @ psi[c]) for r, c, sg in lst)
    for (i, j), lst in od.items():
        gd[i, j] = sum(sg * float(psi[:, r] @ psi[:, c]) for r, c, sg in lst)
    return float(w[0]), gu, gd
