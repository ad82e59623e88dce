"""CPU checks of the sigma adjoint's oracle and of the map from MPO-term derivatives to integral
derivatives (RDMs) of the ab initio Hamiltonian."""
import itertools

import numpy as np
import pytest

from chain import SiteProblem
from oracle import oracle as O
from mqdmrg import models, qc


@pytest.mark.parametrize("model,L,site,M,seed", [("hubbard", 6, 2, 6, 4), ("qc", 6, 2, 6, 6),
                                                  ("heisenberg", 8, 3, 6, 1)])
def test_oracle_term_gradient_identities(model, L, site, M, seed):
    """psi . H_eff psi is linear in every term scale: sum_t scale_t grad_t reproduces it (Euler),
    and a finite step in one scale changes it by exactly step * grad_t."""
    p = SiteProblem(model, L=L, site=site, Mmax=M, seed=seed)
    e = float(p.psi @ O.SigmaOracle(p.site_desc, p.mpo_desc, p.Lb, p.Rb).apply(p.psi))
    g = O.sigma_term_gradient(p.site_desc, p.mpo_desc, p.Lb, p.Rb, p.psi)
    assert abs(float(g @ p.mpo_desc.term_scale) - e) <= 1e-12 * abs(e)
    rng = np.random.default_rng(seed)
    for t in rng.choice(len(g), size=min(3, len(g)), replace=False):
        sc = p.mpo_desc.term_scale.copy()
        sc[t] += 0.25
        e2 = float(p.psi @ O.SigmaOracle(p.site_desc, p.mpo_desc.with_scales(sc), p.Lb, p.Rb).apply(p.psi))
        assert abs((e2 - e) - 0.25 * g[t]) <= 1e-12 * max(abs(e), 1.0), (t, e2 - e, 0.25 * g[t])


def _sector_states(phys, L, target):
    q = []
    for c, d in zip(phys.charges, phys.sizes):
        q += [c] * d
    sel = []
    for k, states in enumerate(itertools.product(range(len(q)), repeat=L)):
        tot = tuple(sum(q[s][i] for s in states) for i in range(len(target)))
        if tot == tuple(target):
            sel.append(k)
    return np.array(sel)


def _dense_term_gradients(mpo, phys, psi, e0):
    """dE/d(scale) of every term of every site of a dense chain state psi (normalised, energy e0),
    from dense left / right environments; checks the Euler identity per site."""
    L = len(mpo)
    d = phys.total()
    Ws = [W.dense(phys) for W in mpo]
    # left[j]: (a, O, I) product of sites < j; right[j]: (b, O, I) product of sites > j
    left = [np.ones((1, 1, 1))]
    for j in range(L - 1):
        nb = np.einsum("aOI,abji->bOjIi", left[-1], Ws[j])
        left.append(nb.reshape(nb.shape[0], nb.shape[1] * d, nb.shape[3] * d))
    right = [None] * L
    right[L - 1] = np.ones((1, 1, 1))
    for j in range(L - 1, 0, -1):
        nb = np.einsum("abji,bOI->ajOiI", Ws[j], right[j])
        right[j - 1] = nb.reshape(nb.shape[0], d * nb.shape[2], d * nb.shape[4])
    grads = []
    for j, W in enumerate(mpo):
        X = psi.reshape(d ** j, d, d ** (L - j - 1))
        T1 = np.einsum("aPQ,PoR->aQoR", left[j], X)
        T2 = np.einsum("bRS,QiS->bRQi", right[j], X)
        Gj = np.einsum("aQoR,bRQi->aboi", T1, T2)     # dE/dW_j[a, b, out, in]
        b1s, b2s, ois, scs = W.sorted_terms()
        dops = [op.dense(phys) for op in W.ops]
        g = np.array([np.sum(Gj[b1s[t], b2s[t]] * dops[ois[t]]) for t in range(len(b1s))])
        # Euler: the Hamiltonian is linear in each site's coefficients
        assert abs(g @ scs - e0) < 1e-9
        grads.append(g)
    return grads


def test_mpo_integral_sources_give_fci_energy_gradient():
    """Exact ground state of a 4-orbital ab initio chain (dense MPO): the derivatives with respect
    to each MPO term scale (dense single-term chains), mapped through the integral sources of the
    switch coefficients, equal central finite differences of the FCI energy with respect to every
    h_pq and (pq|rs) entry (step 1e-5) to 1e-6 relative."""
    L, nup, ndn = 4, 2, 2
    h, eri = qc.synthetic_integrals(L, seed=3)
    basis, mpo, _, src = qc.qc_mpo(h, eri, sources=True)
    phys = basis.phys
    H = models.mpo_dense_chain(mpo, phys)
    sel = _sector_states(phys, L, (nup, ndn))
    w, v = np.linalg.eigh(H[np.ix_(sel, sel)])
    e0 = w[0]
    assert abs(e0 - qc.fci_energy(h, eri, nup, ndn)) < 1e-10
    psi = np.zeros(H.shape[0])
    psi[sel] = v[:, 0]
    grads = _dense_term_gradients(mpo, phys, psi, e0)
    dh, deri, _ = qc.mpo_gradient_to_integrals(mpo, src, grads, L)
    fd = qc.fci_energy_gradient_fd(h, eri, nup, ndn)
    fh = np.array([[fd[("h", (p, q))] for q in range(L)] for p in range(L)])
    fe = np.zeros((L,) * 4)
    for idx in itertools.product(range(L), repeat=4):
        fe[idx] = fd[("eri", idx)]
    assert np.linalg.norm(dh - fh) <= 1e-6 * np.linalg.norm(fh)
    assert np.linalg.norm(deri - fe) <= 1e-6 * np.linalg.norm(fe)
    # spin-summed 1-RDM trace = N
    assert abs(np.trace(dh) - (nup + ndn)) < 1e-8


def test_hubbard_ladder_onebody_matrix_from_term_gradients():
    """Hubbard ladder of 2 rungs (U=8, 2 up + 2 down electrons) as ab initio integrals with every
    one-body entry a term (h_tol < 0, zeros included): the term derivatives of the exact ground
    state, mapped to dE/dh_pq, equal the spin-summed one-body density matrix <c+_p c_q> of a direct
    diagonalisation (the quantity the reference's one-body density matrix measurement reports)."""
    from lattices import ladder_hopping, hubbard_integrals, exact_hubbard
    hop = ladder_hopping(2)
    h, eri = hubbard_integrals(hop, 8.0)
    L, nup, ndn = 4, 2, 2
    basis, mpo, _, src = qc.qc_mpo(h, eri, sources=True, h_tol=-1.0)
    phys = basis.phys
    H = models.mpo_dense_chain(mpo, phys)
    sel = _sector_states(phys, L, (nup, ndn))
    w, v = np.linalg.eigh(H[np.ix_(sel, sel)])
    e_ex, gu, gd = exact_hubbard(hop, 8.0, nup, ndn)
    assert abs(w[0] - e_ex) < 1e-10
    assert w[1] - w[0] > 1e-6   # a unique ground state: its density matrix is well defined
    psi = np.zeros(H.shape[0])
    psi[sel] = v[:, 0]
    grads = _dense_term_gradients(mpo, phys, psi, w[0])
    dh, _, _ = qc.mpo_gradient_to_integrals(mpo, src, grads, L)
    np.testing.assert_allclose(dh, gu + gd, atol=1e-10)
