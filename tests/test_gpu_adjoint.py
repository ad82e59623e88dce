"""GPU checks of the sigma adjoint (mq_sigma_adjoint): parity with the oracle's brute-force term
derivatives, the linearity identities at middle-of-chain shapes, and the 1-/2-RDMs of an ab
initio ground state (dE/dh, dE/d(pq|rs)) against central finite differences of the FCI energy,
also through a torch.autograd.Function."""
import itertools

import numpy as np
import pytest
import torch

from chain import SiteProblem
from oracle import oracle as O
from mqdmrg import engine as E, qc
from mqdmrg.dmrg import TwoSiteDMRG

pytestmark = pytest.mark.gpu

GRAD_RTOL = 1e-12   # FP64, same bar as sigma


@pytest.fixture(scope="module")
def eng():
    e = E.Engine(0)
    yield e
    e.close()


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda:0")


CASES = [("heisenberg", 8, 3, 6, 1), ("heisenberg", 10, 0, 8, 2), ("hubbard", 8, 3, 10, 5), ("hubbard", 6, 2, 6, 4),
         ("bosehubbard", 8, 3, 6, 8), ("qc", 6, 2, 8, 6), ("qc", 8, 0, 8, 7)]


@pytest.mark.parametrize("model,L,site,M,seed", CASES)
def test_sigma_adjoint_matches_oracle(eng, model, L, site, M, seed):
    print("case", model, L, site, M, seed, flush=True)
    p = SiteProblem(model, L=L, site=site, Mmax=M, seed=seed)
    ref = O.sigma_term_gradient(p.site_desc, p.mpo_desc, p.Lb, p.Rb, p.psi)
    Lb = E.DeviceBoundary.upload(eng, p.Lb.index, p.Lb.cohorts)
    Rb = E.DeviceBoundary.upload(eng, p.Rb.index, p.Rb.cohorts)
    plan = E.SigmaPlan(eng, p.site_desc, p.mpo_desc, Lb, Rb, algorithm=E.SIGMA_GROUPED, adjoint=True)
    g = plan.adjoint(dev(p.psi))
    err = np.linalg.norm(g - ref) / np.linalg.norm(ref)
    print("  terms", len(g), "relative error", err, flush=True)
    assert err <= GRAD_RTOL, err
    # the adjoint leaves the plan's sigma intact
    out = torch.zeros_like(dev(p.psi))
    plan.apply(dev(p.psi), out)
    eng.sync()
    sref = O.SigmaOracle(p.site_desc, p.mpo_desc, p.Lb, p.Rb).apply(p.psi)
    assert np.linalg.norm(out.cpu().numpy() - sref) <= 1e-12 * np.linalg.norm(sref)


def test_adjoint_requires_flag(eng):
    p = SiteProblem("hubbard", L=6, site=2, Mmax=6, seed=4)
    Lb = E.DeviceBoundary.upload(eng, p.Lb.index, p.Lb.cohorts)
    Rb = E.DeviceBoundary.upload(eng, p.Rb.index, p.Rb.cohorts)
    plan = E.SigmaPlan(eng, p.site_desc, p.mpo_desc, Lb, Rb, algorithm=E.SIGMA_GROUPED)
    with pytest.raises(Exception):
        plan.adjoint(dev(p.psi))
    with pytest.raises(Exception):
        E.SigmaPlan(eng, p.site_desc, p.mpo_desc, Lb, Rb, algorithm=E.SIGMA_SCHEDULE, adjoint=True)


@pytest.mark.parametrize("model,L,m", [("qc", 12, 128), ("hubbard", 16, 256)])
def test_sigma_adjoint_linearity_at_synthetic_shapes(eng, model, L, m):
    """Middle-of-chain shapes: sum_t scale_t grad_t = psi . sigma (Euler), and changing one term's
    scale by s changes psi . sigma by s grad_t (H_eff is linear in every scale)."""
    from mqdmrg.workloads import SyntheticSite
    print("case", model, L, m, flush=True)
    ws = SyntheticSite(model, L, L // 2 - 1, m, seed=3)
    Lb = E.DeviceBoundary.upload(eng, ws.left_index, ws.random_cohorts(ws.left_index))
    Rb = E.DeviceBoundary.upload(eng, ws.right_index, ws.random_cohorts(ws.right_index))
    psi = dev(ws.random_psi())
    plan = E.SigmaPlan(eng, ws.site_desc, ws.mpo_desc, Lb, Rb, algorithm=E.SIGMA_GROUPED, adjoint=True)
    g = plan.adjoint(psi)

    def quad(desc):
        pl = E.SigmaPlan(eng, ws.site_desc, desc, Lb, Rb, algorithm=E.SIGMA_GROUPED)
        out = torch.zeros_like(psi)
        pl.apply(psi, out)
        eng.sync()
        pl.close()
        return float(torch.dot(psi, out))
    e = quad(ws.mpo_desc)
    sc = ws.mpo_desc.term_scale
    print("  terms", len(g), "euler", float(g @ sc), e, flush=True)
    assert abs(float(g @ sc) - e) <= 1e-11 * abs(e)
    rng = np.random.default_rng(0)
    for t in rng.choice(len(g), size=3, replace=False):
        s2 = sc.copy()
        s2[t] += 0.5
        e2 = quad(ws.mpo_desc.with_scales(s2))
        assert abs((e2 - e) - 0.5 * g[t]) <= 1e-11 * abs(e), (t, e2 - e, 0.5 * g[t])


def _qc_dmrg(eng, h, eri, nup, ndn, m=64, sweeps=4):
    L = h.shape[0]
    basis, mpo, _, src = qc.qc_mpo(h, eri, sources=True)
    d = TwoSiteDMRG(eng, "qc", L, m=m, cutoff=1e-14, init_m=8, seed=2, jd_tol=1e-12, jd_maxiter=20,
                    mpo=mpo, basis=basis, target=(nup, ndn))
    e = d.run(sweeps)
    grads, energies = d.mpo_gradients()
    assert max(abs(x - e) for x in energies) < 1e-9, (e, energies)   # same state at every center
    dh, deri, _ = qc.mpo_gradient_to_integrals(mpo, src, grads, L)
    return e, energies, dh, deri


class DMRGEnergy(torch.autograd.Function):
    """E(h, eri) of the DMRG ground state; backward from the sigma adjoint (dE/dh, dE/d(pq|rs))."""

    @staticmethod
    def forward(ctx, h, eri, eng, nup, ndn):
        e, _, dh, deri = _qc_dmrg(eng, h.detach().cpu().numpy(), eri.detach().cpu().numpy(), nup, ndn)
        ctx.save_for_backward(torch.from_numpy(dh), torch.from_numpy(deri))
        return torch.tensor(e, dtype=torch.float64)

    @staticmethod
    def backward(ctx, go):
        dh, deri = ctx.saved_tensors
        return go * dh, go * deri, None, None, None


def test_rdm_from_adjoint_matches_fci_finite_differences(eng):
    """Exact two-site DMRG of a 4-orbital ab initio Hamiltonian (N=4): the adjoint sweep's dE/dh_pq
    (spin-summed 1-RDM) and dE/d(pq|rs) (2-RDM / 2) equal central finite differences of the FCI
    energy (step 1e-5) to 1e-6 relative, through torch.autograd."""
    L, nup, ndn = 4, 2, 2
    h, eri = qc.synthetic_integrals(L, seed=3)
    efci = qc.fci_energy(h, eri, nup, ndn)
    ht = torch.tensor(h, requires_grad=True)
    et = torch.tensor(eri, requires_grad=True)
    e = DMRGEnergy.apply(ht, et, eng, nup, ndn)
    e.backward()
    ev = float(e.detach())
    print("E", ev, "FCI", efci, flush=True)
    assert abs(ev - efci) < 1e-9
    fd = qc.fci_energy_gradient_fd(h, eri, nup, ndn)
    fh = np.array([[fd[("h", (p, q))] for q in range(L)] for p in range(L)])
    fe = np.zeros((L,) * 4)
    for idx in itertools.product(range(L), repeat=4):
        fe[idx] = fd[("eri", idx)]
    eh = np.linalg.norm(ht.grad.numpy() - fh) / np.linalg.norm(fh)
    ee = np.linalg.norm(et.grad.numpy() - fe) / np.linalg.norm(fe)
    print("dE/dh rel err", eh, "dE/deri rel err", ee, flush=True)
    assert eh <= 1e-6 and ee <= 1e-6
    assert abs(np.trace(ht.grad.numpy()) - (nup + ndn)) < 1e-8


def test_rdm1_measurement_equals_adjoint_gradient(eng):
    """The spin-summed 1-RDM measured as expectation values of one-body MPOs (TwoSiteDMRG.rdm1,
    the measurement route) equals dE/dh from the sigma adjoint of the same converged state
    (Hellmann-Feynman; the adjoint's dE/dh is pinned to FCI finite differences above) to 1e-8, and
    its trace is the electron number."""
    L, nup, ndn = 4, 2, 2
    h, eri = qc.synthetic_integrals(L, seed=3)
    basis, mpo, _, src = qc.qc_mpo(h, eri, sources=True)
    d = TwoSiteDMRG(eng, "qc", L, m=64, cutoff=1e-14, init_m=8, seed=2, jd_tol=1e-12, jd_maxiter=20,
                    mpo=mpo, basis=basis, target=(nup, ndn))
    d.run(4)
    gamma = d.rdm1()
    grads, _ = d.mpo_gradients()
    dh, _, _ = qc.mpo_gradient_to_integrals(mpo, src, grads, L)
    dh = 0.5 * (dh + dh.T)
    print("trace", np.trace(gamma), "max |gamma - dE/dh|", np.abs(gamma - dh).max(), flush=True)
    assert abs(np.trace(gamma) - (nup + ndn)) < 1e-9
    assert np.abs(gamma - dh).max() < 1e-8
