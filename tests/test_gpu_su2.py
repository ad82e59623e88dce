"""GPU parity of the spin-adapted (SU2U1) path against the CPU oracle (through the C-ABI), and
two-site DMRG energies of the SU2U1 ab initio Hamiltonian against exact energies of fixed total
spin.  Tolerances: FP64 sigma / boundary steps 1e-12 relative; energies 1e-8 Eh."""
import numpy as np
import pytest
import torch

from su2_chain import SU2BlockProblem
from oracle import oracle as O
from mqdmrg import engine as E, mps, su2
from mqdmrg.dmrg import TwoSiteDMRG
from mqdmrg.workloads import SyntheticSite

pytestmark = pytest.mark.gpu

RTOL = 1e-12


@pytest.fixture(scope="module")
def eng():
    e = E.Engine(0)
    yield e
    e.close()


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda:0")


def upload(eng, hb):
    return E.DeviceBoundary.upload(eng, hb.index, hb.cohorts)


def _cmp_boundaries(hb, db):
    assert db.index.n_cohorts == hb.index.n_cohorts
    np.testing.assert_array_equal(db.index.offsets, hb.index.offsets)
    got = db.download()
    num = sum(float(np.sum((g - h) ** 2)) for g, h in zip(got, hb.cohorts))
    den = sum(float(np.sum(h ** 2)) for h in hb.cohorts)
    assert np.sqrt(num / max(den, 1e-300)) <= RTOL


BLOCKS = [(4, 4, 1, 0), (5, 5, 1, 1), (6, 6, 2, 0), (6, 6, 0, 0), (6, 6, 4, 0), (7, 6, 3, 2)]


@pytest.mark.parametrize("algorithm", [E.SIGMA_SCHEDULE, E.SIGMA_GROUPED])
@pytest.mark.parametrize("L,N,site,two_s", BLOCKS)
def test_su2_sigma_matches_oracle(eng, L, N, site, two_s, algorithm):
    print("case", L, N, site, two_s, algorithm, flush=True)
    p = SU2BlockProblem(L, N, site, two_s=two_s, seed=L + site)
    ref = O.SigmaOracle(p.site_desc, p.mpo_desc, p.Lb, p.Rb).apply(p.psi)
    plan = E.SigmaPlan(eng, p.site_desc, p.mpo_desc, upload(eng, p.Lb), upload(eng, p.Rb), algorithm=algorithm)
    psi = dev(p.psi)
    out = torch.full_like(psi, np.nan)
    plan.apply(psi, out)
    eng.sync()
    err = np.linalg.norm(out.cpu().numpy() - ref) / np.linalg.norm(ref)
    assert err <= RTOL, err
    out2 = torch.zeros_like(psi)
    plan.apply(psi, out2)
    eng.sync()
    assert torch.equal(out, out2)


@pytest.mark.parametrize("algorithm", [E.SIGMA_SCHEDULE, E.SIGMA_GROUPED])
@pytest.mark.parametrize("L,N,site,two_s", [(6, 6, 3, 0), (7, 6, 3, 2)])
def test_su2_boundary_steps_match_oracle(eng, L, N, site, two_s, algorithm):
    print("case", L, N, site, two_s, algorithm, flush=True)
    p = SU2BlockProblem(L, N, site, two_s=two_s, seed=3)
    sd, v_lp, v_rp = p.left_sites[-1]
    md = E.MpoDesc(p.mpo[site - 1])
    ref = O.left_step(sd, md, v_lp, v_rp, p.left_bounds[-2], True)
    got = E.left_step(eng, sd, md, dev(v_lp), dev(v_rp), upload(eng, p.left_bounds[-2]), True, algorithm=algorithm)
    _cmp_boundaries(ref, got)
    sd, v = p.right_sites[-1]
    md = E.MpoDesc(p.mpo[site + 2])
    ref = O.right_step(sd, md, v, v, p.right_bounds[-2], True)
    got = E.right_step(eng, sd, md, dev(v), dev(v), upload(eng, p.right_bounds[-2]), True, algorithm=algorithm)
    _cmp_boundaries(ref, got)


@pytest.mark.parametrize("algorithm", [E.SIGMA_SCHEDULE, E.SIGMA_GROUPED])
@pytest.mark.parametrize("L,m", [(10, 48), (12, 96)])
def test_su2_sigma_synthetic_shapes_match_oracle(eng, L, m, algorithm):
    """Middle-of-chain SU2U1 shapes with truncated random bonds and random boundary data."""
    print("case", L, m, algorithm, flush=True)
    ws = SyntheticSite("qc_su2", L, L // 2 - 1, m, seed=5)
    Lh = O.HostBoundary(ws.left_index, ws.random_cohorts(ws.left_index))
    Rh = O.HostBoundary(ws.right_index, ws.random_cohorts(ws.right_index))
    psi = ws.random_psi()
    ref = O.SigmaOracle(ws.site_desc, ws.mpo_desc, Lh, Rh).apply(psi)
    plan = E.SigmaPlan(eng, ws.site_desc, ws.mpo_desc, upload(eng, Lh), upload(eng, Rh), algorithm=algorithm)
    out = torch.zeros(len(psi), dtype=torch.float64, device="cuda:0")
    plan.apply(dev(psi), out)
    eng.sync()
    err = np.linalg.norm(out.cpu().numpy() - ref) / np.linalg.norm(ref)
    assert err <= RTOL, err


@pytest.mark.parametrize("L,N,two_s", [(6, 6, 0), (5, 5, 1), (6, 4, 2)])
def test_su2_dmrg_matches_fci(eng, L, N, two_s):
    """Two-site SU2U1 DMRG (the reference's qc_su2 model, two-site MPO and ts_reduction) reaches
    the exact lowest energy of the requested total spin."""
    print("case", L, N, two_s, flush=True)
    d = TwoSiteDMRG(eng, "qc_su2", L, m=256, cutoff=1e-14, init_m=4, seed=2, jd_tol=1e-10, N=N, two_s=two_s,
                    integral_seed=7)
    e = d.run(4)
    from mqdmrg import qc
    h, eri = qc.synthetic_integrals(L, seed=7)
    ef = su2.fci_spin_energy(h, eri, N, two_s)
    print("E_dmrg %.12f E_fci %.12f" % (e, ef), flush=True)
    assert abs(e - ef) < 1e-8


def test_su2_rdm_from_adjoint_matches_fci_finite_differences(eng):
    """Exact two-site SU2U1 DMRG of a 4-orbital singlet (N=4): the adjoint sweep gives dE/d(element)
    for every integral element; split over each element's permutation class it equals central
    finite differences (step 1e-5) of the exact singlet energy with respect to single entries of h
    (each stepped with its Hermitian partner) and the class means of those of (pq|rs), to 1e-6
    relative.  The Euler identity
    sum_x x dE/dx = E holds to 1e-10."""
    from mqdmrg import qc
    L, N = 4, 4
    h, eri = qc.synthetic_integrals(L, seed=3)
    basis, mpo, _, elements, src = su2.qc_su2_mpo(h, eri, N, sources=True)
    d = TwoSiteDMRG(eng, "qc_su2", L, m=64, cutoff=1e-14, init_m=8, seed=2, jd_tol=1e-12, jd_maxiter=20,
                    mpo=mpo, basis=basis, target=(N, 0))
    e = d.run(4)
    ef = su2.fci_spin_energy(h, eri, N, 0)
    assert abs(e - ef) < 1e-9, (e, ef)
    grads, energies = d.mpo_gradients()
    assert max(abs(x - e) for x in energies) < 1e-9, (e, energies)
    g = su2.mpo_gradient_to_elements(mpo, src, grads, len(elements))
    vals = np.array([x[0] for x in elements])
    assert abs(float(vals @ g) - e) < 1e-10
    dh, deri, _ = su2.element_gradient_to_tensors(elements, g, L)
    step = 1e-5

    def fd(kind, idx):
        # the entry and its Hermitian partner together, halved (qc.fci_energy_gradient_fd)
        cls = {tuple(idx), qc.hermitian_partner(kind, tuple(idx))}
        out = []
        for sgn in (+1, -1):
            hh, ee = h.copy(), eri.copy()
            for x in cls:
                (hh if kind == "h" else ee)[x] += sgn * step
            out.append(su2.fci_spin_energy(hh, ee, N, 0))
        return (out[0] - out[1]) / (2 * step) / len(cls)
    import itertools
    fh = np.array([[fd("h", (p, q)) for q in range(L)] for p in range(L)])
    fe = np.zeros((L,) * 4)
    for idx in itertools.product(range(L), repeat=4):
        fe[idx] = fd("eri", idx)
    # the spin-adapted generator reads one element per 8-fold permutation class of (pq|rs), so only
    # class sums of the entry derivatives are defined for it (entries of one class need not share
    # their derivative: the 2-RDM has the 4-fold symmetry only); element_gradient_to_tensors
    # spreads the class sum evenly, so compare with the class means of the finite differences
    fe_cls = np.zeros_like(fe)
    for idx in itertools.product(range(L), repeat=4):
        p_, q_, r_, s_ = idx
        cls = {(p_, q_, r_, s_), (q_, p_, r_, s_), (p_, q_, s_, r_), (q_, p_, s_, r_), (r_, s_, p_, q_),
               (s_, r_, p_, q_), (r_, s_, q_, p_), (s_, r_, q_, p_)}
        fe_cls[idx] = sum(fe[c] for c in cls) / len(cls)
    eh = np.linalg.norm(dh - fh) / np.linalg.norm(fh)
    ee = np.linalg.norm(deri - fe_cls) / np.linalg.norm(fe_cls)
    print("dE/dh rel err", eh, "dE/deri rel err", ee, flush=True)
    assert eh <= 1e-6 and ee <= 1e-6
    assert abs(np.trace(dh) - N) < 1e-8


def test_su2_rdm1_measurement_equals_adjoint_gradient(eng):
    """SU2U1 singlet (L=4, N=4): the spin-summed 1-RDM measured as expectation values of one-body
    spin-adapted MPOs (TwoSiteDMRG.rdm1) equals dE/dh from the sigma adjoint of the same converged
    state (pinned to FCI finite differences above) to 1e-8, with trace N."""
    from mqdmrg import qc
    L, N = 4, 4
    h, eri = qc.synthetic_integrals(L, seed=3)
    basis, mpo, _, elements, src = su2.qc_su2_mpo(h, eri, N, sources=True)
    d = TwoSiteDMRG(eng, "qc_su2", L, m=64, cutoff=1e-14, init_m=8, seed=2, jd_tol=1e-12, jd_maxiter=20,
                    mpo=mpo, basis=basis, target=(N, 0))
    d.run(4)
    gamma = d.rdm1()
    grads, _ = d.mpo_gradients()
    g = su2.mpo_gradient_to_elements(mpo, src, grads, len(elements))
    dh, _, _ = su2.element_gradient_to_tensors(elements, g, L)
    dh = 0.5 * (dh + dh.T)
    print("trace", np.trace(gamma), "max |gamma - dE/dh|", np.abs(gamma - dh).max(), flush=True)
    assert abs(np.trace(gamma) - N) < 1e-9
    assert np.abs(gamma - dh).max() < 1e-8


def test_su2_rdm2_measurement_equals_adjoint_gradient(eng):
    """SU2U1 singlet (L=4, N=4): the two-body expectation value of every (pq|rs) permutation class,
    measured as the expectation value of the MPO built from that integral element alone
    (TwoSiteDMRG.rdm2_classes), equals dE/d(element) from the sigma adjoint to 1e-8; the energy
    rebuilt from the measured 1- and 2-RDM (sum over elements of value x measured expectation)
    equals the DMRG energy to 1e-9."""
    from mqdmrg import qc
    L, N = 4, 4
    h, eri = qc.synthetic_integrals(L, seed=3)
    basis, mpo, _, elements, src = su2.qc_su2_mpo(h, eri, N, sources=True)
    d = TwoSiteDMRG(eng, "qc_su2", L, m=64, cutoff=1e-14, init_m=8, seed=2, jd_tol=1e-12, jd_maxiter=20,
                    mpo=mpo, basis=basis, target=(N, 0))
    e = d.run(4)
    gamma = d.rdm1()
    r2 = d.rdm2_classes()
    grads, _ = d.mpo_gradients()
    g = su2.mpo_gradient_to_elements(mpo, src, grads, len(elements))
    two = [(m, x) for m, x in enumerate(elements) if not (x[3] == -1 and x[4] == -1)]
    worst = max(abs(r2[x[1:]] - g[m]) for m, x in two)
    e1 = sum(h[i, j] * (1.0 if i == j else 2.0) * gamma[i, j] for i in range(L) for j in range(i + 1))
    e2 = sum(x[0] * r2[x[1:]] for _, x in two)
    print("classes", len(r2), "max |measured - dE/d(element)|", worst, "E", e, "E from RDMs", e1 + e2, flush=True)
    assert worst < 1e-8
    assert abs(e1 + e2 - e) < 1e-9
