"""End-to-end two-site DMRG on the engine: energies against exact diagonalisation and the
reference's own regression energy, and sigma against the oracle at every site of a sweep."""
import numpy as np
import pytest
import scipy.sparse.linalg as sla
import torch

from mqdmrg import models
from mqdmrg.dmrg import TwoSiteDMRG
from mqdmrg import engine as E
from oracle import oracle as O

pytestmark = pytest.mark.gpu

# dmrg/regression/applications/dmrg/ref/test_coded_bh_ts.h5, /spectrum/results/Energy/mean/value
# (coded boson Hubbard, open chain L=10, Nmax=2, t=1, U=8, N=5, two-site, m=200, truncation 1e-10)
BH_REFERENCE_ENERGY = -6.8810903606393472


@pytest.fixture(scope="module")
def eng():
    e = E.Engine(0)
    yield e
    e.close()


def ed_energy(model, L, target, **kw):
    basis, mpo, _ = models.chain(model, L, **kw)
    H = models.mpo_dense_chain(mpo, basis.phys)
    # restrict to the target charge sector
    ch = [c for c, s in zip(basis.phys.charges, basis.phys.sizes) for _ in range(s)]
    tot = np.zeros((1, 2), dtype=int)
    for _ in range(L):
        tot = (tot[:, None, :] + np.asarray(ch)[None, :, :]).reshape(-1, 2)
    sel = np.where((tot[:, 0] == target[0]) & (tot[:, 1] == target[1]))[0]
    Hs = H[np.ix_(sel, sel)]
    return float(np.linalg.eigvalsh(Hs)[0])


def test_heisenberg_matches_exact(eng):
    L = 10
    d = TwoSiteDMRG(eng, "heisenberg", L, m=64, cutoff=1e-12, init_m=8, seed=3, jd_tol=1e-10)
    e = d.run(3)
    assert abs(e - ed_energy("heisenberg", L, (0, 0))) < 1e-8


def test_hubbard_matches_exact(eng):
    L = 6
    d = TwoSiteDMRG(eng, "hubbard", L, m=128, cutoff=1e-12, init_m=8, seed=4, jd_tol=1e-10, U=4.0)
    e = d.run(3)
    assert abs(e - ed_energy("hubbard", L, (3, 3), U=4.0)) < 1e-8


def test_boson_hubbard_reference_energy(eng):
    """The reference's regression energy (same model, sizes, bond dimension and truncation)."""
    d = TwoSiteDMRG(eng, "bosehubbard", 10, m=200, cutoff=1e-10, init_m=8, seed=5, jd_tol=1e-10,
                    nmax=2, t=1.0, U=8.0, N=5)
    e = d.run(4)
    assert abs(e - BH_REFERENCE_ENERGY) < 1e-8, e


def test_sigma_parity_at_every_site_heisenberg_L32(eng):
    """Config 1 (Heisenberg L=32, U1, two-site): sigma of the engine equals the oracle's on the
    identical psi and boundaries at every site pair of a sweep."""
    errs = []

    def hook(dm, i, theta, wdesc, plan):
        Lh = O.HostBoundary(dm.Lb[i].index, dm.Lb[i].download())
        Rh = O.HostBoundary(dm.Rb[i + 2].index, dm.Rb[i + 2].download())
        psi = theta.data.cpu().numpy()
        ref = O.SigmaOracle(theta.desc, wdesc, Lh, Rh).apply(psi)
        import torch
        out = torch.zeros_like(theta.data)
        plan.apply(theta.data, out)
        got = out.cpu().numpy()
        errs.append(np.linalg.norm(got - ref) / np.linalg.norm(ref))

    d = TwoSiteDMRG(eng, "heisenberg", 32, m=128, cutoff=1e-10, init_m=16, seed=1, hook=hook)
    d.init_boundaries()
    d.sweep()
    assert len(errs) == 2 * 31
    assert max(errs) <= 1e-12, max(errs)


def test_ab_initio_matches_fci(eng):
    """Synthetic ab initio Hamiltonian (the benchmark integral generator), L=6 orbitals, N=6:
    two-site DMRG at a bond dimension covering the full space equals FCI."""
    from mqdmrg import qc
    L = 6
    h, eri = qc.synthetic_integrals(L, seed=11)
    e_fci = qc.fci_energy(h, eri, 3, 3, ecore=0.5)
    d = TwoSiteDMRG(eng, "qc", L, m=400, cutoff=1e-14, init_m=8, seed=2, jd_tol=1e-11, jd_maxiter=12,
                    integral_seed=11, ecore=0.5, N=6)
    e = d.run(3)
    assert abs(e - e_fci) < 1e-8, (e, e_fci)


@pytest.mark.parametrize("precision", [E.PREC_F32, E.PREC_BF16X3])
def test_mixed_precision_dmrg_energy(eng, precision):
    """Two-site DMRG on the ab initio Hamiltonian (synthetic integrals, L=8, N=8) with sigma in FP32
    or BF16x3 and FP64 residual refinement ends within 1e-6 Eh of the same run in FP64."""
    L = 8
    kw = dict(m=64, cutoff=1e-12, init_m=8, seed=3, jd_tol=1e-10, jd_maxiter=12, integral_seed=5, ecore=0.0, N=8)
    e64 = TwoSiteDMRG(eng, "qc", L, **kw).run(3)
    d = TwoSiteDMRG(eng, "qc", L, precision=precision, max_outer=4, **kw)
    e = d.run(3)
    print("E64 %.12f  E(%d) %.12f  dE %.2e  sigma calls %s" % (e64, precision, e, e - e64, d.sigma_calls), flush=True)
    assert abs(e - e64) <= 1e-6, (e, e64)
    assert d.sigma_calls["reduced"] > d.sigma_calls["fp64"]


def test_gram_split_matches_svd_split_in_a_truncating_sweep(eng, monkeypatch):
    """Ab initio chain (synthetic integrals, L=12, N=12) at m=48, where the truncation binds: at
    every split of two sweeps, the batched Gram-eigh factors the product uses equal a per-sector
    LAPACK-style SVD of the same two-site tensor on the GPU (same kept states per sector, same
    truncated weight, truncated products within 1e-10 of the tensor's norm, orthonormal factor to
    1e-13), and the energy equals that of the same run with every split on exact SVDs."""
    import torch
    import mqdmrg.dmrg as DM
    kw = dict(m=48, cutoff=1e-10, init_m=8, seed=6, jd_tol=1e-10, jd_maxiter=10, integral_seed=3, ecore=0.0, N=12)
    orig = DM.split_factors
    seen = []

    def checked(mats, keys, direction, device, Mmax, cutoff):
        fac, keeps, svals, tw, meth = orig(mats, keys, direction, device, Mmax, cutoff)
        full, sv = DM.svd_factors(mats, keys, direction, device)
        k2, tw2 = DM.truncate([sv[c] for c in keys], Mmax, cutoff)
        diff = nrm = orth = 0.0
        for c, ka, kb in zip(keys, keeps, k2):
            X = mats[c] if direction > 0 else mats[c].t()
            A = fac[c][0] @ fac[c][1] if ka else torch.zeros_like(X)
            if ka:
                Q = fac[c][0]
                orth = max(orth, (Q.t() @ Q - torch.eye(ka, dtype=Q.dtype, device=Q.device)).abs().max().item())
            Bm = full[c][0][:, :kb] @ full[c][1][:kb] if kb else torch.zeros_like(X)
            diff += ((A - Bm) ** 2).sum().item()
            nrm += (X ** 2).sum().item()
        seen.append((meth, keeps == k2, abs(tw - tw2), (diff / nrm) ** 0.5, orth))
        return fac, keeps, svals, tw, meth

    monkeypatch.setattr(DM, "split_factors", checked)
    e = TwoSiteDMRG(eng, "qc", 12, **kw).run(2)
    monkeypatch.setattr(DM, "split_factors", orig)
    gram = [s for s in seen if s[0] == "gram"]
    assert len(gram) >= len(seen) // 2, [s[0] for s in seen]
    for meth, same, dtw, dth, orth in seen:
        assert same and dtw <= 1e-12 and dth <= 1e-10 and orth <= 1e-13, (meth, same, dtw, dth, orth)
    monkeypatch.setattr(DM, "GRAM_RESOLVED", float("inf"))
    e_svd = TwoSiteDMRG(eng, "qc", 12, **kw).run(2)
    assert abs(e - e_svd) <= 1e-9, (e, e_svd)


def test_expval_by_left_boundary_propagation(eng):
    """expval / multi_expval / norm2 over the whole chain on the device: after converged DMRG of
    the Hubbard chain (L=6, exact at m=128) <psi|psi> = 1, <psi|H|psi> equals the sweep energy and
    the exact ground-state energy, and <psi|2 I|psi> = 2 (an MPO other than the Hamiltonian)."""
    L = 6
    d = TwoSiteDMRG(eng, "hubbard", L, m=128, cutoff=1e-12, init_m=8, seed=4, jd_tol=1e-10, U=4.0)
    e = d.run(3)
    assert abs(d.norm2() - 1.0) < 1e-12
    h = d.expval()
    assert abs(h - e) < 1e-10, (h, e)
    assert abs(h - ed_energy("hubbard", L, (3, 3), U=4.0)) < 1e-8
    two = models.identity_mpo(d.phys, L)
    two[0].terms[(0, 0)] = [(op, 2.0) for op, _ in two[0].terms[(0, 0)]]
    assert abs(d.expval(two) - 2.0) < 1e-12
    assert d.multi_expval(two).shape == (1,)


def test_hubbard_L64_m1024_sigma_and_steps_match_oracle():
    """Fermi-Hubbard chain L=64 (t=1, U=4, half filling, TwoU1) at bond dimension m=1024, middle site:
    the grouped FP64 sigma and both boundary steps equal the oracle's to 1e-12 relative."""
    from oracle import oracle as O
    from mqdmrg import mps
    from mqdmrg.symmetry import SiteLayout
    from mqdmrg.workloads import SyntheticSite
    eng = E.Engine(0)
    try:
        ws = SyntheticSite("hubbard", 64, 31, 1024, seed=4)
        Lc, Rc = ws.random_cohorts(ws.left_index), ws.random_cohorts(ws.right_index)
        hl = O.HostBoundary(ws.left_index, [c if c is not None else np.zeros(0) for c in Lc])
        hr = O.HostBoundary(ws.right_index, [c if c is not None else np.zeros(0) for c in Rc])
        Lb = E.DeviceBoundary.upload(eng, ws.left_index, Lc)
        Rb = E.DeviceBoundary.upload(eng, ws.right_index, Rc)
        psi_h = ws.random_psi()
        ref = O.SigmaOracle(ws.site_desc, ws.mpo_desc, hl, hr).apply(psi_h)
        plan = E.SigmaPlan(eng, ws.site_desc, ws.mpo_desc, Lb, Rb, algorithm=E.SIGMA_GROUPED)
        psi = torch.from_numpy(psi_h).to("cuda:0")
        out = torch.full_like(psi, np.nan)
        plan.apply(psi, out)
        eng.sync()
        err = np.linalg.norm(out.cpu().numpy() - ref) / np.linalg.norm(ref)
        print("sigma rel err %.2e (%d elements)" % (err, len(psi_h)), flush=True)
        assert err <= 1e-12
        sym, phys = ws.basis.sym, ws.basis.phys
        rng = np.random.default_rng(1)

        def cmp(hb, db):
            got = db.download()
            num = sum(float(np.sum((g - h) ** 2)) for g, h in zip(got, hb.cohorts))
            den = sum(float(np.sum(h ** 2)) for h in hb.cohorts)
            return np.sqrt(num / den)
        k = ws.site
        sd, md = E.SiteDesc(sym, ws.bonds[k], phys, ws.bonds[k + 1]), E.MpoDesc(ws.mpo[k])
        lay = SiteLayout(ws.bonds[k], phys, ws.bonds[k + 1])
        v = rng.standard_normal(lay.rp_total)
        lp = mps.rp_to_lp(lay, v)
        el = cmp(O.left_step(sd, md, lp, v, hl, True),
                 E.left_step(eng, sd, md, torch.from_numpy(lp).to("cuda:0"), torch.from_numpy(v).to("cuda:0"), Lb, True))
        k = ws.site + 1
        sd, md = E.SiteDesc(sym, ws.bonds[k], phys, ws.bonds[k + 1]), E.MpoDesc(ws.mpo[k])
        lay = SiteLayout(ws.bonds[k], phys, ws.bonds[k + 1])
        v = rng.standard_normal(lay.rp_total)
        vd = torch.from_numpy(v).to("cuda:0")
        er = cmp(O.right_step(sd, md, v, v, hr, True), E.right_step(eng, sd, md, vd, vd, Rb, True))
        print("left step %.2e right step %.2e" % (el, er), flush=True)
        assert el <= 1e-12 and er <= 1e-12
    finally:
        eng.close()


@pytest.mark.parametrize("L,m,sweeps,bound", [(64, 1024, 4, 1e-6), (24, 1024, 4, 1e-8)])
def test_hubbard_free_fermion_energy(L, m, sweeps, bound):
    """Hubbard chain at U=0, half filling: two-site FP64 DMRG at bond dimension m reaches the closed
    form E = 2 sum_{k=1..L/2} -2 cos(pi k / (L+1)) from above (variational), within the truncation
    error of m: 1e-6 at L=64 (the benchmark chain; two spin species of a critical chain need
    m ~ 1024^2 for 1e-8 there), 1e-8 at L=24."""
    exact = 2.0 * sum(-2.0 * np.cos(np.pi * k / (L + 1.0)) for k in range(1, L // 2 + 1))
    eng = E.Engine(0)
    try:
        d = TwoSiteDMRG(eng, "hubbard", L, m=m, cutoff=1e-14, init_m=16, seed=5, jd_tol=1e-10, jd_maxiter=10, U=0.0)
        e = d.run(sweeps)
        print("L %d m %d: E %.12f exact %.12f diff %.3e per sweep %s" % (L, m, e, exact, e - exact, d.energies),
              flush=True)
        assert -1e-10 <= e - exact <= bound
    finally:
        eng.close()
