"""The density-matrix noise of the split on the GPU (noise_density_left / _right over the engine's
grouped boundary steps, dm_factors): the noise term against the dense einsum construction
(abelian chains) and against the same functions run over the oracle's steps (SU2U1 block), and
two-site DMRG with noise against exact energies."""
import numpy as np
import pytest
import torch

from chain import SiteProblem
from su2_chain import SU2BlockProblem
from oracle import oracle as O
from mqdmrg import dmrg as DM, engine as E, mps, su2
from mqdmrg.dmrg import TwoSiteDMRG
from test_cpu_noise import _HostBoundary, _HostEngine, _lp_rows, _rp_cols
from test_gpu_dmrg import ed_energy

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = E.Engine(0)
    yield e
    e.close()


def upload(eng, hb):
    return E.DeviceBoundary.upload(eng, hb.index, hb.cohorts)


def _mats_lp(lay, vec, dev):
    lp = mps.rp_to_lp(lay, vec)
    out = {}
    for rb, rc in enumerate(lay.right.charges):
        n, rs = lay.lp_rows[rb], lay.right.sizes[rb]
        o = lay.lp_offset[rb]
        out[rc] = torch.from_numpy(lp[o:o + n * rs].reshape(rs, n).T.copy()).to(dev)
    return out


def _mats_rp(lay, vec, dev):
    out = {}
    for lb, lc in enumerate(lay.left.charges):
        ls, cols = lay.left.sizes[lb], lay.rp_cols[lb]
        o = lay.rp_offset[lb]
        out[lc] = torch.from_numpy(vec[o:o + ls * cols].reshape(cols, ls).T.copy()).to(dev)
    return out


@pytest.mark.parametrize("model,seed", [("heisenberg", 1), ("hubbard", 2)])
def test_noise_density_left_on_device_equals_dense(eng, model, seed):
    print("case left", model, seed, flush=True)
    p = SiteProblem(model, L=8, site=3, Mmax=6, seed=seed)
    lay = p.layouts[3]
    md = DM.plain_mpo_desc(p.mpo[3])
    mids = list(lay.right.charges)
    got = DM.noise_density_left(eng, p.sym, lay.left, lay.phys, mids, _mats_lp(lay, p.vecs[3], eng.torch_device),
                                md, upload(eng, p.Lb))
    T = np.einsum("xpa,asc->xpsc", O.dense_boundary(p.Lb), O.dense_site(lay, p.vecs[3]))
    T = np.einsum("xyqs,xpsc->ypqc", p.mpo[3].dense(lay.phys), T)
    scale = np.abs(T).max() ** 2
    for rb, rc in enumerate(mids):
        rows = _lp_rows(lay, rb)
        if rows:
            A = np.stack([T[:, a, s, :] for a, s in rows], axis=0).reshape(len(rows), -1)
            np.testing.assert_allclose(got[rc].cpu().numpy(), A @ A.T, rtol=0, atol=1e-12 * scale)


@pytest.mark.parametrize("model,seed", [("heisenberg", 4), ("hubbard", 5)])
def test_noise_density_right_on_device_equals_dense(eng, model, seed):
    print("case right", model, seed, flush=True)
    p = SiteProblem(model, L=8, site=2, Mmax=6, seed=seed)
    lay = p.layouts[3]
    md = DM.plain_mpo_desc(p.mpo[3])
    mids = list(lay.left.charges)
    got = DM.noise_density_right(eng, p.sym, lay.right, lay.phys, mids, _mats_rp(lay, p.vecs[3], eng.torch_device),
                                 md, upload(eng, p.Rb))
    T = np.einsum("asc,ycd->yasd", O.dense_site(lay, p.vecs[3]), O.dense_boundary(p.Rb))
    T = np.einsum("xyqs,yasd->xaqd", p.mpo[3].dense(lay.phys), T)
    scale = np.abs(T).max() ** 2
    for lb, lc in enumerate(mids):
        cols = _rp_cols(lay, lb)
        if cols:
            A = np.stack([T[:, :, q, d] for q, d in cols], axis=0).reshape(len(cols), -1)
            np.testing.assert_allclose(got[lc].cpu().numpy(), A @ A.T, rtol=0, atol=1e-12 * scale)


def test_su2_noise_density_matches_oracle_steps(eng, monkeypatch):
    """SU2U1 block (L=6, N=6, singlet): the device noise terms equal the same functions run over the
    oracle's SU2 boundary steps, to 1e-12 relative."""
    p = SU2BlockProblem(6, 6, 3, two_s=0, seed=3)
    sd, v_lp, v_rp = p.left_sites[-1]
    lay = sd.layout
    md = DM.plain_mpo_desc(p.mpo[2])
    mids = list(lay.right.charges)
    mats = _mats_lp(lay, v_rp, eng.torch_device)
    got = DM.noise_density_left(eng, p.basis.sym, lay.left, lay.phys, mids, mats, md,
                                upload(eng, p.left_bounds[-2]))
    sdr, vr = p.right_sites[-1]
    layr = sdr.layout
    mdr = DM.plain_mpo_desc(p.mpo[5])
    midr = list(layr.left.charges)
    matr = _mats_rp(layr, vr, eng.torch_device)
    got_r = DM.noise_density_right(eng, p.basis.sym, layr.right, layr.phys, midr, matr, mdr,
                                   upload(eng, p.right_bounds[-2]))
    # the same construction over the oracle's steps, on the host
    host = _HostEngine()
    monkeypatch.setattr(E, "left_step", lambda e, s, m, b, k, l, symmetric=True, algorithm=0: _HostBoundary(
        O.left_step(s, m, b.numpy(), k.numpy(), l, symmetric)))
    monkeypatch.setattr(E, "right_step", lambda e, s, m, b, k, r, symmetric=True, algorithm=0: _HostBoundary(
        O.right_step(s, m, b.numpy(), k.numpy(), r, symmetric)))

    def l2r(e, s, lp, rp):
        rp[:s.layout.rp_total] = torch.from_numpy(mps.lp_to_rp(s.layout, lp.numpy()))
    monkeypatch.setattr(E, "reshape_left_to_right", l2r)
    ref = DM.noise_density_left(host, p.basis.sym, lay.left, lay.phys, mids, {c: m.cpu() for c, m in mats.items()},
                                md, p.left_bounds[-2])
    ref_r = DM.noise_density_right(host, p.basis.sym, layr.right, layr.phys, midr,
                                   {c: m.cpu() for c, m in matr.items()}, mdr, p.right_bounds[-2])
    for g, r in ((got, ref), (got_r, ref_r)):
        num = sum(float(torch.sum((g[c].cpu() - r[c]) ** 2)) for c in r)
        den = sum(float(torch.sum(r[c] ** 2)) for c in r)
        assert den > 0 and np.sqrt(num / den) <= 1e-12


@pytest.mark.parametrize("model,L,target,kw", [("heisenberg", 10, (0, 0), {}), ("hubbard", 6, (3, 3), {"U": 4.0})])
def test_dmrg_with_noise_reaches_exact_energy(eng, model, L, target, kw):
    """Two sweeps with noise (alpha 1e-3, then 1e-4) from a small initial state, then two without:
    the exact ground-state energy to 1e-8; the noisy sweeps use the density-matrix split."""
    print("case", model, L, flush=True)
    d = TwoSiteDMRG(eng, model, L, m=128, cutoff=1e-12, init_m=2, seed=3, jd_tol=1e-10, noise=[1e-3, 1e-4], **kw)
    e = d.run(4)
    assert d.split_methods.get("dm", 0) == 4 * (L - 1)
    ref = ed_energy(model, L, target, **kw)
    print("E %.12f exact %.12f" % (e, ref), flush=True)
    assert abs(e - ref) < 1e-8


def test_su2_dmrg_with_noise_matches_fci(eng):
    L, N = 6, 6
    d = TwoSiteDMRG(eng, "qc_su2", L, m=256, cutoff=1e-14, init_m=2, seed=2, jd_tol=1e-10, N=N, two_s=0,
                    integral_seed=7, noise=[1e-3, 1e-4])
    e = d.run(4)
    from mqdmrg import qc
    h, eri = qc.synthetic_integrals(L, seed=7)
    ef = su2.fci_spin_energy(h, eri, N, 0)
    print("E_dmrg %.12f E_fci %.12f" % (e, ef), flush=True)
    assert d.split_methods.get("dm", 0) == 4 * (L - 1)
    assert abs(e - ef) < 1e-8
