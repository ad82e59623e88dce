"""CPU tests of the density-matrix split with noise (mqdmrg.dmrg.noise_density_left / _right and
dm_factors, the restatement of alpha_dm_direct(_right) and heev_truncate).  The noise functions
run here with the oracle's boundary steps standing in for the device steps (same arguments, host
arrays), and are compared with the noise term formed densely with einsum from the dense boundary,
MPO and site tensor."""
import numpy as np
import pytest
import torch

from chain import SiteProblem
from oracle import oracle as O
from mqdmrg import dmrg as DM, engine as E, mps
from mqdmrg.symmetry import fuse, sub


class _HostBoundary:
    """What the noise functions use of a DeviceBoundary, over an oracle boundary."""

    def __init__(self, hb):
        self.index, self.hb = hb.index, hb

    def cohort(self, ci):
        return torch.from_numpy(np.ascontiguousarray(self.hb.cohorts[ci], np.float64))

    def close(self):
        pass


class _HostEngine:
    torch_device = torch.device("cpu")


@pytest.fixture
def oracle_steps(monkeypatch):
    def left_step(eng, site, md, bra_lp, ket_rp, left, symmetric=True, algorithm=0):
        return _HostBoundary(O.left_step(site, md, bra_lp.numpy(), ket_rp.numpy(), left, symmetric))

    def right_step(eng, site, md, bra_rp, ket_rp, right, symmetric=True, algorithm=0):
        return _HostBoundary(O.right_step(site, md, bra_rp.numpy(), ket_rp.numpy(), right, symmetric))

    def l2r(eng, site, lp, rp):
        rp[:site.layout.rp_total] = torch.from_numpy(mps.lp_to_rp(site.layout, lp.numpy()))

    monkeypatch.setattr(E, "left_step", left_step)
    monkeypatch.setattr(E, "right_step", right_step)
    monkeypatch.setattr(E, "reshape_left_to_right", l2r)
    return _HostEngine()


def _lp_rows(lay, rb):
    """(left full index, phys full index) of the left-paired rows of right sector rb."""
    ol = np.cumsum([0] + lay.left.sizes)
    op = np.cumsum([0] + lay.phys.sizes)
    rc = lay.right.charges[rb]
    out = [None] * lay.lp_rows[rb]
    for s, (pc, d) in enumerate(zip(lay.phys.charges, lay.phys.sizes)):
        lc = sub(rc, pc)
        if not lay.left.has(lc):
            continue
        lb = lay.left.position(lc)
        ls = lay.left.sizes[lb]
        r0 = lay.lpb(pc, lc)
        for ss in range(d):
            for i in range(ls):
                out[r0 + ss * ls + i] = (ol[lb] + i, op[s] + ss)
    return out


def _rp_cols(lay, lb):
    """(phys full index, right full index) of the right-paired columns of left sector lb."""
    op = np.cumsum([0] + lay.phys.sizes)
    orr = np.cumsum([0] + lay.right.sizes)
    lc = lay.left.charges[lb]
    out = [None] * lay.rp_cols[lb]
    for s, (pc, d) in enumerate(zip(lay.phys.charges, lay.phys.sizes)):
        rc = fuse(lc, pc)
        if not lay.right.has(rc):
            continue
        rb = lay.right.position(rc)
        rs = lay.right.sizes[rb]
        c0 = lay.rpb(pc, rc)
        for ss in range(d):
            for j in range(rs):
                out[c0 + ss * rs + j] = (op[s] + ss, orr[rb] + j)
    return out


@pytest.mark.parametrize("model,seed", [("heisenberg", 1), ("hubbard", 2), ("bosehubbard", 3)])
def test_noise_density_left_equals_dense(oracle_steps, model, seed):
    p = SiteProblem(model, L=8, site=3, Mmax=6, seed=seed)
    k = 3
    lay = p.layouts[k]
    md = DM.plain_mpo_desc(p.mpo[k])
    vec = p.vecs[k]
    lp = mps.rp_to_lp(lay, vec)
    mids = list(lay.right.charges)
    mats = {}
    for rb, rc in enumerate(mids):
        n, rs = lay.lp_rows[rb], lay.right.sizes[rb]
        o = lay.lp_offset[rb]
        mats[rc] = torch.from_numpy(lp[o:o + n * rs].reshape(rs, n).T.copy())
    got = DM.noise_density_left(oracle_steps, p.sym, lay.left, lay.phys, mids, mats, md, p.Lb)
    # dense: T[y, p, q, c] = sum L[x, p, a] W[x, y, q, s] X[a, s, c]; N = sum_{y, c} T T^T
    X = O.dense_site(lay, vec)
    T = np.einsum("xpa,asc->xpsc", O.dense_boundary(p.Lb), X)
    T = np.einsum("xyqs,xpsc->ypqc", p.mpo[k].dense(lay.phys), T)
    scale, checked = np.abs(T).max() ** 2, 0
    for rb, rc in enumerate(mids):
        rows = _lp_rows(lay, rb)
        if not rows:
            continue
        A = np.stack([T[:, a, s, :] for a, s in rows], axis=0).reshape(len(rows), -1)
        ref = A @ A.T
        np.testing.assert_allclose(got[rc].numpy(), ref, rtol=0, atol=1e-12 * scale)
        checked += int(np.abs(ref).max() > 1e-8 * scale)
    assert checked > 0


@pytest.mark.parametrize("model,seed", [("heisenberg", 4), ("hubbard", 5), ("bosehubbard", 6)])
def test_noise_density_right_equals_dense(oracle_steps, model, seed):
    p2 = SiteProblem(model, L=8, site=2, Mmax=6, seed=seed)   # its Rb covers sites >= 4
    k, Rb = 3, p2.Rb
    lay = p2.layouts[k]
    md = DM.plain_mpo_desc(p2.mpo[k])
    vec = p2.vecs[k]
    mids = list(lay.left.charges)
    mats = {}
    for lb, lc in enumerate(mids):
        ls, cols = lay.left.sizes[lb], lay.rp_cols[lb]
        o = lay.rp_offset[lb]
        mats[lc] = torch.from_numpy(vec[o:o + ls * cols].reshape(cols, ls).T.copy())
    got = DM.noise_density_right(oracle_steps, p2.sym, lay.right, lay.phys, mids, mats, md, Rb)
    # dense: T[x, a, q, d] = sum X[a, s, c] W[x, y, q, s] R[y, c, d]; N = sum_{x, a} T^T T
    X = O.dense_site(lay, vec)
    T = np.einsum("asc,ycd->yasd", X, O.dense_boundary(Rb))
    T = np.einsum("xyqs,yasd->xaqd", p2.mpo[k].dense(lay.phys), T)
    scale, checked = np.abs(T).max() ** 2, 0
    for lb, lc in enumerate(mids):
        cols = _rp_cols(lay, lb)
        if not cols:
            continue
        A = np.stack([T[:, :, q, d] for q, d in cols], axis=0).reshape(len(cols), -1)
        ref = A @ A.T
        np.testing.assert_allclose(got[lc].numpy(), ref, rtol=0, atol=1e-12 * scale)
        checked += int(np.abs(ref).max() > 1e-8 * scale)
    assert checked > 0


def test_dm_factors_without_noise_equal_truncated_svd():
    """alpha = 0: the eigenvectors of X X^T are the left singular vectors, sqrt of the eigenvalues
    the singular values, so keeps, truncated weight and the rank-k product Q_k B_k equal the
    truncated SVD's; with noise the kept Q_k stays orthonormal and B_k = Q_k^T X."""
    rng = np.random.default_rng(3)
    mats, keys = {}, []
    for c, (m, n) in enumerate([(7, 12), (15, 6), (9, 9), (4, 20)]):
        s = np.exp(-rng.uniform(0, 8, min(m, n)))
        U, _ = np.linalg.qr(rng.standard_normal((m, m)))
        V, _ = np.linalg.qr(rng.standard_normal((n, n)))
        mats[(c, 0)] = torch.from_numpy((U[:, :len(s)] * s) @ V[:, :len(s)].T)
        keys.append((c, 0))
    for direction in (+1, -1):
        zero = {c: torch.zeros((mats[c].shape[0] if direction > 0 else mats[c].shape[1],) * 2, dtype=torch.float64)
                for c in keys}
        fac, keeps, svals, tw = DM.dm_factors(mats, keys, direction, "cpu", 12, 1e-3, zero, 0.0)
        ref_s = [np.linalg.svd((mats[c] if direction > 0 else mats[c].t()).numpy(), compute_uv=False) for c in keys]
        ref_keeps, ref_tw = DM.truncate(ref_s, 12, 1e-3)
        assert keeps == ref_keeps
        assert abs(tw - ref_tw) < 1e-10
        for c, k in zip(keys, keeps):
            for a, b in zip(svals[keys.index(c)][:k], ref_s[keys.index(c)][:k]):
                assert abs(a - b) < 1e-7 * ref_s[keys.index(c)][0]
            if k == 0:
                continue
            X = (mats[c] if direction > 0 else mats[c].t()).numpy()
            U, S, Vh = np.linalg.svd(X, full_matrices=False)
            Q, B = fac[c]
            np.testing.assert_allclose((Q @ B).numpy(), (U[:, :k] * S[:k]) @ Vh[:k], atol=1e-10 * S[0])
    noise = {}
    for c in keys:
        G = rng.standard_normal((mats[c].shape[0], 30))
        noise[c] = torch.from_numpy(G @ G.T)
    fac, keeps, _, _ = DM.dm_factors(mats, keys, +1, "cpu", 12, 1e-3, noise, 1e-2)
    for c, k in zip(keys, keeps):
        if k:
            Q, B = fac[c]
            np.testing.assert_allclose((Q.t() @ Q).numpy(), np.eye(k), atol=1e-12)
            np.testing.assert_allclose(B.numpy(), (Q.t() @ mats[c]).numpy(), atol=1e-13)
