"""CPU tests (no GPU): the oracle pinned by dense brute force, closed forms and sympy fixtures;
the product's host planner against the oracle's restatement; the C-ABI library surface."""
import ctypes as C
import json
import os
import re

import numpy as np
import pytest
import torch

from chain import SiteProblem
from oracle import oracle as O
from mqdmrg import capi, engine as E, models, mps
from mqdmrg import dmrg as DM
from mqdmrg.dmrg import truncate

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


# --------------------------------------------------------------------------- oracle pins
@pytest.mark.parametrize("model,L,site,M,seed", [("heisenberg", 6, 2, 5, 1), ("heisenberg", 8, 0, 6, 2),
                                                 ("heisenberg", 8, 6, 6, 3), ("hubbard", 6, 2, 5, 4),
                                                 ("bosehubbard", 6, 2, 6, 5)])
def test_oracle_sigma_equals_dense_brute_force(model, L, site, M, seed):
    p = SiteProblem(model, L=L, site=site, Mmax=M, seed=seed)
    sig = O.SigmaOracle(p.site_desc, p.mpo_desc, p.Lb, p.Rb).apply(p.psi)
    Ld, Rd = O.dense_boundary(p.Lb), O.dense_boundary(p.Rb)
    ref = O.dense_to_vec(p.layout, O.dense_sigma(Ld, p.W12.dense(p.phys12), Rd, O.dense_site(p.layout, p.psi)))
    assert np.linalg.norm(sig - ref) <= 1e-13 * np.linalg.norm(ref)


@pytest.mark.parametrize("model", ["heisenberg", "hubbard", "bosehubbard"])
def test_oracle_boundary_steps_equal_dense_brute_force(model):
    p = SiteProblem(model, L=8, site=3, Mmax=6, seed=7)
    k = 4   # Rb covers sites >= 5, Lb sites < 3
    lay = p.layouts[k]
    sd = E.SiteDesc(p.sym, lay.left, lay.phys, lay.right)
    md = E.MpoDesc(p.mpo[k])
    X = O.dense_site(lay, p.vecs[k])
    Wd = p.mpo[k].dense(lay.phys)
    r = O.right_step(sd, md, p.vecs[k], p.vecs[k], p.Rb, True)
    dn = O.dense_right_step(O.dense_boundary(p.Rb), Wd, X, X)
    assert np.abs(O.dense_boundary(r) - dn).max() <= 1e-13 * np.abs(dn).max()
    p2 = SiteProblem(model, L=8, site=4, Mmax=6, seed=8)
    lay = p2.layouts[4]
    sd = E.SiteDesc(p2.sym, lay.left, lay.phys, lay.right)
    md = E.MpoDesc(p2.mpo[4])
    X = O.dense_site(lay, p2.vecs[4])
    lft = O.left_step(sd, md, mps.rp_to_lp(lay, p2.vecs[4]), p2.vecs[4], p2.Lb, True)
    dn = O.dense_left_step(O.dense_boundary(p2.Lb), p2.mpo[4].dense(lay.phys), X, X)
    assert np.abs(O.dense_boundary(lft) - dn).max() <= 1e-13 * np.abs(dn).max()


@pytest.mark.parametrize("model", ["heisenberg", "hubbard", "bosehubbard"])
def test_hermitian_complement_equals_full_computation(model):
    """complement_transpose's transposed references (boundary.h:161-191) reproduce the blocks a
    computation without Hermitian skipping produces."""
    import copy
    basis, mpo, target = models.chain(model, 6)
    rng = np.random.default_rng(2)
    bonds, _, _ = mps.allowed_sectors([basis.phys] * 6, target, 5)
    lays = mps.layouts_for_chain(bonds, [basis.phys] * 6)
    vecs = [mps.random_site_vector(l, rng) for l in lays]

    def strip(W):
        W2 = copy.copy(W)
        W2.left_herm, W2.right_herm = [None] * W.row_dim, [None] * W.col_dim
        return W2
    res = []
    for herm in (True, False):
        Lb = O.trivial_boundary(bonds[0], mpo[0].row_dim)
        for k in range(3):
            sd = E.SiteDesc(basis.sym, lays[k].left, lays[k].phys, lays[k].right)
            md = E.MpoDesc(mpo[k] if herm else strip(mpo[k]))
            Lb = O.left_step(sd, md, mps.rp_to_lp(lays[k], vecs[k]), vecs[k], Lb, True)
        Rb = O.trivial_boundary(bonds[6], mpo[5].col_dim)
        for k in range(5, 2, -1):
            sd = E.SiteDesc(basis.sym, lays[k].left, lays[k].phys, lays[k].right)
            md = E.MpoDesc(mpo[k] if herm else strip(mpo[k]))
            Rb = O.right_step(sd, md, vecs[k], vecs[k], Rb, True)
        res.append((O.dense_boundary(Lb), O.dense_boundary(Rb)))
    for a, b in zip(res[0], res[1]):
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


def test_oracle_ninej_matches_sympy_fixture():
    rows = json.load(open(os.path.join(GOLDEN, "ninej_sympy.json")))["rows"]
    err = max(abs(O.wigner9j(*r[:9]) - r[9]) for r in rows)
    assert err <= 1e-14
    # a triangle violation gives zero
    assert O.wigner9j(0, 0, 2, 0, 0, 0, 0, 0, 2) == 0.0


def test_oracle_ninej_special_cases():
    """Closed forms: with one zero spin the 9j reduces to a 6j; {1/2..} known value."""
    # {a b c; d e f; g h 0} = (-1)^(b+c+d+g) / sqrt((2c+1)(2g+1)) {a b c; e d g}  (c == f, g == h)
    from sympy import Rational
    from sympy.physics.wigner import wigner_6j
    for (a, b, c, d, e, g) in [(1, 1, 2, 1, 1, 0), (2, 1, 1, 1, 2, 1), (2, 2, 2, 2, 2, 2), (3, 1, 2, 1, 3, 2)]:
        lhs = O.wigner9j(a, b, c, d, e, c, g, g, 0)
        six = float(wigner_6j(*[Rational(x, 2) for x in (a, b, c, e, d, g)]))
        sign = (-1) ** ((b + c + d + g) // 2)
        assert abs(lhs - sign * six / np.sqrt((c + 1) * (g + 1))) <= 1e-14


def test_product_su2_coupling_matches_oracle():
    L = capi.lib()
    out = (C.c_double * 4)()
    rows = json.load(open(os.path.join(GOLDEN, "ninej_sympy.json")))["rows"]
    for r in rows[::7]:
        assert abs(L.mq_wigner9j(*r[:9]) - r[9]) <= 1e-14
    for a, c, g, i in [(1, 1, 1, 1), (2, 2, 1, 3), (4, 3, 5, 4), (0, 1, 1, 0), (3, 3, 3, 3)]:
        for d, e, f in [(0, 1, 1), (1, 1, 0), (2, 1, 1), (1, 2, 1), (0, 0, 0)]:
            capi.check(L.mq_su2_coupling(a, abs(a - c), c, d, e, f, g, abs(g - i), i, 0.7, out))
            ref = O.set_coupling(a, abs(a - c), c, d, e, f, g, abs(g - i), i, 0.7)
            assert np.allclose(list(out), ref, rtol=0, atol=1e-14)


def test_oracle_subspace_eigensolver():
    rng = np.random.default_rng(1)
    for n in [1, 2, 5, 8]:
        A = rng.standard_normal((n, n))
        A = A + A.T
        val, vec = O.subspace_min(A)
        w, v = np.linalg.eigh(A)
        assert abs(val - w[0]) <= 1e-12
        assert min(np.linalg.norm(vec - v[:, 0]), np.linalg.norm(vec + v[:, 0])) <= 1e-10


def test_oracle_jd_converges_to_lowest_eigenvalue():
    p = SiteProblem("heisenberg", L=6, site=2, Mmax=4, seed=9)
    so = O.SigmaOracle(p.site_desc, p.mpo_desc, p.Lb, p.Rb)
    n = p.layout.rp_total
    # explicit H_eff by applying sigma to unit vectors of the valid (non-padding) entries
    idx = np.nonzero(p.psi)[0]
    H = np.zeros((len(idx), len(idx)))
    for j, k in enumerate(idx):
        e = np.zeros(n)
        e[k] = 1.0
        H[:, j] = so.apply(e)[idx]
    assert np.allclose(H, H.T, atol=1e-10)
    e0 = np.linalg.eigvalsh(H)[0]
    e, v, it = so.jd(p.psi, tol=1e-12, max_iter=40)
    assert abs(e - e0) <= 1e-9


# --------------------------------------------------------------------------- product planner
@pytest.mark.parametrize("model,L,site,M,seed", [("heisenberg", 8, 3, 6, 1), ("hubbard", 6, 2, 6, 2),
                                                 ("bosehubbard", 8, 3, 6, 3), ("hubbard", 8, 0, 8, 4)])
def test_product_sigma_schedule_equals_oracle_schedule(model, L, site, M, seed):
    p = SiteProblem(model, L=L, site=site, Mmax=M, seed=seed)
    sp, fp, _ = E.build_schedule("sigma", p.site_desc, p.mpo_desc, p.Lb.index, p.Rb.index)
    so, fo = O.schedule_sigma(p.site_desc, p.mpo_desc, p.Lb, p.Rb)
    assert fp == fo
    for k in so:
        if isinstance(so[k], np.ndarray):
            np.testing.assert_array_equal(sp[k], so[k], err_msg=k)


@pytest.mark.parametrize("model", ["heisenberg", "hubbard"])
def test_product_step_indices_equal_oracle(model):
    p = SiteProblem(model, L=8, site=3, Mmax=6, seed=5)
    k = 4
    lay = p.layouts[k]
    sd = E.SiteDesc(p.sym, lay.left, lay.phys, lay.right)
    md = E.MpoDesc(p.mpo[k])
    _, _, ri = E.build_schedule("right", sd, md, bindex_r=p.Rb.index)
    ro = O.right_step(sd, md, p.vecs[k], p.vecs[k], p.Rb, True)
    np.testing.assert_array_equal(ri.offsets, ro.index.offsets)
    np.testing.assert_array_equal(ri.tr, ro.index.tr)
    np.testing.assert_array_equal(ri.n_blocks, ro.index.n_blocks)
    np.testing.assert_array_equal(ri.conj_scales, ro.index.conj_scales)


def _same_bindex(a, b):
    assert a.n_cohorts == b.n_cohorts
    for k in ("bra_pos", "ket_pos", "tr", "n_blocks", "offsets", "conj_scales", "transposes"):
        np.testing.assert_array_equal(getattr(a, k), getattr(b, k), err_msg=k)


@pytest.mark.parametrize("model,symmetric", [("heisenberg", True), ("heisenberg", False), ("hubbard", True),
                                             ("qc", True), ("bosehubbard", True)])
def test_index_only_steps_equal_full_builders(model, symmetric):
    """The index-only left / right steps give the same boundary index as the full schedules, at
    every site of a chain (the right step always complements, so it runs with symmetric=True)."""
    p = SiteProblem(model, L=8, site=3, Mmax=8, seed=9)
    import mqdmrg.workloads as W
    left = W.trivial_bindex(p.bonds[0], p.mpo[0].row_dim)
    right = W.trivial_bindex(p.bonds[p.L], p.mpo[p.L - 1].col_dim)
    for k in range(p.L):
        lay = p.layouts[k]
        sd = E.SiteDesc(p.sym, lay.left, lay.phys, lay.right)
        md = E.MpoDesc(p.mpo[k])
        _, _, full = E.build_schedule("left", sd, md, bindex_l=left, symmetric=symmetric)
        fast = E.step_index("left", sd, md, left, symmetric=symmetric)
        _same_bindex(full, fast)
        left = full
    for k in range(p.L - 1, -1, -1):
        lay = p.layouts[k]
        sd = E.SiteDesc(p.sym, lay.left, lay.phys, lay.right)
        md = E.MpoDesc(p.mpo[k])
        _, _, full = E.build_schedule("right", sd, md, bindex_r=right)
        fast = E.step_index("right", sd, md, right)
        _same_bindex(full, fast)
        right = full


def test_oracle_block_sample_is_additive():
    """The bounded CPU sample (sigma of a subset of MPS blocks) partitions sigma and its flops."""
    p = SiteProblem("hubbard", L=8, site=3, Mmax=10, seed=5)
    full = O.SigmaOracle(p.site_desc, p.mpo_desc, p.Lb, p.Rb)
    n = p.site_desc.desc.left_i.n
    a = O.SigmaOracle(p.site_desc, p.mpo_desc, p.Lb, p.Rb, only_blocks=range(0, n, 2))
    b = O.SigmaOracle(p.site_desc, p.mpo_desc, p.Lb, p.Rb, only_blocks=range(1, n, 2))
    r = full.apply(p.psi)
    np.testing.assert_allclose(a.apply(p.psi) + b.apply(p.psi), r, rtol=0, atol=1e-13 * np.abs(r).max())
    assert a.flops + b.flops == full.flops and 0 < a.flops < full.flops


def test_planner_rejects_mismatched_boundary():
    p = SiteProblem("heisenberg", L=8, site=3, Mmax=6, seed=1)
    q = SiteProblem("heisenberg", L=8, site=2, Mmax=6, seed=1)
    with pytest.raises(capi.EngineError, match="does not match"):
        E.build_schedule("sigma", p.site_desc, p.mpo_desc, p.Lb.index, q.Rb.index)


def test_planner_rejects_bad_descriptors():
    p = SiteProblem("heisenberg", L=6, site=2, Mmax=4, seed=1)
    W = p.W12
    md = E.MpoDesc(W)
    md.desc.term_b1[0] = W.row_dim + 5    # out of range bond index
    with pytest.raises(capi.EngineError, match="out of range"):
        E.build_schedule("sigma", p.site_desc, md, p.Lb.index, p.Rb.index)


def test_truncation_rule():
    """estimate_truncation: cut = max(cutoff*s_max, s[Mmax]); keep s >= cut per block."""
    s = [np.array([1.0, 0.5, 1e-12]), np.array([0.7, 0.5]), np.array([])]
    keeps, tw = truncate(s, Mmax=10, cutoff=1e-10)
    assert keeps == [2, 2, 0]
    keeps, _ = truncate(s, Mmax=1, cutoff=1e-10)   # cut = s[1] = 0.7
    assert keeps == [1, 1, 0]
    keeps, _ = truncate(s, Mmax=2, cutoff=1e-10)   # cut = s[2] = 0.5: the tie keeps both 0.5's
    assert keeps == [2, 2, 0]
    keeps, _ = truncate(s, Mmax=10, cutoff=0.6)    # relative cutoff 0.6 * s_max
    assert keeps == [1, 1, 0]


def _graded(rng, n, w, smin_exp, rank=None):
    r = min(n, w) if rank is None else rank
    U, _ = np.linalg.qr(rng.standard_normal((n, r)))
    V, _ = np.linalg.qr(rng.standard_normal((w, r)))
    s = np.logspace(0, smin_exp, r) * rng.uniform(0.5, 1.0)
    return (U * s) @ V.T


def _sector_mats(seed):
    rng = np.random.default_rng(seed)
    shapes = {0: (37, 20, -13, None), 1: (20, 37, -9, None), 2: (64, 64, -6, 10), 3: (1, 1, 0, None),
              4: (0, 5, 0, None), 5: (90, 3, -2, None), 6: (3, 90, -2, None), 7: (200, 150, -12, None),
              8: (33, 40, -11, 25)}
    return {c: torch.from_numpy(_graded(rng, n, w, e, r)) if n and w else torch.zeros(n, w, dtype=torch.float64)
            for c, (n, w, e, r) in shapes.items()}


@pytest.mark.parametrize("direction", [1, -1])
def test_gram_spectrum_and_factors_against_lapack_svd(direction):
    """The batched Gram-eigh factorisation of the split against numpy's SVD (LAPACK gesdd) on
    graded spectra down to 1e-13 of the largest value, sectors wider and taller than they are long
    (Gram on either side, so both the eigenvector and the Cholesky-QR branch), ragged sizes sharing
    a batch, rank-deficient, empty and 1x1 sectors: each singular value within the Gram bound
    8 eps s_max^2 / s above GRAM_RESOLVED s_max; kept factors with Q_k^T Q_k = 1 to 1e-14 and
    Q_k B_k equal to LAPACK's rank-k truncation."""
    mats = _sector_mats(11)
    keys = list(mats)
    state, sv = DM.gram_spectrum(mats, keys, direction, "cpu")
    ref = {c: np.linalg.svd(mats[c].numpy(), compute_uv=False) if mats[c].numel() else np.zeros(0) for c in keys}
    smax = max(v[0] for v in ref.values() if len(v))
    eps = np.finfo(np.float64).eps
    keeps = {}
    for c in keys:
        s = sv[c]
        assert s.shape == ref[c].shape and np.all(np.diff(s) <= 0)
        big = ref[c] >= DM.GRAM_RESOLVED * smax
        assert np.all(np.abs(s - ref[c])[big] <= 8 * eps * smax ** 2 / ref[c][big]), c
        keeps[c] = int(np.sum(ref[c] >= 1e-4 * smax))
    assert sum(keeps.values()) > 30
    fac = DM.gram_truncated(state, keeps, "cpu")
    assert set(fac) == {c for c, k in keeps.items() if k}
    for c, (Q, B) in fac.items():
        k = keeps[c]
        Q, B = Q.numpy(), B.numpy()
        assert Q.shape[1] == k and B.shape[0] == k
        assert np.abs(Q.T @ Q - np.eye(k)).max() <= 1e-14, c
        U, S, Vh = np.linalg.svd(mats[c].numpy(), full_matrices=False)
        trunc = (U[:, :k] * S[:k]) @ Vh[:k]
        got = Q @ B if direction > 0 else (Q @ B).T
        assert np.abs(got - trunc).max() <= 1e-12 * smax, c


@pytest.mark.parametrize("direction", [1, -1])
@pytest.mark.parametrize("Mmax,cutoff,method", [(40, 1e-10, "gram"), (12, 1e-10, "gram"), (1000, 1e-3, "gram"),
                                                (1000, 1e-10, "svd"), (1000, 1e-7, "svd")])
def test_split_factors_truncate_like_lapack_svd(direction, Mmax, cutoff, method):
    """split_factors keeps the same number of states per sector as a LAPACK SVD with the
    reference's truncation rule, the same truncated weight, orthonormal Q_k and the same truncated
    product U_k S_k V_k^T (to 1e-12 of s_max), on the Gram path when the threshold is resolved and
    on the SVD fallback when it is not."""
    mats = _sector_mats(12)
    keys = list(mats)
    fac, keeps, svals, tw, used = DM.split_factors(mats, keys, direction, "cpu", Mmax, cutoff)
    assert used == method
    ref_s = [np.linalg.svd(mats[c].numpy(), compute_uv=False) if mats[c].numel() else np.zeros(0) for c in keys]
    k_ref, tw_ref = truncate(ref_s, Mmax, cutoff)
    smax = max(v[0] for v in ref_s if len(v))
    assert keeps == k_ref and abs(tw - tw_ref) <= 1e-13
    for c, kc in zip(keys, keeps):
        if kc == 0:
            assert c not in fac
            continue
        U, S, Vh = np.linalg.svd(mats[c].numpy(), full_matrices=False)
        ref = (U[:, :kc] * S[:kc]) @ Vh[:kc]
        Q, B = (t.numpy() for t in fac[c])
        assert np.abs(Q.T @ Q - np.eye(kc)).max() <= 1e-13
        got = Q @ B if direction > 0 else (Q @ B).T
        assert np.abs(got - ref).max() <= 1e-12 * smax, (c, kc)


def test_slab_copy_matches_blockwise_copies():
    """slab_copy (the merge / split data movement) against one slice copy per block, for row-major
    and column-major slabs, empty slabs and a slab list with no elements."""
    rng = np.random.default_rng(5)
    src = torch.from_numpy(rng.standard_normal(4000))
    dst, ref = torch.zeros(3000, dtype=torch.float64), torch.zeros(3000, dtype=torch.float64)
    slabs = [(0, 1, 7, 100, 9, 1, 7, 5),      # column-major target, row-major source
             (40, 11, 1, 0, 1, 13, 4, 11),    # row-major target, column-major source
             (500, 1, 3, 2000, 1, 3, 3, 0),   # empty
             (900, 1, 20, 3000, 20, 1, 20, 30)]
    for d0, dsi, dsj, s0, ssi, ssj, ni, nj in slabs:
        for i in range(ni):
            for j in range(nj):
                ref[d0 + i * dsi + j * dsj] = src[s0 + i * ssi + j * ssj]
    DM.slab_copy(dst, src, slabs)
    assert torch.equal(dst, ref)
    DM.slab_copy(dst, src, [])
    assert torch.equal(dst, ref)


def test_identity_mpo_is_the_identity():
    """models.identity_mpo: bond dimension 1 and the identity operator on every site (the MPO of
    TwoSiteDMRG.norm2)."""
    from mqdmrg import models
    basis, _, _ = models.chain("hubbard", 3, U=4.0)
    mpo = models.identity_mpo(basis.phys, 3)
    assert len(mpo) == 3
    for W in mpo:
        D = W.dense(basis.phys)
        assert D.shape[:2] == (1, 1)
        assert np.array_equal(D[0, 0], np.eye(basis.phys.total()))


def test_gram_buckets_cover_every_size_once():
    sizes = [3, 700, 64, 65, 300, 290, 31, 1, 1200, 256]
    groups = DM._gram_buckets(sizes)
    seen = sorted(p for _, ps in groups for p in ps)
    assert seen == list(range(len(sizes)))
    for nb, ps in groups:
        assert nb % 32 == 0 and all(sizes[p] <= nb for p in ps)


# --------------------------------------------------------------------------- C-ABI surface
def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "mq_sigma.h")).read()
    declared = set(re.findall(r"^\s*(?:int|double|void\*|const char\*)\s+(mq_\w+)\s*\(", hdr, re.M))
    assert len(declared) >= 40
    lib = capi.lib()
    missing = [n for n in declared if not hasattr(lib, n)]
    assert not missing, missing
    assert declared <= set(capi.SIGNATURES), declared - set(capi.SIGNATURES)


def test_error_codes_not_exit():
    L = capi.lib()
    assert L.mq_sigma_apply(None, None, None) == 1
    assert b"bad argument" in L.mq_last_error()


def test_wrappers_reject_bad_buffers_before_any_c_call(monkeypatch):
    """A float32 buffer would be read as n doubles (out of bounds on the device) and a host
    tensor would hand a host pointer to the kernels: the wrappers reject both before calling into
    the library."""
    class FakeEngine:
        torch_device = torch.device("cuda:0")
        ctx = None

    class NoCalls:
        def __getattr__(self, name):
            def call(*args):
                raise AssertionError("C-ABI %s called with an unchecked buffer" % name)
            return call
    monkeypatch.setattr(E, "lib", lambda: NoCalls())
    eng = FakeEngine()
    plan = E.SigmaPlan.__new__(E.SigmaPlan)
    plan.engine, plan.n, plan.h = eng, 8, None
    good_shape = torch.zeros(8, dtype=torch.float64)
    with pytest.raises(ValueError, match="float64"):
        plan.apply(torch.zeros(8, dtype=torch.float32), good_shape)
    with pytest.raises(ValueError, match="cuda:0"):
        plan.apply(good_shape, good_shape)
    with pytest.raises(TypeError):
        E.dev_ptr(eng, np.zeros(8), 8)
    with pytest.raises(ValueError):
        E.Engine.axpy(eng, 1.0, good_shape, good_shape)
    with pytest.raises(ValueError):
        E.apply_batched(eng, [plan], [torch.zeros(4, dtype=torch.float64)], [good_shape])


def test_oracle_jd_gmres_and_ortho_vectors():
    """The oracle's Jacobi-Davidson with the GMRES correction reaches the lowest eigenvalue of the
    dense H_eff; with the ground state as ortho vector it reaches the second one, orthogonal to it."""
    p = SiteProblem("heisenberg", L=8, site=3, Mmax=6, seed=21)
    so = O.SigmaOracle(p.site_desc, p.mpo_desc, p.Lb, p.Rb)
    lay = p.layout
    idx = np.concatenate([np.arange(lay.rp_offset[b], lay.rp_offset[b] + lay.left.sizes[b] * lay.rp_cols[b])
                          for b in range(len(lay.left.sizes))])
    H = np.zeros((len(idx), len(idx)))
    for c, i in enumerate(idx):
        e = np.zeros(lay.rp_total)
        e[i] = 1.0
        H[:, c] = so.apply(e)[idx]
    w = np.linalg.eigvalsh(0.5 * (H + H.T))
    e0, v0, it0 = so.jd(p.psi, tol=1e-11, max_iter=40)
    e1, v1, it1 = so.jd(p.psi, tol=1e-11, max_iter=40, gmres_iter=4)
    assert abs(e0 - w[0]) <= 1e-9 * abs(w[0]) and abs(e1 - w[0]) <= 1e-9 * abs(w[0]), (e0, e1, w[0])
    e2, v2, it2 = so.jd(p.psi, tol=1e-11, max_iter=60, ortho=[v0])
    assert abs(e2 - w[1]) <= 1e-8 * abs(w[1]), (e2, w[1])
    assert abs(float(v2 @ v0)) <= 1e-10
