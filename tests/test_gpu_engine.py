"""GPU parity of the HIP engine against the CPU oracle (called through the C-ABI).

Tolerances: sigma and boundary steps are FP64 and must agree to 1e-12 relative (the bar the
engine is specified to); Jacobi-Davidson energies to 1e-10 absolute."""
import numpy as np
import pytest
import torch

from chain import SiteProblem
from oracle import oracle as O
from mqdmrg import engine as E, mps
from mqdmrg.symmetry import SiteLayout

pytestmark = pytest.mark.gpu

SIGMA_RTOL = 1e-12
# Reduced-precision sigma: FP32 products carry ~1e-7 relative rounding per term, the bf16 split
# x = hi + lo leaves ~2^-17 relative per operand; sums over a few hundred terms stay well inside
# these bounds.
MIXED_RTOL = {E.PREC_F32: 1e-5, E.PREC_BF16X3: 1e-4}


@pytest.fixture(scope="module")
def eng():
    e = E.Engine(0)
    yield e
    e.close()


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda:0")


def upload(eng, hb):
    return E.DeviceBoundary.upload(eng, hb.index, hb.cohorts)


CASES = [
    ("heisenberg", 8, 3, 6, 1),
    ("heisenberg", 10, 0, 8, 2),     # left edge (trivial left boundary)
    ("heisenberg", 10, 8, 8, 3),     # right edge
    ("hubbard", 6, 2, 6, 4),
    ("hubbard", 8, 3, 10, 5),
    ("bosehubbard", 8, 3, 6, 8),
    ("qc", 6, 2, 8, 6),
    ("qc", 8, 3, 10, 7),
]


@pytest.mark.parametrize("algorithm", [E.SIGMA_SCHEDULE, E.SIGMA_GROUPED])
@pytest.mark.parametrize("model,L,site,M,seed", CASES)
def test_sigma_matches_oracle(eng, model, L, site, M, seed, algorithm):
    print("case", model, L, site, M, seed, algorithm, flush=True)
    p = SiteProblem(model, L=L, site=site, Mmax=M, seed=seed)
    so = O.SigmaOracle(p.site_desc, p.mpo_desc, p.Lb, p.Rb)
    ref = so.apply(p.psi)
    Lb, Rb = upload(eng, p.Lb), upload(eng, p.Rb)
    plan = E.SigmaPlan(eng, p.site_desc, p.mpo_desc, Lb, Rb, algorithm=algorithm)
    assert plan.n == p.layout.rp_total
    if algorithm == E.SIGMA_SCHEDULE:
        assert plan.flops == so.flops
    psi = dev(p.psi)
    out = torch.full_like(psi, np.nan)
    plan.apply(psi, out)
    eng.sync()
    got = out.cpu().numpy()
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err <= SIGMA_RTOL, err
    # run twice: bitwise deterministic (no atomics)
    out2 = torch.zeros_like(psi)
    plan.apply(psi, out2)
    eng.sync()
    assert torch.equal(out, out2)


@pytest.mark.parametrize("model,L,m", [("heisenberg", 24, 512), ("qc", 12, 128), ("hubbard", 16, 256)])
def test_sigma_synthetic_shapes_match_oracle(eng, model, L, m):
    """Middle-of-chain shapes (blocks from 2 to ~200 rows; MPO bond up to ~430) reach every GEMM
    tile configuration: 128x128, 64x64, 32x128, 16x128 and the split-K 16x16 tiles."""
    from mqdmrg.workloads import SyntheticSite
    print("case", model, L, m, flush=True)
    ws = SyntheticSite(model, L, L // 2 - 1, m, seed=3)
    Lc, Rc = ws.random_cohorts(ws.left_index), ws.random_cohorts(ws.right_index)
    hl = O.HostBoundary(ws.left_index, [c if c is not None else np.zeros(0) for c in Lc])
    hr = O.HostBoundary(ws.right_index, [c if c is not None else np.zeros(0) for c in Rc])
    psi_h = ws.random_psi()
    ref = O.SigmaOracle(ws.site_desc, ws.mpo_desc, hl, hr).apply(psi_h)
    Lb = E.DeviceBoundary.upload(eng, ws.left_index, Lc)
    Rb = E.DeviceBoundary.upload(eng, ws.right_index, Rc)
    psi = dev(psi_h)
    for alg in (E.SIGMA_SCHEDULE, E.SIGMA_GROUPED):
        plan = E.SigmaPlan(eng, ws.site_desc, ws.mpo_desc, Lb, Rb, algorithm=alg)
        precs = [E.PREC_F64] if alg == E.SIGMA_SCHEDULE else [E.PREC_F64, E.PREC_F32, E.PREC_BF16X3]
        for prec in precs:
            out = torch.full_like(psi, np.nan)
            plan.apply(psi, out, precision=prec)
            eng.sync()
            err = np.linalg.norm(out.cpu().numpy() - ref) / np.linalg.norm(ref)
            print("  algorithm", alg, "precision", prec, "relative error", err, flush=True)
            assert err <= (SIGMA_RTOL if prec == E.PREC_F64 else MIXED_RTOL[prec]), (alg, prec, err)
        plan.close()


def test_autotuned_plan_matches_oracle(eng):
    """Tile configurations chosen by the autotuner change the schedule of the T and final stages,
    not the result: sigma still matches the oracle (FP64 to 1e-13 of the untuned result); a second
    plan of the same block-size histogram takes the cached decision without timing; graphs captured
    after tuning replay to the tuned eager result."""
    from mqdmrg.workloads import SyntheticSite
    ws = SyntheticSite("qc", 12, 5, 128, seed=3)
    Lc, Rc = ws.random_cohorts(ws.left_index), ws.random_cohorts(ws.right_index)
    hl = O.HostBoundary(ws.left_index, [c if c is not None else np.zeros(0) for c in Lc])
    hr = O.HostBoundary(ws.right_index, [c if c is not None else np.zeros(0) for c in Rc])
    psi_h = ws.random_psi()
    ref = O.SigmaOracle(ws.site_desc, ws.mpo_desc, hl, hr).apply(psi_h)
    Lb = E.DeviceBoundary.upload(eng, ws.left_index, Lc)
    Rb = E.DeviceBoundary.upload(eng, ws.right_index, Rc)
    psi = dev(psi_h)
    for alg in (E.SIGMA_SCHEDULE, E.SIGMA_GROUPED):
        precs = [E.PREC_F64] if alg == E.SIGMA_SCHEDULE else [E.PREC_F64, E.PREC_F32, E.PREC_BF16X3]
        for prec in precs:
            print("case", alg, prec, flush=True)
            plan = E.SigmaPlan(eng, ws.site_desc, ws.mpo_desc, Lb, Rb, algorithm=alg)
            plan.set_graphs(False)
            before = torch.full_like(psi, np.nan)
            plan.apply(psi, before, precision=prec)
            assert plan.autotune(prec) > 0
            after = torch.full_like(psi, np.nan)
            plan.apply(psi, after, precision=prec)
            eng.sync()
            got = after.cpu().numpy()
            err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
            assert err <= (SIGMA_RTOL if prec == E.PREC_F64 else MIXED_RTOL[prec]), (alg, prec, err)
            if prec == E.PREC_F64:
                d = np.linalg.norm(got - before.cpu().numpy()) / np.linalg.norm(ref)
                assert d <= 1e-13, d
            plan.set_graphs(True)
            for rep in range(3):
                out = torch.full_like(psi, np.nan)
                plan.apply(psi, out, precision=prec)
                eng.sync()
                assert torch.equal(out, after), rep
            twin = E.SigmaPlan(eng, ws.site_desc, ws.mpo_desc, Lb, Rb, algorithm=alg)
            assert twin.autotune(prec) == 0   # same histogram: cached decision
            out = torch.full_like(psi, np.nan)
            twin.apply(psi, out, precision=prec)
            eng.sync()
            assert torch.equal(out, after)
            twin.close()
            plan.close()


@pytest.mark.parametrize("prec", [E.PREC_F32, E.PREC_BF16X3])
def test_library_solutions_adopted_by_shape_class(prec, monkeypatch):
    """A library GEMM solution tuned on one plan is adopted, through the shape-class cache, by the
    GEMMs of a plan of nearby shapes (the same chain at a slightly larger bond dimension) when the
    library supports them: that plan's library-only autotune times fewer candidates than the first
    plan's, and its sigma still matches the oracle to the precision's tolerance.  The plans run
    every S stage on the library (MQ_S_DENSE=1), so FP32 has library GEMMs to tune."""
    from mqdmrg.workloads import SyntheticSite
    monkeypatch.setenv("MQ_S_DENSE", "1")
    e = E.Engine(0)
    try:
        counts = []
        for m, seed in ((128, 3), (136, 4)):
            ws = SyntheticSite("qc", 12, 5, m, seed=seed)
            Lc, Rc = ws.random_cohorts(ws.left_index), ws.random_cohorts(ws.right_index)
            hl = O.HostBoundary(ws.left_index, [c if c is not None else np.zeros(0) for c in Lc])
            hr = O.HostBoundary(ws.right_index, [c if c is not None else np.zeros(0) for c in Rc])
            psi_h = ws.random_psi()
            ref = O.SigmaOracle(ws.site_desc, ws.mpo_desc, hl, hr).apply(psi_h)
            Lb = E.DeviceBoundary.upload(e, ws.left_index, Lc)
            Rb = E.DeviceBoundary.upload(e, ws.right_index, Rc)
            plan = E.SigmaPlan(e, ws.site_desc, ws.mpo_desc, Lb, Rb, algorithm=E.SIGMA_GROUPED)
            print("case m", m, "precision", prec, flush=True)
            counts.append(plan.autotune(prec, library_only=True))
            out = torch.full_like(dev(psi_h), np.nan)
            plan.apply(dev(psi_h), out, precision=prec)
            e.sync()
            err = np.linalg.norm(out.cpu().numpy() - ref) / np.linalg.norm(ref)
            assert err <= MIXED_RTOL[prec], (m, err)
            assert plan.autotune(prec, library_only=True) == 0   # every GEMM now cached
            plan.close()
        assert counts[0] > 0 and counts[1] < counts[0], counts
    finally:
        e.close()


def test_sigma_plan_from_caller_schedule(eng):
    """Drop-in seam: the caller hands its own ScheduleNew (here the oracle's) to the engine."""
    p = SiteProblem("hubbard", L=6, site=2, Mmax=6, seed=7)
    ref = O.SigmaOracle(p.site_desc, p.mpo_desc, p.Lb, p.Rb).apply(p.psi)
    sched, _ = O.schedule_sigma(p.site_desc, p.mpo_desc, p.Lb, p.Rb)
    Lb, Rb = upload(eng, p.Lb), upload(eng, p.Rb)
    rows = np.asarray(p.layout.left.sizes, np.uint32)
    cols = np.asarray(p.layout.rp_cols, np.uint32)
    plan = E.SigmaPlan.from_schedule(eng, sched, rows, cols, Lb, Rb)
    psi = dev(p.psi)
    out = torch.zeros_like(psi)
    plan.apply(psi, out)
    eng.sync()
    err = np.linalg.norm(out.cpu().numpy() - ref) / np.linalg.norm(ref)
    assert err <= SIGMA_RTOL, err


def _cmp_boundaries(hb, db):
    assert db.index.n_cohorts == hb.index.n_cohorts
    np.testing.assert_array_equal(db.index.offsets, hb.index.offsets)
    np.testing.assert_array_equal(db.index.tr, hb.index.tr)
    np.testing.assert_allclose(db.index.conj_scales, hb.index.conj_scales, rtol=0, atol=0)
    got = db.download()
    num = sum(float(np.sum((g - h) ** 2)) for g, h in zip(got, hb.cohorts))
    den = sum(float(np.sum(h ** 2)) for h in hb.cohorts)
    assert np.sqrt(num / max(den, 1e-300)) <= SIGMA_RTOL


STEP_ALGS = [E.SIGMA_SCHEDULE, E.SIGMA_GROUPED]


@pytest.mark.parametrize("algorithm", STEP_ALGS)
@pytest.mark.parametrize("model,symmetric", [("heisenberg", True), ("heisenberg", False), ("hubbard", True),
                                             ("hubbard", False), ("qc", True), ("bosehubbard", True)])
def test_left_step_matches_oracle(eng, model, symmetric, algorithm):
    print("case", model, symmetric, algorithm, flush=True)
    p = SiteProblem(model, L=8, site=4, Mmax=8, seed=11)
    k = 4
    lay = p.layouts[k]
    sd = E.SiteDesc(p.sym, lay.left, lay.phys, lay.right)
    md = E.MpoDesc(p.mpo[k])
    bra_lp = mps.rp_to_lp(lay, p.vecs[k])
    ref = O.left_step(sd, md, bra_lp, p.vecs[k], p.Lb, symmetric)
    Lb = upload(eng, p.Lb)
    got = E.left_step(eng, sd, md, dev(bra_lp), dev(p.vecs[k]), Lb, symmetric, algorithm=algorithm)
    _cmp_boundaries(ref, got)


@pytest.mark.parametrize("algorithm", STEP_ALGS)
@pytest.mark.parametrize("model,site", [("heisenberg", 3), ("hubbard", 3), ("qc", 3), ("bosehubbard", 3),
                                        ("heisenberg", 0), ("qc", 5)])
def test_right_step_matches_oracle(eng, model, site, algorithm):
    print("case", model, site, algorithm, flush=True)
    p = SiteProblem(model, L=8, site=site, Mmax=8, seed=12)
    k = site + 1   # p.Rb is the boundary of sites >= site + 2
    lay = p.layouts[k]
    sd = E.SiteDesc(p.sym, lay.left, lay.phys, lay.right)
    md = E.MpoDesc(p.mpo[k])
    ref = O.right_step(sd, md, p.vecs[k], p.vecs[k], p.Rb, True)
    Rb = upload(eng, p.Rb)
    v = dev(p.vecs[k])
    got = E.right_step(eng, sd, md, v, v, Rb, True, algorithm=algorithm)
    _cmp_boundaries(ref, got)


@pytest.mark.parametrize("model,L,m", [("qc", 12, 128), ("hubbard", 16, 256)])
def test_grouped_steps_match_schedule_steps_at_synthetic_shapes(eng, model, L, m):
    """Middle-of-chain shapes: the grouped left and right steps equal the schedule steps (random
    boundary data and site tensor; the oracle is too slow here, the schedule path is pinned by it
    at the small cases above)."""
    from mqdmrg.workloads import SyntheticSite
    print("case", model, L, m, flush=True)
    ws = SyntheticSite(model, L, L // 2 - 1, m, seed=3)
    sym, phys = ws.basis.sym, ws.basis.phys
    rng = np.random.default_rng(1)
    Lb = E.DeviceBoundary.upload(eng, ws.left_index, ws.random_cohorts(ws.left_index))
    Rb = E.DeviceBoundary.upload(eng, ws.right_index, ws.random_cohorts(ws.right_index))
    for k, side in ((ws.site, "left"), (ws.site + 1, "right")):
        sd = E.SiteDesc(sym, ws.bonds[k], phys, ws.bonds[k + 1])
        md = E.MpoDesc(ws.mpo[k])
        lay = SiteLayout(ws.bonds[k], phys, ws.bonds[k + 1])
        dv = dev(rng.standard_normal(max(lay.rp_total, 1)))
        if side == "left":
            lp = torch.zeros(max(lay.lp_total, 1), dtype=torch.float64, device="cuda:0")
            E.reshape_right_to_left(eng, sd, dv, lp)
            a = E.left_step(eng, sd, md, lp, dv, Lb, True, algorithm=E.SIGMA_SCHEDULE)
            b = E.left_step(eng, sd, md, lp, dv, Lb, True, algorithm=E.SIGMA_GROUPED)
        else:
            a = E.right_step(eng, sd, md, dv, dv, Rb, True, algorithm=E.SIGMA_SCHEDULE)
            b = E.right_step(eng, sd, md, dv, dv, Rb, True, algorithm=E.SIGMA_GROUPED)
        print("  ", side, "cohorts", a.index.n_cohorts, flush=True)
        _cmp_boundaries(O.HostBoundary(a.index, a.download()), b)


@pytest.mark.parametrize("model,site", [("qc", 3), ("hubbard", 3)])
def test_grouped_steps_in_many_chunks_match_oracle(eng, model, site, monkeypatch):
    """A workspace budget of one element (MQ_STEP_BUDGET_ELEMS=1) runs every ket sector of a grouped
    step as its own chunk, reusing the context's step workspace across chunks without clearing it;
    the new boundaries equal the one-chunk steps to 1e-14 and the oracle to 1e-12."""
    print("case", model, site, flush=True)
    p = SiteProblem(model, L=8, site=site, Mmax=8, seed=21)
    k = site + 1
    lay = p.layouts[k]
    sd = E.SiteDesc(p.sym, lay.left, lay.phys, lay.right)
    md = E.MpoDesc(p.mpo[k])
    v = dev(p.vecs[k])
    Rb = upload(eng, p.Rb)
    one_r = E.right_step(eng, sd, md, v, v, Rb, True, algorithm=E.SIGMA_GROUPED)
    kl = site
    layl = p.layouts[kl]
    sdl = E.SiteDesc(p.sym, layl.left, layl.phys, layl.right)
    mdl = E.MpoDesc(p.mpo[kl])
    bra_lp = mps.rp_to_lp(layl, p.vecs[kl])
    Lb = upload(eng, p.Lb)
    one_l = E.left_step(eng, sdl, mdl, dev(bra_lp), dev(p.vecs[kl]), Lb, True, algorithm=E.SIGMA_GROUPED)
    assert len(lay.left) > 1 and len(layl.right) > 1   # several chunks
    monkeypatch.setenv("MQ_STEP_BUDGET_ELEMS", "1")
    many_r = E.right_step(eng, sd, md, v, v, Rb, True, algorithm=E.SIGMA_GROUPED)
    many_l = E.left_step(eng, sdl, mdl, dev(bra_lp), dev(p.vecs[kl]), Lb, True, algorithm=E.SIGMA_GROUPED)
    monkeypatch.delenv("MQ_STEP_BUDGET_ELEMS")
    _cmp_boundaries(O.right_step(sd, md, p.vecs[k], p.vecs[k], p.Rb, True), many_r)
    _cmp_boundaries(O.left_step(sdl, mdl, bra_lp, p.vecs[kl], p.Lb, True), many_l)
    for a, b in ((one_r, many_r), (one_l, many_l)):
        np.testing.assert_array_equal(a.index.offsets, b.index.offsets)
        for x, y in zip(a.download(), b.download()):
            assert np.abs(x - y).max(initial=0.0) <= 1e-14 * max(1.0, np.abs(x).max(initial=0.0))


def test_reshape_roundtrip(eng):
    p = SiteProblem("hubbard", L=6, site=2, Mmax=6, seed=13)
    lay = p.layouts[2]
    sd = E.SiteDesc(p.sym, lay.left, lay.phys, lay.right)
    rp = dev(p.vecs[2])
    lp = torch.zeros(lay.lp_total, dtype=torch.float64, device="cuda:0")
    E.reshape_right_to_left(eng, sd, rp, lp)
    np.testing.assert_array_equal(lp.cpu().numpy(), mps.rp_to_lp(lay, p.vecs[2]))
    back = torch.zeros_like(rp)
    E.reshape_left_to_right(eng, sd, lp, back)
    np.testing.assert_array_equal(back.cpu().numpy(), p.vecs[2])


def test_vector_kernels(eng):
    rng = np.random.default_rng(5)
    for n in [1, 127, 128, 100003, 3 * 2 ** 20 + 5]:
        x, y = rng.standard_normal(n), rng.standard_normal(n)
        dx, dy = dev(x), dev(y)
        assert abs(eng.dot(dx, dy) - x @ y) <= 1e-12 * max(1.0, abs(x) @ abs(y))
        assert abs(eng.nrm2(dx) - np.linalg.norm(x)) <= 1e-12 * np.linalg.norm(x)
        eng.axpy(0.5, dx, dy)
        eng.scal(-2.0, dx)
        eng.sync()
        np.testing.assert_allclose(dy.cpu().numpy(), y + 0.5 * x, rtol=0, atol=1e-14)
        np.testing.assert_array_equal(dx.cpu().numpy(), -2.0 * x)


@pytest.mark.parametrize("model", ["heisenberg", "hubbard", "qc"])
def test_jacobi_davidson_matches_oracle(eng, model):
    p = SiteProblem(model, L=8, site=3, Mmax=6, seed=21)
    so = O.SigmaOracle(p.site_desc, p.mpo_desc, p.Lb, p.Rb)
    e_ref, v_ref, it_ref = so.jd(p.psi, tol=1e-10, max_iter=8)
    Lb, Rb = upload(eng, p.Lb), upload(eng, p.Rb)
    plan = E.SigmaPlan(eng, p.site_desc, p.mpo_desc, Lb, Rb)
    psi = dev(p.psi)
    out = torch.zeros_like(psi)
    e, it = plan.jd(psi, out, tol=1e-10, max_iter=8)
    assert it == it_ref
    assert abs(e - e_ref) <= 1e-10 * max(1.0, abs(e_ref))
    v = out.cpu().numpy()
    assert min(np.linalg.norm(v - v_ref), np.linalg.norm(v + v_ref)) <= 1e-8


@pytest.mark.parametrize("model", ["heisenberg", "qc"])
def test_jacobi_davidson_gmres_and_ortho_vectors(eng, model):
    """The GMRES correction (gmres_iter > 0) and ortho vectors follow the oracle's restatement of
    ietl's GMRES-JD (same iteration count, energies to 1e-10 relative); on the Heisenberg block the
    GMRES run reaches the lowest eigenvalue of the dense H_eff, and with the ground state as ortho
    vector the second one, orthogonal to it (excited states, solver.h:65-69).  (On the ab initio block
    with random boundaries neither correction converges in 30 steps; the trajectories still agree.)"""
    p = SiteProblem(model, L=8, site=3, Mmax=6, seed=21)
    so = O.SigmaOracle(p.site_desc, p.mpo_desc, p.Lb, p.Rb)
    Lb, Rb = upload(eng, p.Lb), upload(eng, p.Rb)
    plan = E.SigmaPlan(eng, p.site_desc, p.mpo_desc, Lb, Rb)
    psi = dev(p.psi)
    out0, out1, out2 = torch.zeros_like(psi), torch.zeros_like(psi), torch.zeros_like(psi)
    e0, it0 = plan.jd(psi, out0, tol=1e-11, max_iter=30)
    e1, it1 = plan.jd(psi, out1, tol=1e-11, max_iter=30, gmres_iter=4)
    r1, _, rit1 = so.jd(p.psi, tol=1e-11, max_iter=30, gmres_iter=4)
    print("plain: E %.12f in %d; gmres: E %.12f in %d (oracle %.12f in %d)" % (e0, it0, e1, it1, r1, rit1), flush=True)
    assert it1 == rit1 and abs(e1 - r1) <= 1e-10 * abs(r1)
    e2, it2 = plan.jd(psi, out2, tol=1e-10, max_iter=60, ortho=[out0])
    r2, v2, rit2 = so.jd(p.psi, tol=1e-10, max_iter=60, ortho=[out0.cpu().numpy()])
    ov = float(torch.dot(out2, out0))
    print("with ortho: E %.12f in %d (oracle %.12f in %d), overlap %.2e" % (e2, it2, r2, rit2, ov), flush=True)
    assert it2 == rit2 and abs(e2 - r2) <= 1e-10 * abs(r2)
    assert abs(ov) <= 1e-9 * float(torch.dot(out0, out0))
    if model == "heisenberg":
        lay = p.layout
        idx = np.concatenate([np.arange(lay.rp_offset[b], lay.rp_offset[b] + lay.left.sizes[b] * lay.rp_cols[b])
                              for b in range(len(lay.left.sizes))])
        H = np.zeros((len(idx), len(idx)))
        for c, i in enumerate(idx):
            x = np.zeros(lay.rp_total)
            x[i] = 1.0
            H[:, c] = so.apply(x)[idx]
        w = np.linalg.eigvalsh(0.5 * (H + H.T))
        assert abs(e1 - w[0]) <= 1e-9 * abs(w[0]) and abs(e2 - w[1]) <= 1e-8 * abs(w[1]), (e1, e2, w[:2])


@pytest.mark.parametrize("model,L,m", [("qc_su2", 24, 1000), ("qc", 20, 512)])
def test_block_sparse_s_stage_matches_dense_and_oracle(eng, model, L, m, monkeypatch):
    """Ab initio middle-of-chain shapes: the block-sparse S stage (keys in pattern-class order, the
    s_stage.hip kernel over the nonzero 16-key tiles of every group's W_D) does fewer flops than the
    dense library GEMMs of the same plan built with MQ_S_DENSE=1; its sigma equals the dense plan's
    to 1e-13 (FP64), 2e-6 (FP32) and 2e-5 (BF16x3, against the dense FP64 sigma), runs bitwise
    reproducibly, and equals the oracle's on a block sample to 1e-12."""
    from mqdmrg.workloads import SyntheticSite
    print("case", model, L, m, flush=True)
    ws = SyntheticSite(model, L, L // 2 - 1, m, seed=9)
    Lc, Rc = ws.random_cohorts(ws.left_index), ws.random_cohorts(ws.right_index)
    Lb = E.DeviceBoundary.upload(eng, ws.left_index, Lc)
    Rb = E.DeviceBoundary.upload(eng, ws.right_index, Rc)
    psi_h = ws.random_psi()
    psi = dev(psi_h)
    sparse = E.SigmaPlan(eng, ws.site_desc, ws.mpo_desc, Lb, Rb, algorithm=E.SIGMA_GROUPED)
    monkeypatch.setenv("MQ_S_DENSE", "1")
    dense = E.SigmaPlan(eng, ws.site_desc, ws.mpo_desc, Lb, Rb, algorithm=E.SIGMA_GROUPED)
    monkeypatch.delenv("MQ_S_DENSE")
    tmp = torch.zeros_like(psi)
    _, fl_sparse, _ = sparse.profile(psi, tmp, reps=1)
    _, fl_dense, _ = dense.profile(psi, tmp, reps=1)
    print("  S-stage flops sparse %.3e dense %.3e" % (fl_sparse[1], fl_dense[1]), flush=True)
    assert fl_sparse[1] < 0.9 * fl_dense[1]
    outs = {}
    for name, plan in (("sparse", sparse), ("dense", dense)):
        for prec in (E.PREC_F64, E.PREC_F32, E.PREC_BF16X3):
            out = torch.full_like(psi, np.nan)
            plan.apply(psi, out, precision=prec)
            eng.sync()
            outs[name, prec] = out.cpu().numpy()
            again = torch.full_like(psi, np.nan)
            plan.apply(psi, again, precision=prec)
            eng.sync()
            assert torch.equal(out, again), (name, prec)
    ref = outs["dense", E.PREC_F64]
    nr = np.linalg.norm(ref)
    e64 = np.linalg.norm(outs["sparse", E.PREC_F64] - ref) / nr
    e32 = np.linalg.norm(outs["sparse", E.PREC_F32] - ref) / nr
    ebf = np.linalg.norm(outs["sparse", E.PREC_BF16X3] - ref) / nr
    ebf_lib = np.linalg.norm(outs["dense", E.PREC_BF16X3] - ref) / nr
    print("  sparse vs dense FP64: FP64 %.2e FP32 %.2e BF16x3 %.2e (library BF16x3 %.2e)" % (e64, e32, ebf, ebf_lib),
          flush=True)
    assert e64 <= 1e-13 and e32 <= 2e-6 and ebf <= 2e-5
    lay = ws.layout
    nblk = len(lay.left.sizes)
    blocks = sorted(np.random.default_rng(0).choice(nblk, size=min(6, nblk), replace=False).tolist())
    hl = O.HostBoundary(ws.left_index, [c if c is not None else np.zeros(0) for c in Lc])
    hr = O.HostBoundary(ws.right_index, [c if c is not None else np.zeros(0) for c in Rc])
    # the oracle's sigma of the sampled ket blocks alone against the GPU sigma of psi restricted to them
    oref = O.SigmaOracle(ws.site_desc, ws.mpo_desc, hl, hr, only_blocks=blocks).apply(psi_h)
    masked = np.zeros_like(psi_h)
    for b in blocks:
        o, n = lay.rp_offset[b], lay.left.sizes[b] * lay.rp_cols[b]
        masked[o:o + n] = psi_h[o:o + n]
    out = torch.full_like(psi, np.nan)
    sparse.apply(dev(masked), out)
    eng.sync()
    eo = np.linalg.norm(out.cpu().numpy() - oref) / np.linalg.norm(oref)
    print("  sparse vs oracle, ket blocks", blocks, "%.2e" % eo, flush=True)
    assert eo <= 1e-12


@pytest.mark.parametrize("model,L,m", [("qc_su2", 16, 256), ("qc", 12, 128)])
def test_reused_workspace_needs_no_zero_fill(eng, model, L, m, monkeypatch):
    """Grouped sigma plans and grouped boundary steps take a reused workspace without clearing it.
    With MQ_POISON_WS=1 a reused workspace (and split-K partial buffer) is filled with NaN instead:
    a plan built on it must give sigma bitwise equal to the same plan built before, in every
    precision, and the left and right steps of the site bitwise equal boundaries, so no kernel
    reads a workspace element it has not written in the same application."""
    from mqdmrg.workloads import SyntheticSite
    print("case", model, L, m, flush=True)
    ws = SyntheticSite(model, L, L // 2 - 1, m, seed=4)
    Lb = E.DeviceBoundary.upload(eng, ws.left_index, ws.random_cohorts(ws.left_index))
    Rb = E.DeviceBoundary.upload(eng, ws.right_index, ws.random_cohorts(ws.right_index))
    psi = dev(ws.random_psi())
    precs = (E.PREC_F64, E.PREC_F32, E.PREC_BF16X3)

    def run(plan):
        outs = {}
        for prec in precs:
            out = torch.full_like(psi, np.nan)
            plan.apply(psi, out, precision=prec)
            eng.sync()
            outs[prec] = out
        return outs

    def steps():
        sym, phys = ws.basis.sym, ws.basis.phys
        res = []
        for k, side in ((ws.site, "left"), (ws.site + 1, "right")):
            sd = E.SiteDesc(sym, ws.bonds[k], phys, ws.bonds[k + 1])
            md = E.MpoDesc(ws.mpo[k])
            lay = SiteLayout(ws.bonds[k], phys, ws.bonds[k + 1])
            dv = dev(np.random.default_rng(k).standard_normal(max(lay.rp_total, 1)))
            if side == "left":
                lp = torch.zeros(max(lay.lp_total, 1), dtype=torch.float64, device="cuda:0")
                E.reshape_right_to_left(eng, sd, dv, lp)
                b = E.left_step(eng, sd, md, lp, dv, Lb, True, algorithm=E.SIGMA_GROUPED)
            else:
                b = E.right_step(eng, sd, md, dv, dv, Rb, True, algorithm=E.SIGMA_GROUPED)
            res.append([np.asarray(c) for c in b.download()])
        return res

    first = E.SigmaPlan(eng, ws.site_desc, ws.mpo_desc, Lb, Rb, algorithm=E.SIGMA_GROUPED)
    ref = run(first)
    first.close()   # its workspace returns to the pool and is taken by the next plan
    ref_steps = steps()
    monkeypatch.setenv("MQ_POISON_WS", "1")
    poisoned = E.SigmaPlan(eng, ws.site_desc, ws.mpo_desc, Lb, Rb, algorithm=E.SIGMA_GROUPED)
    got = run(poisoned)
    poisoned.close()
    got_steps = steps()
    monkeypatch.delenv("MQ_POISON_WS")
    for prec in precs:
        assert torch.isfinite(got[prec]).all(), prec
        assert torch.equal(got[prec], ref[prec]), prec
    for a, b in zip(ref_steps, got_steps):
        assert len(a) == len(b)
        for x, y in zip(a, b):
            assert np.all(np.isfinite(y)) and np.array_equal(x, y)


@pytest.mark.parametrize("precision", [E.PREC_F32, E.PREC_BF16X3])
@pytest.mark.parametrize("model,L,site,M,seed", [("hubbard", 8, 3, 10, 5), ("qc", 8, 3, 10, 7),
                                                  ("heisenberg", 10, 0, 8, 2)])
def test_mixed_sigma_error(eng, model, L, site, M, seed, precision):
    print("case", model, L, site, M, seed, precision, flush=True)
    p = SiteProblem(model, L=L, site=site, Mmax=M, seed=seed)
    ref = O.SigmaOracle(p.site_desc, p.mpo_desc, p.Lb, p.Rb).apply(p.psi)
    Lb, Rb = upload(eng, p.Lb), upload(eng, p.Rb)
    plan = E.SigmaPlan(eng, p.site_desc, p.mpo_desc, Lb, Rb, algorithm=E.SIGMA_GROUPED)
    psi = dev(p.psi)
    out = torch.full_like(psi, np.nan)
    plan.apply(psi, out, precision=precision)
    eng.sync()
    got = out.cpu().numpy()
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print("relative error", err, flush=True)
    assert err <= MIXED_RTOL[precision], err
    assert err > 1e-13   # the reduced-precision kernels ran, not the FP64 ones
    # the FP64 path of the same plan is unaffected by the FP32 workspace
    plan.apply(psi, out)
    eng.sync()
    assert np.linalg.norm(out.cpu().numpy() - ref) <= SIGMA_RTOL * np.linalg.norm(ref)


@pytest.mark.parametrize("precision", [E.PREC_F32, E.PREC_BF16X3])
def test_reduced_precision_first_use_after_fp64(eng, precision):
    """A plan's FP32 workspace is allocated at its first reduced-precision application; device
    memory released before that (here filled with NaN) must not leak into sigma."""
    poison = torch.full((1 << 25,), np.nan, dtype=torch.float64, device="cuda:0")
    del poison
    torch.cuda.empty_cache()
    p = SiteProblem("qc", L=8, site=0, Mmax=64, seed=7)
    ref = O.SigmaOracle(p.site_desc, p.mpo_desc, p.Lb, p.Rb).apply(p.psi)
    Lb, Rb = upload(eng, p.Lb), upload(eng, p.Rb)
    plan = E.SigmaPlan(eng, p.site_desc, p.mpo_desc, Lb, Rb, algorithm=E.SIGMA_GROUPED)
    psi = dev(p.psi)
    out = torch.zeros_like(psi)
    plan.apply(psi, out)
    plan.apply(psi, out, precision=precision)
    eng.sync()
    got = out.cpu().numpy()
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert np.isfinite(got).all() and err <= MIXED_RTOL[precision], err


def test_mixed_precision_rejected_on_schedule_plan(eng):
    p = SiteProblem("hubbard", L=6, site=2, Mmax=6, seed=4)
    Lb, Rb = upload(eng, p.Lb), upload(eng, p.Rb)
    plan = E.SigmaPlan(eng, p.site_desc, p.mpo_desc, Lb, Rb, algorithm=E.SIGMA_SCHEDULE)
    psi = dev(p.psi)
    out = torch.zeros_like(psi)
    with pytest.raises(RuntimeError, match="GROUPED"):
        plan.apply(psi, out, precision=E.PREC_F32)


def _dense_lowest(so, layout):
    """Lowest eigenvalue of H_eff assembled column by column from the oracle's sigma, over the
    stored (non-padding) entries of the padded layout."""
    live = np.concatenate([off + np.arange(ls * c) for off, ls, c in
                           zip(layout.rp_offset, layout.left.sizes, layout.rp_cols)]).astype(np.int64)
    H = np.empty((live.size, live.size))
    e = np.zeros(layout.rp_total)
    for j, k in enumerate(live):
        e[k] = 1.0
        H[:, j] = so.apply(e)[live]
        e[k] = 0.0
    return np.linalg.eigvalsh(0.5 * (H + H.T))[0]


@pytest.mark.parametrize("precision", [E.PREC_F32, E.PREC_BF16X3])
@pytest.mark.parametrize("model", ["heisenberg", "qc"])
def test_refined_jd_reaches_fp64_energy(eng, model, precision):
    """Iterative refinement over the reduced-precision sigma converges to the FP64 eigenvalue, on
    problems where the plain FP64 solver converges within the same budget (checked first)."""
    p = SiteProblem(model, L=8, site=3, Mmax=6, seed=21)
    so = O.SigmaOracle(p.site_desc, p.mpo_desc, p.Lb, p.Rb)
    assert p.layout.rp_total <= 4096
    e_exact = _dense_lowest(so, p.layout)
    Lb, Rb = upload(eng, p.Lb), upload(eng, p.Rb)
    plan = E.SigmaPlan(eng, p.site_desc, p.mpo_desc, Lb, Rb, algorithm=E.SIGMA_GROUPED)
    psi = dev(p.psi)
    out = torch.zeros_like(psi)
    tol, max_iter, max_outer = 1e-9, 32, 6
    e64, c64, rn64 = plan.jd_refined(psi, out, tol=tol, max_iter=max_iter, precision=E.PREC_F64,
                                     max_outer=max_outer)
    assert rn64 <= tol * abs(e64), (e64, c64, rn64)
    e, counts, rn = plan.jd_refined(psi, out, tol=tol, max_iter=max_iter, precision=precision, max_outer=max_outer)
    print("refined", model, precision, e, e64, e_exact, counts, c64, rn, flush=True)
    assert rn <= tol * abs(e) or rn < tol
    # these boundaries come from random MPS, so the spectrum's scale is arbitrary: relative bar
    assert abs(e - e_exact) <= 1e-9 * max(1.0, abs(e_exact))
    assert counts["fp64"] == counts["outer"] and counts["reduced"] >= counts["fp64"]
    # the returned vector is the normalised Ritz vector whose FP64 Rayleigh quotient is e
    v = out.cpu().numpy()
    assert abs(np.linalg.norm(v) - 1.0) <= 1e-12
    assert abs(v @ so.apply(v) - e) <= 1e-10 * max(1.0, abs(e))


@pytest.mark.parametrize("precision", [E.PREC_F64, E.PREC_F32])
def test_graph_replay_equals_eager(eng, precision):
    """Captured applications replay to the eager result, bit for bit, on every replay, with eager
    work and synchronisations of the program between the replays."""
    p = SiteProblem("hubbard", L=8, site=3, Mmax=10, seed=5)
    Lb, Rb = upload(eng, p.Lb), upload(eng, p.Rb)
    plan = E.SigmaPlan(eng, p.site_desc, p.mpo_desc, Lb, Rb, algorithm=E.SIGMA_GROUPED)
    psi = dev(p.psi)
    psi2 = dev(p.psi[::-1].copy())
    plan.set_graphs(False)
    eager = torch.zeros_like(psi)
    plan.apply(psi, eager, precision=precision)
    eager2 = torch.zeros_like(psi)
    plan.apply(psi2, eager2, precision=precision)
    eng.sync()
    plan.set_graphs(True)
    out = torch.zeros_like(psi)
    out2 = torch.zeros_like(psi)
    for rep in range(4):
        out.fill_(np.nan)
        plan.apply(psi, out, precision=precision)      # captured on the first pass, replayed after
        eng.sync()
        assert torch.equal(out, eager), rep
        plan.set_graphs(False) if rep == 2 else None  # eager work between replays
        plan.apply(psi2, out2, precision=precision)
        eng.sync()
        assert torch.equal(out2, eager2), rep
        plan.set_graphs(True)


@pytest.mark.parametrize("precision", [E.PREC_F64, E.PREC_BF16X3])
def test_batched_apply_equals_individual(eng, precision):
    """Independent instances (different models and chains) applied as one batch, eagerly and then as
    a captured graph replayed several times, equal their individual applications bit for bit."""
    probs = [SiteProblem("hubbard", L=8, site=3, Mmax=10, seed=5), SiteProblem("qc", L=8, site=3, Mmax=8, seed=7),
             SiteProblem("heisenberg", L=10, site=0, Mmax=8, seed=2), SiteProblem("qc", L=6, site=2, Mmax=8, seed=6),
             SiteProblem("bosehubbard", L=8, site=3, Mmax=6, seed=8)]
    plans, psis, refs = [], [], []
    for p in probs:
        Lb, Rb = upload(eng, p.Lb), upload(eng, p.Rb)
        plan = E.SigmaPlan(eng, p.site_desc, p.mpo_desc, Lb, Rb, algorithm=E.SIGMA_GROUPED)
        psi = dev(p.psi)
        ref = torch.zeros_like(psi)
        plan.set_graphs(False)
        plan.apply(psi, ref, precision=precision)
        plans.append(plan)
        psis.append(psi)
        refs.append(ref)
    eng.sync()
    outs = [torch.zeros_like(x) for x in psis]
    for rep in range(4):   # eager, capture, replay, replay
        for o in outs:
            o.fill_(np.nan)
        E.apply_batched(eng, plans, psis, outs, precision=precision)
        eng.sync()
        for i, (o, r) in enumerate(zip(outs, refs)):
            assert torch.equal(o, r), (rep, i)


@pytest.mark.parametrize("model", ["qc_su2", "qc"])
def test_batched_apply_at_batch_shapes_equals_individual(eng, model):
    """16 independent ab initio instances at the batch workload's shapes (L uniform in [12, 24],
    m=256): the FP64 batch (library GEMMs of the S stage on four side streams, one rocBLAS handle
    and workspace per stream) equals the individual applications bit for bit, eagerly and on
    graph replays."""
    from mqdmrg.workloads import SyntheticSite
    rng = np.random.default_rng(7)
    plans, psis, refs, keep = [], [], [], []
    for i in range(16):
        L = int(rng.integers(12, 25))
        print("instance", i, model, L, flush=True)
        ws = SyntheticSite(model, L, L // 2 - 1, 256, seed=100 + i)
        lb = E.DeviceBoundary.upload(eng, ws.left_index, ws.random_cohorts(ws.left_index))
        rb = E.DeviceBoundary.upload(eng, ws.right_index, ws.random_cohorts(ws.right_index))
        plan = E.SigmaPlan(eng, ws.site_desc, ws.mpo_desc, lb, rb, algorithm=E.SIGMA_GROUPED)
        psi = dev(ws.random_psi())
        ref = torch.zeros_like(psi)
        plan.set_graphs(False)
        plan.apply(psi, ref)
        plans.append(plan)
        psis.append(psi)
        refs.append(ref)
        keep += [lb, rb]
    eng.sync()
    outs = [torch.zeros_like(x) for x in psis]
    for rep in range(3):   # eager, capture, replay
        for o in outs:
            o.fill_(np.nan)
        E.apply_batched(eng, plans, psis, outs)
        eng.sync()
        for i, (o, r) in enumerate(zip(outs, refs)):
            assert torch.equal(o, r), (rep, i)
    for p in plans:
        p.close()


def test_multi_dot_matches_numpy(eng):
    """Batched dot products over k contiguous vectors (more than one pass of 16) equal numpy's to
    rounding, and repeat bitwise."""
    rng = np.random.default_rng(4)
    n, k, ldv = 100003, 37, 100032
    V = rng.standard_normal(k * ldv)
    x = rng.standard_normal(n)
    got = eng.multi_dot(dev(V), k, ldv, dev(x))
    ref = np.array([V[j * ldv:j * ldv + n] @ x for j in range(k)])
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.array_equal(got, eng.multi_dot(dev(V), k, ldv, dev(x)))
