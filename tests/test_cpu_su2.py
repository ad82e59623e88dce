"""CPU tests of the spin-adapted (SU2U1) path: Wigner symbols against exact sympy values, the
two-site reduction's orthogonality, and the oracle's SU2 sigma / left / right contractions with
the SU2U1 ab initio MPO pinned by exact energies of fixed total spin (determinant-space FCI with
the S^2 projection)."""
import json
import os

import numpy as np
import pytest

from su2_chain import SU2BlockProblem, exact_bonds
from oracle import oracle as O
from mqdmrg import qc, su2, mps
from mqdmrg.symmetry import Index, SiteLayout

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_wigner_symbols_match_sympy_fixture():
    """The front-end's 9j (three-6j sum over exact Racah 6j) equals sympy's exact values, and
    the oracle's independent 9j agrees."""
    rows = json.load(open(os.path.join(ROOT, "tests", "golden", "ninej_sympy.json")))["rows"]
    err = max(abs(su2.ninej(*r[:9]) - r[9]) for r in rows)
    assert err <= 1e-14
    err_o = max(abs(O.lib().mqo_wigner9j(*r[:9]) - r[9]) for r in rows[::7])
    assert err_o <= 1e-14


def test_sixj_closed_forms():
    """{a b c; 0 c b} = (-1)^(a+b+c) / sqrt((2b+1)(2c+1)) (doubled spins) and the orthogonality
    sum_x (2x+1)(2c+1) {a b x; d e c}{a b x; d e c'} = delta_cc'."""
    for a in range(0, 7):
        for b in range(0, 7):
            for c in range(abs(a - b), a + b + 1, 2):
                ref = (-1) ** ((a + b + c) // 2) / np.sqrt((b + 1.0) * (c + 1.0))
                assert abs(su2.sixj(a, b, c, 0, c, b) - ref) < 1e-14
    a, b, d, e = 3, 2, 4, 3
    cs = [c for c in range(0, 12) if su2._tri(a, e, c) and su2._tri(d, b, c)]
    for c in cs:
        for cp in cs:
            s = sum((x + 1.0) * (c + 1.0) * su2.sixj(a, b, x, d, e, c) * su2.sixj(a, b, x, d, e, cp)
                    for x in range(0, 12))
            assert abs(s - (1.0 if c == cp else 0.0)) < 1e-13


@pytest.mark.parametrize("L,N,two_s", [(4, 4, 0), (6, 6, 0), (5, 5, 1)])
def test_ts_reduction_is_orthogonal(L, N, two_s):
    """unreduce_right(reduce_right(x)) = x and ||reduce_right(x)|| = ||x|| on every two-site
    block of the chain (ts_reduction.h: the recoupling is an orthogonal change of basis)."""
    site = su2.su2_site()
    lb, rb = exact_bonds(site.phys, L, N, (N, two_s))
    phys12, _ = su2.product_index(site.phys, site.phys)
    rng = np.random.default_rng(3)
    for k in range(L - 1):
        left, right = lb[k], rb[k + 2]
        keep = {c: s for c, s in zip(left.charges, left.sizes) if rb[k].has(c)}
        left = Index.from_dict(keep)
        lay = SiteLayout(left, phys12, right)
        segs = su2.reduction_segments(left, site.phys, right)
        # a product of two site tensors has no weight on slabs whose coupling path passes a
        # negative spin; those are the slabs reduce_right skips
        x = np.zeros(lay.rp_total)
        full = mps.random_site_vector(lay, rng)
        for _, i, n, _ in segs:
            x[i:i + n] = full[i:i + n]
        red = su2.apply_segments_np(segs, x, lay.rp_total)
        back = su2.apply_segments_np(su2.unreduction_segments(left, site.phys, right), red, lay.rp_total)
        assert np.abs(back - x).max() <= 1e-14 * np.abs(x).max()
        assert abs(np.linalg.norm(red) - np.linalg.norm(x)) <= 1e-13 * np.linalg.norm(x)


@pytest.mark.parametrize("L,N,site,two_s", [(3, 2, 1, 0), (4, 4, 0, 0), (4, 4, 1, 0), (5, 4, 2, 0),
                                            (5, 5, 1, 1), (4, 4, 1, 2), (6, 6, 2, 0)])
def test_oracle_su2_heff_spectrum_equals_fci(L, N, site, two_s):
    """With complete bond spaces, the lowest eigenvalue of the oracle's two-site SU2 H_eff is the
    exact lowest energy of total spin two_s/2 (singlets, a doublet, a triplet; blocks at the left
    end, in the middle and at the right end)."""
    p = SU2BlockProblem(L, N, site, two_s=two_s)
    orc = O.SigmaOracle(p.site_desc, p.mpo_desc, p.Lb, p.Rb)
    H = p.dense_heff(orc.apply)
    assert np.abs(H - H.T).max() <= 1e-13 * np.abs(H).max()
    e = np.linalg.eigvalsh(0.5 * (H + H.T))[0]
    ef = su2.fci_spin_energy(p.h, p.eri, N, two_s)
    assert abs(e - ef) <= 1e-12 * max(1.0, abs(ef))


def test_su2_terms_cover_every_integral_class():
    """Every integral index class of the reference's term generator occurs at L=6 (so the FCI
    checks above exercise all of them), and no element is dropped."""
    h, eri = qc.synthetic_integrals(6, seed=2)
    elems = su2.integral_elements(h, eri, 0.0)
    terms, ec = su2.qc_su2_terms(elems, 6, 6)
    names = {n for key in terms for _, n in key}
    for n in ("count", "docc", "e2d", "d2e", "flip_to_S2", "flip_to_S0", "flip_S0", "count_fill",
              "create_fill", "destroy", "create_fill_couple_down", "destroy_couple_up", "create_count",
              "destroy_fill_count"):
        assert n in names, n
    assert max(len(k) for k in terms) == 4


@pytest.mark.parametrize("site", [0, 2, 3])
def test_twosite_su2_mpo_equals_the_direct_formula(site):
    """The two-site SU2 MPO (make_twosite_mpo, ts_ops.hpp:38-168) built with one sparse product per
    middle bond spin equals the direct sum over the three bond states: for every operator pair
    (o1, o2) and coupled rank k, W12(b1, b3) = sum_b2 f(j(b1), j(b2), j(b3)) C1(b1, b2) C2(b2, b3)
    with f = sqrt((j_m + 1)(k + 1)) {j j' k; k2 k1 j_m} (-1)^((j + j' + k1 + k2) / 2), evaluated here
    entry by entry on dense matrices; every nonzero (o1, o2, k) with a nonempty coupled operator is
    present, and nothing else."""
    from mqdmrg import models
    basis, mpo, _ = models.chain("qc_su2", 6)
    W1, W2 = mpo[site], mpo[site + 1]
    W12, phys12, pb = su2.twosite_su2(W1, W2, basis.phys)
    ls, ms, rs = W1.left_spin, W1.right_spin, W2.right_spin
    got = {op.name: C.toarray() for op, C in zip(W12.ops, W12.coeffs)}
    want = {}
    for o1, C1 in zip(W1.ops, W1.coeffs):
        A = C1.toarray()
        for o2, C2 in zip(W2.ops, W2.coeffs):
            B = C2.toarray()
            k1, k2 = o1.spin, o2.spin
            for k in range(abs(k1 - k2), k1 + k2 + 1, 2):
                M = np.zeros((A.shape[0], B.shape[1]))
                for b1, b2 in zip(*np.nonzero(A)):
                    for b3 in np.nonzero(B[b2])[0]:
                        j, jm, jp = ls[b1], ms[b2], rs[b3]
                        f = np.sqrt((jm + 1.0) * (k + 1.0)) * su2.sixj(j, jp, k, k2, k1, jm)
                        if ((j + jp + k1 + k2) // 2) % 2:
                            f = -f
                        M[b1, b3] += f * A[b1, b2] * B[b2, b3]
                if np.any(M != 0) and su2.op_kron_su2(o1, o2, k, basis.phys, phys12, pb).blocks:
                    want["%s*%s/%d" % (o1.name, o2.name, k)] = M
    assert set(got) == set(want)
    for name, M in want.items():
        np.testing.assert_allclose(got[name], M, rtol=0, atol=1e-13 * max(1.0, np.abs(M).max()), err_msg=name)
