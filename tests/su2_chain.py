"""Test fixtures for the spin-adapted (SU2U1) ab initio chain: exact (untruncated) bond spaces,
random isometries on both sides of a two-site block, boundaries built by the oracle, and the
dense effective Hamiltonian of the block.  With complete bond spaces the lowest eigenvalue of the
block's H_eff is the exact lowest energy of the chosen total spin, which pins the MPO, the
two-site MPO and every SU2 coupling of the sigma / left / right contractions at once."""
import numpy as np

from mqdmrg import qc, su2, mps
from mqdmrg.engine import MpoDesc, SiteDesc
from mqdmrg.symmetry import SiteLayout, Index, SYM_SU2U1, fuse, sub
from oracle import oracle as O


def exact_bonds(phys, L, N, target):
    """Bond indices reachable from the left (physical spins only) and from the right end."""
    left = [{(0, 0): 1}]
    for _ in range(L):
        acc = {}
        for lc, ls in left[-1].items():
            for pc in phys.charges:
                c = fuse(lc, pc)
                if c[1] >= 0 and c[0] <= N:
                    acc[c] = acc.get(c, 0) + ls
        left.append(acc)
    right = [None] * (L + 1)
    right[L] = {tuple(target): 1}
    for k in range(L - 1, -1, -1):
        acc = {}
        for rc, rs in right[k + 1].items():
            for pc in phys.charges:
                c = sub(rc, pc)
                if c[1] >= 0 and c[0] >= 0:
                    acc[c] = acc.get(c, 0) + rs
        right[k] = acc
    return [Index.from_dict(b) for b in left], [Index.from_dict(b) for b in right]


class SU2BlockProblem:
    """Two-site block at (site, site+1) of an L-orbital, N-electron chain of total spin two_s/2,
    with complete left / right bond spaces (unitary site tensors on both sides)."""

    def __init__(self, L, N, site, two_s=0, seed=1, integral_seed=None):
        rng = np.random.default_rng(seed)
        self.h, self.eri = qc.synthetic_integrals(L, seed=seed if integral_seed is None else integral_seed)
        basis, mpo, _ = su2.qc_su2_mpo(self.h, self.eri, N=N)
        self.basis, self.mpo, self.L, self.N, self.two_s, self.site = basis, mpo, L, N, two_s, site
        phys = basis.phys
        lb, rb = exact_bonds(phys, L, N, (N, two_s))
        # keep the sectors reachable from both ends: a bond index and the boundary built on it then
        # agree at every cut (a sector only one end reaches carries no weight of the target state)
        both = lambda a, b: Index.from_dict({c: n for c, n in zip(a.charges, a.sizes) if b.has(c)})
        lb, rb = [both(l, r) for l, r in zip(lb, rb)], [both(r, l) for l, r in zip(lb, rb)]
        self.Lb = O.trivial_boundary(lb[0], mpo[0].row_dim)
        self.left_sites, self.left_bounds = [], [self.Lb]
        for k in range(site):
            left, right = lb[k], lb[k + 1]
            lay = SiteLayout(left, phys, right)
            v = np.zeros(lay.lp_total)
            for b, rs in enumerate(right.sizes):
                rows = lay.lp_rows[b]
                q, _ = np.linalg.qr(rng.standard_normal((rows, rs)))
                v[lay.lp_offset[b]:lay.lp_offset[b] + rows * rs] = q.T.ravel()
            rp = mps.lp_to_rp(lay, v)
            sd = SiteDesc(SYM_SU2U1, left, phys, right)
            self.left_sites.append((sd, v, rp))
            self.Lb = O.left_step(sd, MpoDesc(mpo[k]), v, rp, self.Lb, True)
            self.left_bounds.append(self.Lb)
        self.Rb = O.trivial_boundary(rb[L], mpo[L - 1].col_dim)
        self.right_sites, self.right_bounds = [], [self.Rb]
        for k in range(L - 1, site + 1, -1):
            left, right = rb[k], rb[k + 1]
            lay = SiteLayout(left, phys, right)
            v = np.zeros(lay.rp_total)
            for b, ls in enumerate(left.sizes):
                cols = lay.rp_cols[b]
                q, _ = np.linalg.qr(rng.standard_normal((cols, ls)))
                v[lay.rp_offset[b]:lay.rp_offset[b] + ls * cols] = q.ravel()
            sd = SiteDesc(SYM_SU2U1, left, phys, right)
            self.right_sites.append((sd, v))
            self.Rb = O.right_step(sd, MpoDesc(mpo[k]), v, v, self.Rb, True)
            self.right_bounds.append(self.Rb)
        self.W12, self.phys12, self.pb = su2.twosite_su2(mpo[site], mpo[site + 1], phys)
        self.mpo_desc = MpoDesc(self.W12)
        left, right = lb[site], rb[site + 2]
        keep = {c: s for c, s in zip(left.charges, left.sizes)
                if any(right.has(fuse(c, p)) for p in self.phys12.charges)}
        self.left, self.right = Index.from_dict(keep), right
        self.site_desc = SiteDesc(SYM_SU2U1, self.left, self.phys12, self.right)
        self.layout = SiteLayout(self.left, self.phys12, self.right)
        self.psi = mps.random_site_vector(self.layout, rng)

    def active(self):
        """Positions of the stored (unpadded) elements of the block vector."""
        idx = []
        for b in range(len(self.left)):
            n = self.left.sizes[b] * self.layout.rp_cols[b]
            idx.extend(range(self.layout.rp_offset[b], self.layout.rp_offset[b] + n))
        return np.asarray(idx)

    def dense_heff(self, apply):
        """H_eff as a dense matrix over the active elements, column by column through apply."""
        idx = self.active()
        H = np.zeros((len(idx), len(idx)))
        for c, i in enumerate(idx):
            e = np.zeros(self.layout.rp_total)
            e[i] = 1.0
            H[:, c] = apply(e)[idx]
        return H
