"""Test fixtures: small chains with random MPS tensors and boundaries built by the oracle."""
import numpy as np

from mqdmrg import models, mps
from mqdmrg.engine import MpoDesc, SiteDesc
from mqdmrg.symmetry import SiteLayout, Index
from oracle import oracle as O


class SiteProblem:
    """Two-site sigma problem at sites (i, i+1) of a chain, with host boundaries."""

    def __init__(self, model="heisenberg", L=8, site=3, Mmax=6, seed=1, symmetric=True, U=4.0):
        rng = np.random.default_rng(seed)
        kw = {"U": U} if model == "hubbard" else {}
        basis, mpo, target = models.chain(model, L, **kw)
        self.basis, self.mpo, self.L, self.site = basis, mpo, L, site
        phys = [basis.phys] * L
        bonds, _, _ = mps.allowed_sectors(phys, target, Mmax)
        self.bonds = bonds
        self.layouts = mps.layouts_for_chain(bonds, phys)
        self.vecs = [mps.random_site_vector(l, rng) for l in self.layouts]
        self.sym = basis.sym
        # left boundaries up to site, right boundaries down to site+2
        self.Lb = O.trivial_boundary(bonds[0], mpo[0].row_dim)
        for k in range(site):
            lay = self.layouts[k]
            sd = SiteDesc(self.sym, lay.left, lay.phys, lay.right)
            md = MpoDesc(mpo[k])
            self.Lb = O.left_step(sd, md, mps.rp_to_lp(lay, self.vecs[k]), self.vecs[k], self.Lb, symmetric)
        self.Rb = O.trivial_boundary(bonds[L], mpo[L - 1].col_dim)
        for k in range(L - 1, site + 1, -1):
            lay = self.layouts[k]
            sd = SiteDesc(self.sym, lay.left, lay.phys, lay.right)
            md = MpoDesc(mpo[k])
            self.Rb = O.right_step(sd, md, self.vecs[k], self.vecs[k], self.Rb, symmetric)
        W12, phys12, pb = models.twosite_sparse(mpo[site], mpo[site + 1], basis.phys, basis.phys)
        self.W12, self.phys12 = W12, phys12
        self.mpo_desc = MpoDesc(W12)
        self.layout = SiteLayout(bonds[site], phys12, bonds[site + 2])
        self.site_desc = SiteDesc(self.sym, bonds[site], phys12, bonds[site + 2])
        self.psi = mps.random_site_vector(self.layout, rng)
