"""Synthetic site problems at full benchmark shapes.

The boundary *indices* of a chain are propagated through the host planner (index-only left /
right steps, no device work); the boundary data, the MPS tensors and psi are random.  Sigma's cost
depends only on these shapes, so this gives the benchmark workload of a site in the middle of a
chain at bond dimension m without running sweeps first."""
import numpy as np

from . import models, mps, su2
from .capi import BIndex
from .engine import MpoDesc, SiteDesc, step_index
from .symmetry import SiteLayout, SYM_SU2U1


def trivial_bindex(index, aux_dim):
    """Index of Boundary(ud, ld, ad) (mp_tensors/boundary.h:258-283)."""
    n = len(index.charges)
    bs = [s * s for s in index.sizes]
    offs = np.asarray([[k * bs[i] for k in range(aux_dim)] for i in range(n)], np.int64).ravel()
    return BIndex(index, index, aux_dim, list(range(n)), list(range(n)), [0] * n, [aux_dim] * n, offs,
                  np.ones(n * aux_dim), np.zeros(n * aux_dim, np.uint8))


def model_chain(model, L, **kw):
    return models.chain(model, L, **kw)


class SyntheticSite:
    def __init__(self, model, L, site, m, seed=0, chain=None):
        """chain: a prebuilt (basis, mpo, target) of models.chain(model, L) to reuse."""
        self.model, self.L, self.site, self.m = model, L, site, m
        basis, mpo, target = chain if chain is not None else model_chain(model, L)
        self.chain = (basis, mpo, target)
        self.basis, self.mpo = basis, mpo
        phys = [basis.phys] * L
        sym = basis.sym
        self.bonds = mps.random_bond_dims(phys, target, m, physical=mps.physical_filter(sym))
        bi = trivial_bindex(self.bonds[0], mpo[0].row_dim)
        for k in range(site):
            sd = SiteDesc(sym, self.bonds[k], basis.phys, self.bonds[k + 1])
            bi = step_index("left", sd, MpoDesc(mpo[k]), bi)
        self.left_index = bi
        bi = trivial_bindex(self.bonds[L], mpo[L - 1].col_dim)
        for k in range(L - 1, site + 1, -1):
            sd = SiteDesc(sym, self.bonds[k], basis.phys, self.bonds[k + 1])
            bi = step_index("right", sd, MpoDesc(mpo[k]), bi)
        self.right_index = bi
        if sym == SYM_SU2U1:
            W12, phys12, _ = su2.twosite_su2(mpo[site], mpo[site + 1], basis.phys)
        else:
            W12, phys12, _ = models.twosite_sparse(mpo[site], mpo[site + 1], basis.phys, basis.phys)
        self.W12, self.phys12 = W12, phys12
        self.mpo_desc = MpoDesc(W12)
        self.layout = SiteLayout(self.bonds[site], phys12, self.bonds[site + 2])
        self.site_desc = SiteDesc(sym, self.bonds[site], phys12, self.bonds[site + 2])
        self.rng = np.random.default_rng(seed)

    def random_cohorts(self, bindex):
        out = []
        for ci in range(bindex.n_cohorts):
            n = bindex.cohort_size(ci)
            out.append(self.rng.standard_normal(n) if n else None)
        return out

    def random_psi(self):
        return mps.random_site_vector(self.layout, self.rng)
