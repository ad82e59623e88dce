"""MPS bond indices, random initial states and host-side layout conversions.

allowed_sectors follows the reference's mps_sectors.h:57-150 (left/right reachable charges, each
sector capped at Mmax); random_bond_dims gives a synthetic bond index of total size ~m whose
sector sizes follow the multiplicity of each charge (used to build benchmark shapes)."""
import math

import numpy as np

from .symmetry import Index, SiteLayout, fuse, sub, SYM_SU2U1


def _fuse_index(a, b, sign=+1, cap=None, physical=None):
    acc = {}
    for ca, sa in zip(a.charges, a.sizes):
        for cb, sb in zip(b.charges, b.sizes):
            c = fuse(ca, cb) if sign > 0 else sub(ca, cb)
            if physical is not None and not physical(c):
                continue
            acc[c] = acc.get(c, 0) + sa * sb
    if cap is not None:
        acc = {c: min(cap, s) for c, s in acc.items()}
    return acc


def physical_filter(sym):
    """Charges a bond may carry: SU2U1 spins are non-negative (mps_sectors.h:43-52)."""
    return (lambda c: c[1] >= 0) if sym == SYM_SU2U1 else None


def allowed_sectors(phys_list, right_end, Mmax, left_end=(0, 0), physical=None):
    """Bond indices 0..L of a chain with per-site physical indices phys_list (mps_sectors.h:57-150);
    physical: optional predicate on charges (SU2U1: non-negative spin), applied at every step."""
    L = len(phys_list)
    left = [None] * (L + 1)
    right = [None] * (L + 1)
    left[0] = {tuple(left_end): 1}
    for i in range(L):
        # exact (arbitrary precision) state counts
        left[i + 1] = _fuse_index(Index.from_dict(left[i]), phys_list[i], +1, cap=None, physical=physical)
    right[L] = {tuple(right_end): 1}
    for i in range(L - 1, -1, -1):
        right[i] = _fuse_index(Index.from_dict(right[i + 1]), phys_list[i], -1, cap=None, physical=physical)
    out = []
    for i in range(L + 1):
        common = {c: min(Mmax, left[i][c], right[i][c]) for c in left[i] if c in right[i]}
        out.append(Index.from_dict(common))
    return out, left, right


def random_bond_dims(phys_list, right_end, m, left_end=(0, 0), sigma=2.0, physical=None):
    """Synthetic bond indices of total dimension ~m.  The sector sizes follow a Gaussian in the
    charge distance (in units of the smallest charge step) from the most populated sector, of
    width sigma, the way the Schmidt spectrum of a ground state concentrates around the mean
    charge; each size is capped by the number of states both chain ends can reach."""
    _, left, right = allowed_sectors(phys_list, right_end, 1, left_end, physical=physical)
    L = len(phys_list)
    steps = [abs(c[k]) for I in phys_list for c in I.charges for k in (0, 1) if c[k] != 0]
    unit = min(steps) if steps else 1
    out = []
    for i in range(L + 1):
        common = {c: min(left[i][c], right[i][c]) for c in left[i] if c in right[i]}
        pop = {c: math.log(float(left[i][c])) + math.log(float(right[i][c])) for c in common}
        c0 = max(pop, key=pop.get)
        w = {c: math.exp(-(((c[0] - c0[0]) / unit) ** 2 + ((c[1] - c0[1]) / unit) ** 2) / (2 * sigma ** 2))
             for c in common}
        tot = sum(w.values())
        dims = {}
        for c, wc in w.items():
            d = int(round(m * wc / tot))
            d = min(d, common[c])
            if d > 0:
                dims[c] = d
        if not dims:
            c = max(w, key=w.get)
            dims[c] = 1
        out.append(Index.from_dict(dims))
    return out


def random_site_vector(layout, rng):
    """Random N(0,1) entries in every block of the right-paired padded layout (padding zero)."""
    v = np.zeros(layout.rp_total)
    for lb, ls in enumerate(layout.left.sizes):
        n = ls * layout.rp_cols[lb]
        v[layout.rp_offset[lb]:layout.rp_offset[lb] + n] = rng.standard_normal(n)
    return v


def rp_to_lp(layout, vec):
    """Host reshape right-paired -> left-paired (reshapes.h:197-242 inverse direction)."""
    out = np.zeros(layout.lp_total)
    L_, P_, R_ = layout.left, layout.phys, layout.right
    for lb, (lc, ls) in enumerate(zip(L_.charges, L_.sizes)):
        cols = layout.rp_cols[lb]
        blk = vec[layout.rp_offset[lb]:layout.rp_offset[lb] + ls * cols].reshape(cols, ls).T
        for s, (pc, d) in enumerate(zip(P_.charges, P_.sizes)):
            rc = fuse(lc, pc)
            rb = R_.position(rc)
            if rb == len(R_):
                continue
            rs = R_.sizes[rb]
            rows = layout.lp_rows[rb]
            lpblk = out[layout.lp_offset[rb]:layout.lp_offset[rb] + rows * rs].reshape(rs, rows).T
            c0 = layout.rpb(pc, rc)
            r0 = layout.lpb(pc, lc)
            for ss in range(d):
                lpblk[r0 + ss * ls:r0 + (ss + 1) * ls, :] = blk[:, c0 + ss * rs:c0 + (ss + 1) * rs]
    return out


def lp_to_rp(layout, vec):
    out = np.zeros(layout.rp_total)
    L_, P_, R_ = layout.left, layout.phys, layout.right
    for lb, (lc, ls) in enumerate(zip(L_.charges, L_.sizes)):
        cols = layout.rp_cols[lb]
        blk = out[layout.rp_offset[lb]:layout.rp_offset[lb] + ls * cols].reshape(cols, ls).T
        for s, (pc, d) in enumerate(zip(P_.charges, P_.sizes)):
            rc = fuse(lc, pc)
            rb = R_.position(rc)
            if rb == len(R_):
                continue
            rs = R_.sizes[rb]
            rows = layout.lp_rows[rb]
            lpblk = vec[layout.lp_offset[rb]:layout.lp_offset[rb] + rows * rs].reshape(rs, rows).T
            c0 = layout.rpb(pc, rc)
            r0 = layout.lpb(pc, lc)
            for ss in range(d):
                blk[:, c0 + ss * rs:c0 + (ss + 1) * rs] = lpblk[r0 + ss * ls:r0 + (ss + 1) * ls, :]
    return out


def layouts_for_chain(bonds, phys_list):
    return [SiteLayout(bonds[i], phys_list[i], bonds[i + 1]) for i in range(len(phys_list))]
