"""Two-site DMRG sweeps on the engine (caller side of the sigma / boundary path).

Mirrors the reference's ts_optimize::sweep (optimize/ts_optimize.hpp:114-362): per site pair the
two-site tensor is formed (twositetensor.hpp:131-151), the effective Hamiltonian is solved with
Jacobi-Davidson (solver_interface.hpp:30-77 -> engine mq_jd_solve), the result is split by a
truncated SVD (split_mps_l2r / r2l, twositetensor.hpp:153-200 with estimate_truncation,
block_matrix_algorithms.h:135-183), and the boundary is moved (optimize.h:134-152 -> engine
left/right steps).  Site tensors live on the device in the right-paired padded layout; the small
dense algebra of the split (batched Gram matrices and eigh in place of per-sector SVDs) uses torch
on the engine's stream."""
import itertools
import time

import numpy as np
import torch

from . import engine as E
from . import mps as M
from . import models
from . import su2
from .symmetry import Index, SiteLayout, fuse, sub, product_index, SYM_SU2U1


def slab_copy(dst, src, slabs):
    """dst[d0 + i dsi + j dsj] = src[s0 + i ssi + j ssj] for i < ni, j < nj, for every slab
    (d0, dsi, dsj, s0, ssi, ssj, ni, nj) of `slabs` (flat tensors on one device; the slabs' targets
    are disjoint).  The element indices are expanded on the device and the copy is one gather and
    one scatter, instead of one small copy per block."""
    S = np.asarray(slabs, np.int64).reshape(-1, 8)
    total = int((S[:, 6] * S[:, 7]).sum()) if len(S) else 0
    if total == 0:
        return
    T = torch.from_numpy(S).to(dst.device)
    sz = T[:, 6] * T[:, 7]
    sid = torch.repeat_interleave(torch.arange(len(S), device=dst.device), sz, output_size=total)
    loc = torch.arange(total, device=dst.device) - (torch.cumsum(sz, 0) - sz)[sid]
    P = T[sid]
    i, j = loc % P[:, 6], loc // P[:, 6]
    dst[P[:, 0] + i * P[:, 1] + j * P[:, 2]] = src[P[:, 3] + i * P[:, 4] + j * P[:, 5]]


def apply_segments(x, segs):
    """out[o + t] = sum of c * x[i + t] over the segments (o, i, n, c), t < n (flat tensors on one
    device); at most two segments meet in one output element, so the sum is order independent."""
    out = torch.zeros_like(x)
    if not segs:
        return out
    S = torch.tensor([(o, i, n) for o, i, n, _ in segs], dtype=torch.int64, device=x.device)
    cf = torch.tensor([c for _, _, _, c in segs], dtype=x.dtype, device=x.device)
    n = S[:, 2]
    total = int(n.sum())
    sid = torch.repeat_interleave(torch.arange(len(segs), device=x.device), n, output_size=total)
    loc = torch.arange(total, device=x.device) - (torch.cumsum(n, 0) - n)[sid]
    out.index_add_(0, S[sid, 0] + loc, x[S[sid, 1] + loc] * cf[sid])
    return out


class SiteTensor:
    def __init__(self, eng, sym, left, phys, right, data=None):
        self.eng, self.sym = eng, sym
        self.left, self.phys, self.right = left, phys, right
        self.layout = SiteLayout(left, phys, right)
        self.desc = E.SiteDesc(sym, left, phys, right)
        n = max(self.layout.rp_total, 1)
        self.data = data if data is not None else torch.zeros(n, dtype=torch.float64, device=eng.torch_device)

    def rp_block(self, lb):
        ls, cols = self.left.sizes[lb], self.layout.rp_cols[lb]
        o = self.layout.rp_offset[lb]
        return self.data[o:o + ls * cols].view(cols, ls).t()

    def lp_data(self):
        out = torch.zeros(max(self.layout.lp_total, 1), dtype=torch.float64, device=self.eng.torch_device)
        E.reshape_right_to_left(self.eng, self.desc, self.data, out)
        return out

    def lp_block(self, lp, rb):
        rows, rs = self.layout.lp_rows[rb], self.right.sizes[rb]
        o = self.layout.lp_offset[rb]
        return lp[o:o + rows * rs].view(rs, rows).t()

    @staticmethod
    def from_lp(eng, sym, left, phys, right, blocks):
        """Site tensor from left-paired blocks {rb: (rows x rs) tensor}."""
        t = SiteTensor(eng, sym, left, phys, right)
        lp = torch.zeros(max(t.layout.lp_total, 1), dtype=torch.float64, device=eng.torch_device)
        for rb, blk in blocks.items():
            t.lp_block(lp, rb).copy_(blk)
        E.reshape_left_to_right(eng, t.desc, lp, t.data)
        return t

    @staticmethod
    def from_rp(eng, sym, left, phys, right, blocks):
        t = SiteTensor(eng, sym, left, phys, right)
        for lb, blk in blocks.items():
            t.rp_block(lb).copy_(blk)
        return t


def truncation_cut(svals, Mmax, cutoff):
    """The threshold of estimate_truncation (block_matrix_algorithms.h:135-183):
    max(cutoff * s_max, s[Mmax]) over all blocks' singular values, sorted descending."""
    allv = np.sort(np.concatenate([s for s in svals if len(s)]))[::-1]
    cut = cutoff * allv[0]
    if len(allv) > Mmax:
        cut = max(cut, allv[Mmax])
    return cut, allv


def truncate(svals, Mmax, cutoff):
    """estimate_truncation (block_matrix_algorithms.h:135-183): per-block keep counts."""
    cut, allv = truncation_cut(svals, Mmax, cutoff)
    keeps = [int(np.sum(s >= cut)) for s in svals]
    tw = float(np.sum(allv[allv < cut] ** 2) / np.sum(allv ** 2))
    return keeps, tw


def _gram_buckets(sizes):
    """Group Gram sizes into padded batches for torch.linalg.eigh.  A batched eigh call has a
    fixed cost that grows slowly with the padded size n (~1.5 + 0.02 n ms on MI355X) and a
    per-matrix cost ~0.08 (n/256)^2 ms, so a few large batches beat one call per size.  Sizes are
    rounded up to multiples of 32 and split into contiguous ranges by dynamic programming on that
    cost.  Returns a list of (padded size, [positions into sizes])."""
    order = sorted(range(len(sizes)), key=lambda i: sizes[i])
    rounded = [max(32, -(-sizes[i] // 32) * 32) for i in order]
    cost = lambda nb, cnt: 1.5 + 0.02 * nb + cnt * 0.08 * (nb / 256.0) ** 2
    K = len(order)
    best, cut = [0.0] + [float("inf")] * K, [0] * (K + 1)
    for j in range(1, K + 1):
        for i in range(j):
            c = best[i] + cost(rounded[j - 1], j - i)
            if c < best[j]:
                best[j], cut[j] = c, i
    groups, j = [], K
    while j > 0:
        i = cut[j]
        groups.append((rounded[j - 1], order[i:j]))
        j = i
    return groups[::-1]


def gram_spectrum(mats, keys, direction, device):
    """Singular values of the mid-sector matrices of a two-site split from batched Gram matrices.

    X is the sector matrix M (rows s1 x left, columns s2 x right) for a left-to-right split and
    M^T for a right-to-left one, so the factor that must come out orthonormal (U, or V for the
    right-normalised site) spans X's rows.  Per sector the Gram matrix of the smaller side,
    G = Y Y^T with Y = X (rows <= columns) or Y = X^T, is formed with batched GEMMs; all of them
    are diagonalised by a few padded batched eigh calls (one SVD per sector is instead a chain of
    small latency-bound kernels).  Z = P^T Y with P the eigenvectors; the singular values are the
    row norms of Z, the weights the vectors actually carry.  Padded diagonal entries are set to
    -(1 + trace G), far below the non-negative spectrum, so padding eigenvectors separate exactly
    and sort last.  The eigenvectors are orthonormal to working precision whatever the
    conditioning, but a singular value s carries an error of ~eps s_max^2 / s, so values below
    ~GRAM_RESOLVED s_max are not resolved (split_factors falls back to exact SVDs when the
    truncation threshold lies there).

    Returns (state, {c: singular values (numpy, descending, min(rows, cols) of them)})."""
    f64 = torch.float64
    xs = {c: (mats[c] if direction > 0 else mats[c].t()) for c in keys}
    svals, live = {}, []
    for c in keys:
        n, w = xs[c].shape
        if n == 0 or w == 0:
            svals[c] = np.zeros(0)
        else:
            live.append(c)
    buckets = []
    for nb, pos in _gram_buckets([min(xs[c].shape) for c in live]):
        cs = [live[p] for p in pos]
        trans = [xs[c].shape[0] > xs[c].shape[1] for c in cs]
        wb = max(max(xs[c].shape) for c in cs)
        nbat = len(cs)
        Y = torch.zeros(nbat, nb, wb, dtype=f64, device=device)
        for i, (c, t) in enumerate(zip(cs, trans)):
            y = xs[c].t() if t else xs[c]
            Y[i, :y.shape[0], :y.shape[1]] = y
        G = torch.bmm(Y, Y.transpose(1, 2).contiguous())
        r = torch.tensor([min(xs[c].shape) for c in cs], device=device)
        pos_n = torch.arange(nb, device=device)[None, :]
        diag = G.diagonal(dim1=1, dim2=2)
        trace = diag.sum(dim=1, keepdim=True)
        diag.add_(torch.where(pos_n >= r[:, None], -(1.0 + trace), torch.zeros_like(trace)))
        _, P = torch.linalg.eigh(G)
        P = P.flip(-1)                                   # descending eigenvalues, padding last
        Z = torch.bmm(P.transpose(1, 2).contiguous(), Y)
        sv = torch.where(pos_n < r[:, None], Z.norm(dim=2), torch.full_like(Z[:, :, 0], -1.0))
        sv, order = sv.sort(dim=1, descending=True)
        P = P.gather(2, order[:, None, :].expand(nbat, nb, nb))
        Z = Z.gather(1, order[:, :, None].expand(nbat, nb, wb))
        s_host = sv.cpu().numpy()
        for i, c in enumerate(cs):
            svals[c] = s_host[i, :min(xs[c].shape)]
        buckets.append((cs, trans, Y, P, Z, sv))
    return (xs, buckets), svals


def gram_truncated(state, keeps, device):
    """The kept factors {c: (Q_k, B_k)} of gram_spectrum's sectors, keeps = {c: k}: Q_k (rows of
    X x k) with orthonormal columns and B_k = Q_k^T X (k x columns of X), so that Q_k B_k is the
    rank-k truncation of X.  Where the Gram was taken on X's rows side, Q_k is its leading
    eigenvectors and B_k the leading rows of Z.  Where it was taken on the columns side (X has
    more rows), Q_k = X P_k / s_k, whose columns are orthogonal only to ~eps s_max^2 / (s_i s_j);
    one Cholesky QR pass (C = Q^T Q = L L^T, Q <- Q L^-T, well conditioned because every kept
    s >= GRAM_RESOLVED s_max) makes them orthonormal to working precision, and B_k = Q_k^T X is
    then the exact projection."""
    f64 = torch.float64
    xs, buckets = state
    out = {}
    for cs, trans, Y, P, Z, sv in buckets:
        tb = []
        for i, (c, t) in enumerate(zip(cs, trans)):
            k = keeps.get(c, 0)
            if k == 0:
                continue
            n, w = xs[c].shape
            if not t:
                out[c] = (P[i, :n, :k], Z[i, :k, :w])
            else:
                tb.append((i, c, k))
        if not tb:
            continue
        kmax = max(k for _, _, k in tb)
        nmax = max(xs[c].shape[0] for _, c, _ in tb)
        nt = len(tb)
        idx = torch.tensor([i for i, _, _ in tb], device=device)
        kk = torch.tensor([k for _, _, k in tb], device=device)
        keep = torch.arange(kmax, device=device)[None, :] < kk[:, None]                    # (nt, kmax)
        s_k = torch.where(keep, sv[idx, :kmax], torch.ones_like(sv[idx, :kmax]))
        U = Z[idx, :kmax, :nmax].transpose(1, 2) / s_k[:, None, :]                       # (nt, nmax, kmax)
        U = U * keep[:, None, :]
        C = torch.bmm(U.transpose(1, 2).contiguous(), U)
        C.diagonal(dim1=1, dim2=2).add_((~keep).to(f64))
        Lc = torch.linalg.cholesky(C)
        U = torch.linalg.solve_triangular(Lc.transpose(1, 2).contiguous(), U, upper=True, left=False)
        Bt = torch.bmm(Y[idx, :, :nmax], U)                                               # (nt, w_pad, kmax) = X^T Q
        for j, (i, c, k) in enumerate(tb):
            n, w = xs[c].shape
            out[c] = (U[j, :n, :k], Bt[j, :w, :k].t())
    return out


def svd_factors(mats, keys, direction, device, threads=8):
    """Singular values and full factors {c: (Q, B)} (X = Q B, Q = U, B = S V^T) from one
    LAPACK-style SVD per sector (torch.linalg.svd).  The calls are latency-bound chains of small
    kernels, so on a GPU the sectors are spread over a few host threads, each on its own stream
    (default priority)."""
    def one(c):
        X = mats[c] if direction > 0 else mats[c].t()
        try:
            U, S, Vh = torch.linalg.svd(X, full_matrices=False)
        except torch.linalg.LinAlgError:
            # the device solver's iteration can fail to converge on matrices with many repeated or
            # tiny singular values; the host LAPACK divide-and-conquer SVD of the same matrix does not
            U, S, Vh = (t.to(X.device) for t in torch.linalg.svd(X.cpu(), full_matrices=False))
        return c, (U, S[:, None] * Vh, S)

    if torch.device(device).type != "cuda" or len(keys) < 4:
        res = dict(one(c) for c in keys)
    else:
        from concurrent.futures import ThreadPoolExecutor
        torch.cuda.current_stream(device).synchronize()        # the matrices are written
        order = sorted(keys, key=lambda c: -mats[c].numel())   # largest first
        chunks = [order[k::threads] for k in range(threads)]

        def run(chunk):
            st = torch.cuda.Stream(device=device)
            with torch.cuda.stream(st):
                out = [one(c) for c in chunk]
            st.synchronize()
            return out
        res = {}
        with ThreadPoolExecutor(max_workers=threads) as pool:
            for part in pool.map(run, chunks):
                res.update(part)
    return {c: (U, B) for c, (U, B, _) in res.items()}, {c: S.cpu().numpy() for c, (_, _, S) in res.items()}


GRAM_RESOLVED = 1e-5


def split_factors(mats, keys, direction, device, Mmax, cutoff):
    """Truncated factors of a two-site split (twositetensor.hpp:153-200 with estimate_truncation,
    block_matrix_algorithms.h:135-183).  The spectrum comes from gram_spectrum; if the truncation
    threshold falls below GRAM_RESOLVED times the largest singular value, where the Gram spectrum
    cannot place it, every sector is factorised by svd_factors instead.  Returns
    ({c: (Q_k, B_k)} for the sectors that keep k > 0 states, keeps (per key), singular values
    (per key), truncated weight, method), with Q_k B_k the rank-k truncation of X (M, or M^T
    right-to-left) and Q_k orthonormal."""
    state, sv = gram_spectrum(mats, keys, direction, device)
    svals = [sv[c] for c in keys]
    if not any(len(v) for v in svals):
        return {}, [0] * len(keys), svals, 0.0, "gram"
    cut, allv = truncation_cut(svals, Mmax, cutoff)
    if cut >= GRAM_RESOLVED * allv[0]:
        keeps, tw = truncate(svals, Mmax, cutoff)
        return gram_truncated(state, dict(zip(keys, keeps)), device), keeps, svals, tw, "gram"
    del state
    full, sv = svd_factors(mats, keys, direction, device)
    svals = [sv[c] for c in keys]
    keeps, tw = truncate(svals, Mmax, cutoff)
    fac = {c: (full[c][0][:, :k], full[c][1][:k]) for c, k in zip(keys, keeps) if k > 0}
    return fac, keeps, svals, tw, "svd"


def _stored_cohorts(bnd):
    """(ci, n_blocks, bra position, ket position) of the cohorts that hold blocks.  A step over an
    MPO without Hermitian pairs (plain_mpo_desc) stores every block itself, with no transposed or
    conjugate-scaled references, which the noise terms below rely on."""
    idx = bnd.index
    if idx.tr.any() or idx.transposes.any():
        raise RuntimeError("noise term: boundary with transposed block references (expected a step over "
                           "plain_mpo_desc)")
    out = []
    for ci in range(idx.n_cohorts):
        nb = int(idx.n_blocks[ci])
        if nb:
            out.append((ci, nb, int(idx.bra_pos[ci]), int(idx.ket_pos[ci])))
    return out


def plain_mpo_desc(W):
    """Descriptor of MPO tensor W without its Hermitian bond pairs: a boundary step over it computes
    every block itself, with no transposed references to partner blocks (the reference's
    complement_transpose, boundary.h:161-191, pairs blocks of bra == ket steps only)."""
    import copy
    W2 = copy.copy(W)
    W2.left_herm, W2.right_herm = [None] * W.row_dim, [None] * W.col_dim
    return E.MpoDesc(W2)


def noise_density_left(eng, sym, left, phys, mids, mats, mpo_desc, Lb):
    """The noise term of alpha_dm_direct (boundary_ops/move_boundary.hpp:38-92 with Cohort::lbtm,
    solver/tasks.cpp:346-361) for a left-to-right split, without the factor alpha:

        N_c = sum_b' X'_b' X'_b'^T,   X'_b' = sum_b L_b W_b,b' X    (rows (s1, l) of mid charge c)

    X = mats[c] is the two-site tensor of mid charge c as the left-paired matrix of the left site
    (rows (s1, l), columns (s2, r)), W the left site's MPO tensor (mpo_desc: plain_mpo_desc, bra and
    ket differ) and L its left boundary: the S of a
    left boundary step before the bra contraction.  It is computed as a grouped left step on the
    device whose bra is the identity: the step's site has the mid charges with sizes
    D_c = max(rows, columns), the bra's left-paired block is [I | 0] (rows x D_c) and the ket's
    [X | 0], so that the new boundary's blocks are L'_b'(c, c') = [X'_b'(c, c') 0; 0 0] and N_c sums
    L'[:rows] L'[:rows]^T over the blocks of the cohorts with bra charge c.  Returns {c: N_c}."""
    f64, dev = torch.float64, eng.torch_device
    rows = {c: mats[c].shape[0] for c in mids}
    pad = Index(mids, [max(mats[c].shape[0], mats[c].shape[1], 1) for c in mids])
    ket = SiteTensor(eng, sym, left, phys, pad)
    lay = ket.layout
    ket_lp = torch.zeros(max(lay.lp_total, 1), dtype=f64, device=dev)
    bra_lp = torch.zeros_like(ket_lp)
    for mb, c in enumerate(pad.charges):
        n, D = lay.lp_rows[mb], pad.sizes[mb]
        if n != rows[c]:
            raise RuntimeError("noise term: left-paired rows of mid sector %s: %d, expected %d" % (c, n, rows[c]))
        if n == 0:
            continue
        o = lay.lp_offset[mb]
        ket_lp[o:o + n * D].view(D, n).t()[:, :mats[c].shape[1]] = mats[c]
        bra_lp[o:o + n * D].view(D, n).t()[:, :n] = torch.eye(n, dtype=f64, device=dev)
    E.reshape_left_to_right(eng, ket.desc, ket_lp, ket.data)
    Ln = E.left_step(eng, ket.desc, mpo_desc, bra_lp, ket.data, Lb, symmetric=False)
    out = {c: torch.zeros(rows[c], rows[c], dtype=f64, device=dev) for c in mids}
    idx = Ln.index
    for ci, nb, bp, kp in _stored_cohorts(Ln):
        c = idx.bra.charges[bp]
        n = rows.get(c, 0)
        if n == 0:
            continue
        Db, Dk = idx.bra.sizes[bp], idx.ket.sizes[kp]
        # block b is Db x Dk column-major: as (Dk, Db) row-major it is the transpose
        blk = Ln.cohort(ci).view(nb * Dk, Db)[:, :n]
        out[c] += blk.t() @ blk
    Ln.close()
    return out


def noise_density_right(eng, sym, right, phys, mids, mats, mpo_desc, Rb):
    """The noise term of alpha_dm_direct_right (boundary_ops/move_boundary.hpp:94-145 with
    Cohort::rbtm, solver/tasks.cpp:363-378) for a right-to-left split, without the factor alpha:

        N_c = sum_b X'_b^T X'_b,   X'_b = sum_b' X W_b,b' R_b'    (columns (s2, r) of mid charge c)

    X = mats[c] as in noise_density_left (rows (s1, l), columns (s2, r): the right-paired matrix of
    the right site with the left index (s1, l)), W the right site's MPO tensor (plain_mpo_desc) and
    R its right boundary.  A grouped right step with the identity as bra: site left index of mid charges with
    sizes D_c, bra block [I; 0] (D_c x columns), ket block [X; 0]; the new boundary's block b of
    cohort (c', c) has the ket side (X' rows) as rows and the bra side (columns (s2, r)) as columns,
    and N_c sums L'[:, :cols]^T L'[:, :cols] over the cohorts whose column charge is c."""
    f64, dev = torch.float64, eng.torch_device
    cols = {c: mats[c].shape[1] for c in mids}
    pad = Index(mids, [max(mats[c].shape[0], mats[c].shape[1], 1) for c in mids])
    ket = SiteTensor(eng, sym, pad, phys, right)
    bra = SiteTensor(eng, sym, pad, phys, right)
    lay = ket.layout
    for mb, c in enumerate(pad.charges):
        if lay.rp_cols[mb] != cols[c]:
            raise RuntimeError("noise term: right-paired columns of mid sector %s: %d, expected %d"
                               % (c, lay.rp_cols[mb], cols[c]))
        if cols[c] == 0:
            continue
        ket.rp_block(mb)[:mats[c].shape[0], :] = mats[c]
        bra.rp_block(mb)[:cols[c], :] = torch.eye(cols[c], dtype=f64, device=dev)
    Rn = E.right_step(eng, ket.desc, mpo_desc, bra.data, ket.data, Rb, symmetric=False)
    out = {c: torch.zeros(cols[c], cols[c], dtype=f64, device=dev) for c in mids}
    idx = Rn.index
    for ci, nb, bp, kp in _stored_cohorts(Rn):
        c = idx.ket.charges[kp]
        n = cols.get(c, 0)
        if n == 0:
            continue
        Db, Dk = idx.bra.sizes[bp], idx.ket.sizes[kp]
        # (Dk, Db) row-major per block: the block's transpose; its first n rows are columns < n
        blk = Rn.cohort(ci).view(nb, Dk, Db)[:, :n, :].permute(1, 0, 2).reshape(n, nb * Db)
        out[c] += blk @ blk.t()
    Rn.close()
    return out


def dm_factors(mats, keys, direction, device, Mmax, cutoff, noise, alpha):
    """heev_truncate (block_matrix_algorithms.h:298-354) of the density matrix with noise, as in
    predict_split_l2r / r2l (boundary_ops/prediction.hpp:152-237): per sector X = M (left-to-right)
    or M^T (right-to-left), rho = X X^T + alpha N, eigenvectors by descending eigenvalue, singular
    values sqrt(max(lambda, 0)) into estimate_truncation, Q_k = the kept eigenvectors and
    B_k = Q_k^T X.  Returns ({c: (Q_k, B_k)}, keeps, singular values, truncated weight)."""
    svals, vecs = [], {}
    for c in keys:
        X = mats[c] if direction > 0 else mats[c].t()
        if X.shape[0] == 0 or X.shape[1] == 0:
            svals.append(np.zeros(0))
            continue
        rho = X @ X.t() + alpha * noise[c]
        lam, P = torch.linalg.eigh(0.5 * (rho + rho.t()))
        lam, P = lam.flip(0), P.flip(1)
        vecs[c] = (X, P)
        svals.append(torch.sqrt(torch.clamp(lam, min=0.0)).cpu().numpy())
    if not any(len(v) for v in svals):
        return {}, [0] * len(keys), svals, 0.0
    keeps, tw = truncate(svals, Mmax, cutoff)
    fac = {}
    for c, k in zip(keys, keeps):
        if k > 0:
            X, P = vecs[c]
            Q = P[:, :k]
            fac[c] = (Q, Q.t() @ X)
    return fac, keeps, svals, tw


class TwoSiteDMRG:
    def __init__(self, eng, model, L, m, cutoff=1e-10, init_m=16, seed=1, jd_tol=1e-8, jd_maxiter=8,
                 mpo=None, basis=None, target=None, verbose=False, hook=None, precision=E.PREC_F64, max_outer=1,
                 init_bonds="sectors", tune_library=False, noise=0.0, boundary_cache=False, spill_dir=None,
                 **model_kw):
        """precision != PREC_F64: sigma in that precision inside Jacobi-Davidson, with the FP64
        residual refinement of SigmaPlan.jd_refined (at most max_outer FP64 steps per site).  With
        max_outer = 1 a site costs jd_maxiter reduced-precision products and one FP64 product (the
        Rayleigh quotient and residual), and the sweep follows the FP64 sweep's trajectory; more
        outer steps restart the iteration and suit solves to a tight tolerance.

        tune_library: in FP32 / BF16x3, autotune each site's S-stage library GEMMs that the
        engine's shape and shape-class caches do not cover yet (SigmaPlan.autotune with
        library_only); the timings are counted in tune_timings and their time in "plan".  Off by
        default: at L=30, m=2048 the shapes of successive sites rarely share a class, and the
        first sweep spent 24.6 s timing candidates to save 0.7 s of products.

        noise: density-matrix noise alpha of the split (predict_split_l2r / r2l,
        boundary_ops/prediction.hpp:152-237; the reference's alpha_initial / alpha_main /
        alpha_final), a number for every sweep or a sequence with one value per sweep (0 after its
        end).  With alpha > 0 the split diagonalises rho = X X^T + alpha N (noise_density_left /
        _right, dm_factors) instead of taking the truncated SVD.

        boundary_cache: keep only the boundaries of the current and the next site pair on the
        device (the reference's storage::gpu policy around each site, ts_optimize.hpp:153-182,
        307-345): a boundary is evicted to pinned host memory (spill_dir: to a file there) when the
        sweep leaves it and prefetched on the copy stream one pair before it is read again;
        boundaries the other half-sweep recomputes before reading are freed.

        init_bonds: "sectors" caps every charge sector of the initial MPS at init_m (the reference's
        mps initialisation, mps_sectors.h); "total" draws bond indices of total dimension ~init_m
        (mps.random_bond_dims), the shape a converged state at that bond dimension has."""
        self.eng, self.L, self.m, self.cutoff = eng, L, m, cutoff
        self.jd_tol, self.jd_maxiter, self.verbose = jd_tol, jd_maxiter, verbose
        self.precision, self.max_outer = precision, max_outer
        self.sigma_calls = {"reduced": 0, "fp64": 0}
        self.split_methods = {}
        self.tune_library, self.tune_timings = tune_library, 0
        self.hook = hook
        self.noise = noise
        self.plain_desc = {}
        self.boundary_cache, self.spill_dir = boundary_cache, spill_dir
        self._spills = 0
        if mpo is None:
            basis, mpo, target = models.chain(model, L, **model_kw)
        self.basis, self.mpo, self.target = basis, mpo, target
        self.sym = basis.sym
        self.phys = basis.phys
        self.mpo_desc = [E.MpoDesc(W) for W in mpo]
        self.ts_mpo = {}
        self.su2 = self.sym == SYM_SU2U1
        rng = np.random.default_rng(seed)
        phys_ok = M.physical_filter(self.sym)
        if init_bonds == "total":
            bonds = M.random_bond_dims([self.phys] * L, target, init_m, physical=phys_ok)
        else:
            bonds, _, _ = M.allowed_sectors([self.phys] * L, target, init_m, physical=phys_ok)
        self.sites = []
        for i in range(L):
            lay = SiteLayout(bonds[i], self.phys, bonds[i + 1])
            v = torch.from_numpy(M.random_site_vector(lay, rng)).to(eng.torch_device)
            self.sites.append(SiteTensor(eng, self.sym, bonds[i], self.phys, bonds[i + 1], v))
        self.energies = []
        self.timings = {"merge": 0.0, "plan": 0.0, "jd": 0.0, "split": 0.0, "boundary": 0.0, "sweep": []}

    # ------------------------------------------------------------------ canonical form
    def right_normalize(self):
        """Bring sites 1..L-1 into right-canonical form (LQ per left sector, mps.hpp canonize)."""
        for i in range(self.L - 1, 0, -1):
            A = self.sites[i]
            newdims, Qs, Rs = {}, {}, {}
            for lb, lc in enumerate(A.left.charges):
                blk = A.rp_block(lb)
                if blk.numel() == 0:
                    continue
                q, r = torch.linalg.qr(blk.t().contiguous(), mode="reduced")   # blk^T = q r
                k = q.shape[1]
                newdims[lc] = k
                Qs[lc], Rs[lc] = q.t(), r.t()                                   # blk = r^T q^T
            newleft = Index(list(newdims.keys()), list(newdims.values()))
            self.sites[i] = SiteTensor.from_rp(self.eng, self.sym, newleft, A.phys, A.right,
                                               {newleft.position(c): Qs[c] for c in newdims})
            B = self.sites[i - 1]
            lp = B.lp_data()
            blocks = {}
            for rb, rc in enumerate(B.right.charges):
                if rc not in newdims:
                    continue
                blocks[newleft.position(rc)] = B.lp_block(lp, rb) @ Rs[rc]
            self.sites[i - 1] = SiteTensor.from_lp(self.eng, self.sym, B.left, B.phys, newleft, blocks)
        # normalise site 0
        s0 = self.sites[0]
        nrm = torch.linalg.norm(s0.data)
        s0.data /= nrm

    def move_center_right(self, j):
        """QR of site j (left-paired, per right sector): site j becomes left-canonical and R moves
        into site j+1 (mps.hpp canonize_left_step)."""
        A, B = self.sites[j], self.sites[j + 1]
        lp = A.lp_data()
        dims, Qs, Rs = {}, {}, {}
        for rb, rc in enumerate(A.right.charges):
            blk = A.lp_block(lp, rb)
            if blk.numel() == 0:
                continue
            q, r = torch.linalg.qr(blk.contiguous(), mode="reduced")
            dims[rc], Qs[rc], Rs[rc] = q.shape[1], q, r
        newright = Index(list(dims.keys()), list(dims.values()))
        self.sites[j] = SiteTensor.from_lp(self.eng, self.sym, A.left, A.phys, newright,
                                           {newright.position(c): Qs[c] for c in dims})
        blocks = {}
        for lb, lc in enumerate(B.left.charges):
            if lc in dims:
                blocks[newright.position(lc)] = Rs[lc] @ B.rp_block(lb)
        self.sites[j + 1] = SiteTensor.from_rp(self.eng, self.sym, newright, B.phys, B.right, blocks)

    def mpo_gradients(self):
        """Derivatives of the energy with respect to every MPO term scale of every site, at the
        current state (after run(): orthogonality center at site 0): the center walks left to
        right by QR, and at each site j the one-site adjoint (SigmaPlan.adjoint) of
        psi_j . H_eff psi_j / psi_j . psi_j with the boundaries L_j, R_{j+1} gives dE/d scale for
        the terms of mpo[j] (Hellmann-Feynman at the variational optimum).  Returns (per-site
        gradient arrays in mpo[j].sorted_terms() order, per-site energies psi_j . H psi_j)."""
        grads, energies = [], []
        for j in range(self.L):
            s = self.sites[j]
            if self.boundary_cache and j + 2 <= self.L and self.Rb[j + 2] is not None:
                self.Rb[j + 2].prefetch()   # the next site's right boundary arrives meanwhile
            nrm2 = float(torch.dot(s.data, s.data))
            plan = E.SigmaPlan(self.eng, s.desc, self.mpo_desc[j], self.Lb[j], self.Rb[j + 1],
                               algorithm=E.SIGMA_GROUPED, adjoint=True)
            grads.append(plan.adjoint(s.data) / nrm2)
            out = torch.zeros_like(s.data)
            plan.apply(s.data, out)
            self.eng.sync()
            energies.append(float(torch.dot(s.data, out)) / nrm2)
            plan.close()
            if j < self.L - 1:
                self.move_center_right(j)
                a = self.sites[j]
                self.Lb[j + 1] = E.left_step(self.eng, a.desc, self.mpo_desc[j], a.lp_data(), a.data, self.Lb[j])
                if self.boundary_cache:   # both stay valid for the state; off the device until read again
                    if j >= 1:
                        self._evict(self.Lb[j])
                    if j + 1 < self.L:
                        self._evict(self.Rb[j + 1])
        return grads, energies

    # ------------------------------------------------------------------ two-site tensors
    def twosite_mpo(self, i):
        if i not in self.ts_mpo:
            if self.su2:
                W12, phys12, pb = su2.twosite_su2(self.mpo[i], self.mpo[i + 1], self.phys)
            else:
                W12, phys12, pb = models.twosite_sparse(self.mpo[i], self.mpo[i + 1], self.phys, self.phys)
            self.ts_mpo[i] = (W12, E.MpoDesc(W12), phys12, pb)
        return self.ts_mpo[i]

    def _recouple(self, theta, reduce):
        """SU2U1: the two-site tensor in the coupled two-site basis of the site problem
        (TwoSiteTensor::make_mps -> ts_reduction::reduce_right, twositetensor.hpp:131-151) and
        back before the split (operator<< -> unreduce_right, twositetensor.hpp:296-326)."""
        segs = (su2.reduction_segments if reduce else su2.unreduction_segments)(theta.left, self.phys, theta.right)
        theta.data = apply_segments(theta.data, segs)

    def _pairs(self, A1, A2, phys12, pb):
        """Enumerate (mid, s1, lc, s2, rc) slabs: rows of the (s1, lc) x mid left-paired A1 and
        columns of the mid x (s2, rc) right-paired A2, with their two-site column offset."""
        lay12 = SiteLayout(A1.left, phys12, A2.right)
        for mb, mc in enumerate(A1.right.charges):
            for s1, (p1, d1) in enumerate(zip(self.phys.charges, self.phys.sizes)):
                lc = sub(mc, p1)
                lb = A1.left.position(lc)
                if lb == len(A1.left):
                    continue
                ls = A1.left.sizes[lb]
                r0 = A1.layout.lpb(p1, lc)
                for s2, (p2, d2) in enumerate(zip(self.phys.charges, self.phys.sizes)):
                    rc = fuse(mc, p2)
                    rb = A2.right.position(rc)
                    if rb == len(A2.right):
                        continue
                    rs = A2.right.sizes[rb]
                    c0 = A2.layout.rpb(p2, rc)
                    pc = fuse(p1, p2)
                    col12 = lay12.rpb(pc, rc)
                    for ss1 in range(d1):
                        for ss2 in range(d2):
                            sub_i = pb(p1, p2) + ss1 * d2 + ss2
                            yield (mc, lb, ls, r0 + ss1 * ls, rb, rs, c0 + ss2 * rs, col12 + sub_i * rs)

    def merge(self, i):
        A1, A2 = self.sites[i], self.sites[i + 1]
        W12, wdesc, phys12, pb = self.twosite_mpo(i)
        theta = SiteTensor(self.eng, self.sym, A1.left, phys12, A2.right)
        lp1 = A1.lp_data()
        prods, base, ncols, off = [], {}, {}, 0
        for mb, mc in enumerate(A1.right.charges):
            lb2 = A2.left.position(mc)
            if lb2 == len(A2.left):
                continue
            P = (A1.lp_block(lp1, mb) @ A2.rp_block(lb2)).contiguous()
            prods.append(P.reshape(-1))
            base[mc], ncols[mc] = off, P.shape[1]
            off += P.numel()
        if not prods:
            return theta
        flat = torch.cat(prods)
        # theta's right-paired block lb holds (row i, column col) at rp_offset[lb] + col * ls + i
        slabs = [(theta.layout.rp_offset[lb] + col12 * ls, 1, ls, base[mc] + r * ncols[mc] + c, ncols[mc], 1, ls, rs)
                 for (mc, lb, ls, r, rb, rs, c, col12) in self._pairs(A1, A2, phys12, pb) if mc in base]
        slab_copy(theta.data, flat, slabs)
        if self.su2:
            self._recouple(theta, True)
        return theta

    def alpha(self):
        """Noise of the current sweep (self.energies holds one entry per finished sweep)."""
        k = len(self.energies)
        if np.isscalar(self.noise):
            return float(self.noise)
        return float(self.noise[k]) if k < len(self.noise) else 0.0

    def split(self, i, theta, direction, alpha=0.0):
        """Truncated SVD split of the two-site tensor (split_mps_l2r / split_mps_r2l), or with
        alpha > 0 the density-matrix split with noise (predict_split_l2r / r2l)."""
        A1, A2 = self.sites[i], self.sites[i + 1]
        if self.su2:
            self._recouple(theta, False)
        W12, wdesc, phys12, pb = self.twosite_mpo(i)
        # mid sectors reachable from both sides
        lidx = {}
        for lc, ls in zip(A1.left.charges, A1.left.sizes):
            for p1, d1 in zip(self.phys.charges, self.phys.sizes):
                mc = fuse(lc, p1)
                lidx[mc] = lidx.get(mc, 0) + d1 * ls
        ridx = {}
        for rc, rs in zip(A2.right.charges, A2.right.sizes):
            for p2, d2 in zip(self.phys.charges, self.phys.sizes):
                mc = sub(rc, p2)
                ridx[mc] = ridx.get(mc, 0) + d2 * rs
        mids = [c for c in sorted(lidx, reverse=True) if c in ridx]
        mid_full = Index(mids, [1] * len(mids))
        # layouts of the unsplit factors with a full mid index (only row/col offsets are used)
        L1 = SiteLayout(A1.left, self.phys, Index(mids, [lidx[c] for c in mids]))
        L2 = SiteLayout(Index(mids, [ridx[c] for c in mids]), self.phys, A2.right)
        mbase, off = {}, 0
        for c in mids:
            mbase[c] = off
            off += lidx[c] * ridx[c]
        mflat = torch.zeros(max(off, 1), dtype=torch.float64, device=self.eng.torch_device)
        mats = {c: mflat[mbase[c]:mbase[c] + lidx[c] * ridx[c]].view(lidx[c], ridx[c]) for c in mids}
        lay12 = SiteLayout(A1.left, phys12, A2.right)
        slabs = []
        for mc in mids:
            for p1, d1 in zip(self.phys.charges, self.phys.sizes):
                lc = sub(mc, p1)
                lb = A1.left.position(lc)
                if lb == len(A1.left):
                    continue
                ls = A1.left.sizes[lb]
                r0 = L1.lpb(p1, lc)
                for p2, d2 in zip(self.phys.charges, self.phys.sizes):
                    rc = fuse(mc, p2)
                    rb = A2.right.position(rc)
                    if rb == len(A2.right):
                        continue
                    rs = A2.right.sizes[rb]
                    c0 = L2.rpb(p2, rc)
                    col12 = lay12.rpb(fuse(p1, p2), rc)
                    o = theta.layout.rp_offset[lb]
                    for ss1 in range(d1):
                        for ss2 in range(d2):
                            sub_i = pb(p1, p2) + ss1 * d2 + ss2
                            # mats[mc][r0 + ss1 ls + i, c0 + ss2 rs + j] = theta block lb (i, col12 + sub_i rs + j)
                            slabs.append((mbase[mc] + (r0 + ss1 * ls) * ridx[mc] + c0 + ss2 * rs, ridx[mc], 1,
                                          o + (col12 + sub_i * rs) * ls, 1, ls, ls, rs))
        slab_copy(mflat, theta.data, slabs)
        if alpha > 0:
            k = i if direction > 0 else i + 1
            if k not in self.plain_desc:
                self.plain_desc[k] = plain_mpo_desc(self.mpo[k])
            if direction > 0:
                N = noise_density_left(self.eng, self.sym, A1.left, self.phys, mids, mats, self.plain_desc[k],
                                       self.Lb[i])
            else:
                N = noise_density_right(self.eng, self.sym, A2.right, self.phys, mids, mats, self.plain_desc[k],
                                        self.Rb[i + 2])
            fac, keeps, svals, tw = dm_factors(mats, mids, direction, self.eng.torch_device, self.m, self.cutoff,
                                               N, alpha)
            method = "dm"
        else:
            fac, keeps, svals, tw, method = split_factors(mats, mids, direction, self.eng.torch_device, self.m,
                                                          self.cutoff)
        self.split_methods[method] = self.split_methods.get(method, 0) + 1
        kept = [(c, k) for c, k in zip(mids, keeps) if k > 0]
        newmid = Index([c for c, _ in kept], [k for _, k in kept])
        if method == "dm":
            # the kept eigenvalues include the noise: the norm is that of the projected tensor
            norm = float(torch.sqrt(sum(torch.sum(fac[c][1] ** 2) for c, _ in kept)))
        else:
            norm = np.sqrt(sum(float(np.sum(s[:k] ** 2)) for s, k in zip(svals, keeps)))
        b1, b2 = {}, {}
        for c, k in kept:
            orth, other = fac[c][0], fac[c][1] / norm
            if direction > 0:
                b1[c], b2[c] = orth, other
            else:
                b1[c], b2[c] = other.t(), orth.t()
        # scatter back into the factor layouts: A1 left-paired rows (s1, lc), A2 right-paired
        N1 = SiteLayout(A1.left, self.phys, newmid)
        N2 = SiteLayout(newmid, self.phys, A2.right)
        lp_blocks, rp_blocks = {}, {}
        for c, k in kept:
            mb = newmid.position(c)
            blk1 = torch.zeros(N1.lp_rows[mb], k, dtype=torch.float64, device=self.eng.torch_device)
            for p1, d1 in zip(self.phys.charges, self.phys.sizes):
                lc = sub(c, p1)
                if not A1.left.has(lc):
                    continue
                n = d1 * A1.left.size_of(lc)
                blk1[N1.lpb(p1, lc):N1.lpb(p1, lc) + n] = b1[c][L1.lpb(p1, lc):L1.lpb(p1, lc) + n]
            lp_blocks[mb] = blk1
            blk2 = torch.zeros(k, N2.rp_cols[mb], dtype=torch.float64, device=self.eng.torch_device)
            for p2, d2 in zip(self.phys.charges, self.phys.sizes):
                rc = fuse(c, p2)
                if not A2.right.has(rc):
                    continue
                n = d2 * A2.right.size_of(rc)
                blk2[:, N2.rpb(p2, rc):N2.rpb(p2, rc) + n] = b2[c][:, L2.rpb(p2, rc):L2.rpb(p2, rc) + n]
            rp_blocks[mb] = blk2
        self.sites[i] = SiteTensor.from_lp(self.eng, self.sym, A1.left, self.phys, newmid, lp_blocks)
        self.sites[i + 1] = SiteTensor.from_rp(self.eng, self.sym, newmid, self.phys, A2.right, rp_blocks)
        return tw, sum(newmid.sizes)

    # ------------------------------------------------------------------ boundary residency
    def _evict(self, b):
        if b is None:
            return
        path = None
        if self.spill_dir is not None:
            import os
            self._spills += 1
            path = os.path.join(self.spill_dir, "boundary_%d_%d.bin" % (id(self), self._spills))
        b.evict(path)

    def _drop(self, side, k):
        b = side[k]
        if b is not None:
            b.close()
        side[k] = None

    def _cache_before(self, i, direction):
        """Prefetch the boundary the next site pair reads (it arrives while this pair computes)."""
        if not self.boundary_cache:
            return
        nxt = None
        if direction > 0 and i + 3 <= self.L:
            nxt = self.Rb[i + 3]
        elif direction < 0 and i >= 1:
            nxt = self.Lb[i - 1]
        if nxt is not None:
            nxt.prefetch()

    def _cache_after(self, i, direction):
        """After pair i: the boundary this pair read on the sweep's trailing side is evicted (the
        other half-sweep reads it again), the one on the leading side is freed (the other
        half-sweep recomputes it before reading it).  The trivial end boundaries stay."""
        if not self.boundary_cache:
            return
        if direction > 0:
            if i >= 1:
                self._evict(self.Lb[i])
            if i + 2 < self.L:
                self._drop(self.Rb, i + 2)
        else:
            if i + 2 < self.L:
                self._evict(self.Rb[i + 2])
            if i >= 1:
                self._drop(self.Lb, i)

    def resident_boundaries(self):
        """Number of boundaries whose data is on the device."""
        return sum(1 for b in list(self.Lb) + list(self.Rb) if b is not None and b.residency()[0] == E.DeviceBoundary.DEVICE)

    # ------------------------------------------------------------------ sweeps
    def init_boundaries(self):
        """Optimizer set-up before the first sweep: the two-site MPO cache (make_ts_cache_mpo in
        the ts_optimize constructor, ts_optimize.hpp:59), right normalization and the right
        boundaries."""
        for i in range(self.L - 1):
            self.twosite_mpo(i)
        self.right_normalize()
        L = self.L
        self.Lb = [None] * (L + 1)
        self.Rb = [None] * (L + 1)
        self.Lb[0] = E.DeviceBoundary.trivial(self.eng, self.sites[0].left, self.mpo[0].row_dim)
        self.Rb[L] = E.DeviceBoundary.trivial(self.eng, self.sites[L - 1].right, self.mpo[L - 1].col_dim)
        for k in range(L - 1, 1, -1):
            s = self.sites[k]
            self.Rb[k] = E.right_step(self.eng, s.desc, self.mpo_desc[k], s.data, s.data, self.Rb[k + 1])
            if self.boundary_cache and k + 1 < L and k + 1 > 3:
                self._evict(self.Rb[k + 1])   # read again from the left end of the first sweep

    def multi_expval(self, mpo=None):
        """<psi|W|psi> per column index of the last MPO tensor for an MPO W on the chain (default:
        the Hamiltonian), by left boundary propagation over every site on the device
        (multi_expval, mps_mpo_ops.h:80-100): L_{k+1} = overlap_mpo_left_step(A_k, A_k, L_k, W_k),
        then the traces of the last boundary."""
        mpo = self.mpo if mpo is None else mpo
        descs = self.mpo_desc if mpo is self.mpo else [E.MpoDesc(W) for W in mpo]
        Lb = E.DeviceBoundary.trivial(self.eng, self.sites[0].left, mpo[0].row_dim)
        for k in range(self.L):
            s = self.sites[k]
            Lb = E.left_step(self.eng, s.desc, descs[k], s.lp_data(), s.data, Lb)
        return Lb.traces()

    def expval(self, mpo=None):
        """<psi|W|psi> (unnormalised, as expval, mps_mpo_ops.h:52-77): the trace of the last left
        boundary of multi_expval."""
        return float(np.sum(self.multi_expval(mpo)))

    def norm2(self):
        """<psi|psi>: expval of the identity MPO."""
        return self.expval(models.identity_mpo(self.phys, self.L))

    def rdm1(self):
        """Spin-summed one-particle density matrix gamma_pq = <psi|E_pq|psi> / <psi|psi>, E_pq =
        sum_sigma a+_{p sigma} a_{q sigma}, of an ab initio chain ('qc': TwoU1, 'qc_su2': SU2U1),
        measured as expectation values of one-body MPOs (the route of the reference's 1-RDM
        measurement, tagged_nrankrdm.h, through expval, mps_mpo_ops.h:52-100): the chain's MPO
        builder with h = (e_pq + e_qp) / 2 and no two-body part gives <(E_pq + E_qp) / 2> =
        gamma_pq for a real state.  L (L + 1) / 2 MPOs, each evaluated by left boundary steps over
        the chain on the device."""
        from . import qc, su2 as su2m
        L = self.L
        zero = np.zeros((L, L, L, L))
        nrm = self.norm2()
        gamma = np.zeros((L, L))
        for p in range(L):
            for q in range(p, L):
                h = np.zeros((L, L))
                h[p, q] += 0.5
                h[q, p] += 0.5
                if self.su2:
                    _, mpo, _ = su2m.qc_su2_mpo(h, zero, self.target[0])
                else:
                    _, mpo, _ = qc.qc_mpo(h, zero)
                gamma[p, q] = gamma[q, p] = self.expval(mpo) / nrm
        return gamma

    def rdm2_classes(self):
        """SU2U1 ab initio chain: for every permutation class of (pq|rs) (one element of
        su2.integral_elements, keyed by its aligned indices), <psi|V|psi> / <psi|psi> with V the
        two-body operator that element multiplies in the spin-adapted Hamiltonian: the sum of the
        spin-summed 2-RDM entries of the class as the Hamiltonian weighs them, i.e. dE/d(element).
        Each is the expectation value of the MPO built from that element alone (unit value, no
        one-body part), L^4 / 8 MPOs in all."""
        from . import su2 as su2m
        assert self.su2, "rdm2_classes: SU2U1 chains"
        L = self.L
        zero_h = np.zeros((L, L))
        nrm = self.norm2()
        out = {}
        for i, j, k, l in itertools.product(range(L), repeat=4):
            if i < j or k < l or (i * L + j) < (k * L + l):
                continue
            eri = np.zeros((L, L, L, L))
            for a, b, c, d in ((i, j, k, l), (j, i, k, l), (i, j, l, k), (j, i, l, k),
                               (k, l, i, j), (l, k, i, j), (k, l, j, i), (l, k, j, i)):
                eri[a, b, c, d] = 1.0
            _, mpo, _ = su2m.qc_su2_mpo(zero_h, eri, self.target[0])
            out[su2m._align(i, j, k, l)] = self.expval(mpo) / nrm
        return out

    def optimize_pair(self, i, direction):
        t0 = time.perf_counter()
        self._cache_before(i, direction)
        W12, wdesc, phys12, pb = self.twosite_mpo(i)
        theta = self.merge(i)
        self.timings["merge"] += time.perf_counter() - t0
        t0 = time.perf_counter()
        plan = E.SigmaPlan(self.eng, theta.desc, wdesc, self.Lb[i], self.Rb[i + 2])
        if self.precision != E.PREC_F64 and self.tune_library:
            self.tune_timings += plan.autotune(self.precision, library_only=True)
        self.timings["plan"] += time.perf_counter() - t0
        t0 = time.perf_counter()
        if self.hook is not None:
            self.hook(self, i, theta, wdesc, plan)
        out = torch.zeros_like(theta.data)
        if self.precision == E.PREC_F64:
            e, it = plan.jd(theta.data, out, tol=self.jd_tol, max_iter=self.jd_maxiter)
            self.sigma_calls["fp64"] += it
        else:
            e, cnt, _ = plan.jd_refined(theta.data, out, tol=self.jd_tol, max_iter=self.jd_maxiter,
                                        precision=self.precision, max_outer=self.max_outer)
            it = cnt["reduced"]
            self.sigma_calls["reduced"] += cnt["reduced"]
            self.sigma_calls["fp64"] += cnt["fp64"]
        plan.close()
        self.timings["jd"] += time.perf_counter() - t0
        theta.data = out
        t0 = time.perf_counter()
        tw, bd = self.split(i, theta, direction, self.alpha())
        self.timings["split"] += time.perf_counter() - t0
        t0 = time.perf_counter()
        if direction > 0:
            s = self.sites[i]
            self.Lb[i + 1] = E.left_step(self.eng, s.desc, self.mpo_desc[i], s.lp_data(), s.data, self.Lb[i])
        else:
            s = self.sites[i + 1]
            self.Rb[i + 1] = E.right_step(self.eng, s.desc, self.mpo_desc[i + 1], s.data, s.data, self.Rb[i + 2])
        self._cache_after(i, direction)
        self.timings["boundary"] += time.perf_counter() - t0
        if self.verbose:
            print("pair %d dir %+d E=%.12f iters=%d bond=%d tw=%.2e" % (i, direction, e, it, bd, tw), flush=True)
        return e

    def sweep(self):
        t0 = time.perf_counter()
        e = None
        for i in range(0, self.L - 1):
            e = self.optimize_pair(i, +1)
        for i in range(self.L - 2, -1, -1):
            e = self.optimize_pair(i, -1)
        self.energies.append(e)
        self.timings["sweep"].append(time.perf_counter() - t0)
        return e

    def run(self, n_sweeps):
        self.init_boundaries()
        for _ in range(n_sweeps):
            self.sweep()
        return self.energies[-1]
