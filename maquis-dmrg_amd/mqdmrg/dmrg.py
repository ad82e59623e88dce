"""Two-site DMRG sweeps on the engine (caller side of the sigma / boundary path).

Mirrors the reference's ts_optimize::sweep (optimize/ts_optimize.hpp:114-362): per site pair the
two-site tensor is formed (twositetensor.hpp:131-151), the effective Hamiltonian is solved with
Jacobi-Davidson (solver_interface.hpp:30-77 -> engine mq_jd_solve), the result is split by a
truncated SVD (split_mps_l2r / r2l, twositetensor.hpp:153-200 with estimate_truncation,
block_matrix_algorithms.h:135-183), and the boundary is moved (optimize.h:134-152 -> engine
left/right steps).  Site tensors live on the device in the right-paired padded layout; the small
dense algebra of the split (SVD, QR) uses torch on the engine's stream."""
import time

import numpy as np
import torch

from . import engine as E
from . import mps as M
from . import models
from .symmetry import Index, SiteLayout, fuse, sub, product_index


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


def truncate(svals, Mmax, cutoff):
    """estimate_truncation (block_matrix_algorithms.h:135-183): per-block keep counts."""
    allv = np.sort(np.concatenate([s for s in svals if len(s)]))[::-1]
    cut = cutoff * allv[0]
    if len(allv) > Mmax:
        cut = max(cut, allv[Mmax])
    keeps = [int(np.sum(s >= cut)) for s in svals]
    tw = float(np.sum(allv[allv < cut] ** 2) / np.sum(allv ** 2))
    return keeps, tw


class TwoSiteDMRG:
    def __init__(self, eng, model, L, m, cutoff=1e-10, init_m=16, seed=1, jd_tol=1e-8, jd_maxiter=8,
                 mpo=None, basis=None, target=None, verbose=False, hook=None, precision=E.PREC_F64, max_outer=1,
                 init_bonds="sectors", **model_kw):
        """precision != PREC_F64: sigma in that precision inside Jacobi-Davidson, with the FP64
        residual refinement of SigmaPlan.jd_refined (at most max_outer FP64 steps per site).  With
        max_outer = 1 a site costs jd_maxiter reduced-precision products and one FP64 product (the
        Rayleigh quotient and residual), and the sweep follows the FP64 sweep's trajectory; more
        outer steps restart the iteration and suit solves to a tight tolerance.

        init_bonds: "sectors" caps every charge sector of the initial MPS at init_m (the reference's
        mps initialisation, mps_sectors.h); "total" draws bond indices of total dimension ~init_m
        (mps.random_bond_dims), the shape a converged state at that bond dimension has."""
        self.eng, self.L, self.m, self.cutoff = eng, L, m, cutoff
        self.jd_tol, self.jd_maxiter, self.verbose = jd_tol, jd_maxiter, verbose
        self.precision, self.max_outer = precision, max_outer
        self.sigma_calls = {"reduced": 0, "fp64": 0}
        self.hook = hook
        if mpo is None:
            basis, mpo, target = models.chain(model, L, **model_kw)
        self.basis, self.mpo, self.target = basis, mpo, target
        self.sym = basis.sym
        self.phys = basis.phys
        self.mpo_desc = [E.MpoDesc(W) for W in mpo]
        self.ts_mpo = {}
        rng = np.random.default_rng(seed)
        if init_bonds == "total":
            bonds = M.random_bond_dims([self.phys] * L, target, init_m)
        else:
            bonds, _, _ = M.allowed_sectors([self.phys] * L, target, init_m)
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
        return grads, energies

    # ------------------------------------------------------------------ two-site tensors
    def twosite_mpo(self, i):
        if i not in self.ts_mpo:
            W12, phys12, pb = models.twosite_sparse(self.mpo[i], self.mpo[i + 1], self.phys, self.phys)
            self.ts_mpo[i] = (W12, E.MpoDesc(W12), phys12, pb)
        return self.ts_mpo[i]

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
        prod = {}
        for mb, mc in enumerate(A1.right.charges):
            lb2 = A2.left.position(mc)
            if lb2 == len(A2.left):
                continue
            prod[mc] = A1.lp_block(lp1, mb) @ A2.rp_block(lb2)
        for (mc, lb, ls, r, rb, rs, c, col12) in self._pairs(A1, A2, phys12, pb):
            if mc not in prod:
                continue
            theta.rp_block(lb)[:, col12:col12 + rs] = prod[mc][r:r + ls, c:c + rs]
        return theta

    def _svds(self, mats, keys):
        """FP64 SVD of every sector matrix on the GPU.  Each call is latency-bound (small
        bidiagonalisation kernels with host round trips), so the sectors are spread over a few
        host threads, each on its own stream (default priority), and the results joined before
        use."""
        if len(keys) < 4:
            return {c: torch.linalg.svd(mats[c], full_matrices=False) for c in keys}
        if not hasattr(self, "_svd_pool"):
            from concurrent.futures import ThreadPoolExecutor
            self._svd_pool = ThreadPoolExecutor(max_workers=8)
            self._svd_streams = [torch.cuda.Stream(device=self.eng.torch_device) for _ in range(8)]
        torch.cuda.current_stream(self.eng.torch_device).synchronize()   # the matrices are written
        order = sorted(keys, key=lambda c: -mats[c].numel())              # largest first
        chunks = [order[k::8] for k in range(8)]

        def run(k):
            st = self._svd_streams[k]
            with torch.cuda.stream(st):
                out = {c: torch.linalg.svd(mats[c], full_matrices=False) for c in chunks[k]}
            st.synchronize()
            return out
        res = {}
        for part in self._svd_pool.map(run, range(8)):
            res.update(part)
        return res

    def split(self, i, theta, direction):
        """Truncated SVD split of the two-site tensor (split_mps_l2r / split_mps_r2l)."""
        A1, A2 = self.sites[i], self.sites[i + 1]
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
        mats = {c: torch.zeros(lidx[c], ridx[c], dtype=torch.float64, device=self.eng.torch_device) for c in mids}
        lay12 = SiteLayout(A1.left, phys12, A2.right)
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
                    for ss1 in range(d1):
                        for ss2 in range(d2):
                            sub_i = pb(p1, p2) + ss1 * d2 + ss2
                            mats[mc][r0 + ss1 * ls:r0 + (ss1 + 1) * ls, c0 + ss2 * rs:c0 + (ss2 + 1) * rs] = \
                                theta.rp_block(lb)[:, col12 + sub_i * rs:col12 + (sub_i + 1) * rs]
        svd = self._svds(mats, mids)
        svals = [svd[c][1].cpu().numpy() for c in mids]
        keeps, tw = truncate(svals, self.m, self.cutoff)
        kept = [(c, k) for c, k in zip(mids, keeps) if k > 0]
        newmid = Index([c for c, _ in kept], [k for _, k in kept])
        norm = np.sqrt(sum(float(np.sum(s[:k] ** 2)) for s, k in zip(svals, keeps)))
        b1, b2 = {}, {}
        for c, k in kept:
            U, S, Vh = svd[c]
            U, S, Vh = U[:, :k], S[:k] / norm, Vh[:k]
            if direction > 0:
                b1[c], b2[c] = U, S[:, None] * Vh
            else:
                b1[c], b2[c] = U * S[None, :], Vh
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

    # ------------------------------------------------------------------ sweeps
    def init_boundaries(self):
        self.right_normalize()
        L = self.L
        self.Lb = [None] * (L + 1)
        self.Rb = [None] * (L + 1)
        self.Lb[0] = E.DeviceBoundary.trivial(self.eng, self.sites[0].left, self.mpo[0].row_dim)
        self.Rb[L] = E.DeviceBoundary.trivial(self.eng, self.sites[L - 1].right, self.mpo[L - 1].col_dim)
        for k in range(L - 1, 1, -1):
            s = self.sites[k]
            self.Rb[k] = E.right_step(self.eng, s.desc, self.mpo_desc[k], s.data, s.data, self.Rb[k + 1])

    def optimize_pair(self, i, direction):
        t0 = time.perf_counter()
        W12, wdesc, phys12, pb = self.twosite_mpo(i)
        theta = self.merge(i)
        self.timings["merge"] += time.perf_counter() - t0
        t0 = time.perf_counter()
        plan = E.SigmaPlan(self.eng, theta.desc, wdesc, self.Lb[i], self.Rb[i + 2])
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
        tw, bd = self.split(i, theta, direction)
        self.timings["split"] += time.perf_counter() - t0
        t0 = time.perf_counter()
        if direction > 0:
            s = self.sites[i]
            self.Lb[i + 1] = E.left_step(self.eng, s.desc, self.mpo_desc[i], s.lp_data(), s.data, self.Lb[i])
        else:
            s = self.sites[i + 1]
            self.Rb[i + 1] = E.right_step(self.eng, s.desc, self.mpo_desc[i + 1], s.data, s.data, self.Rb[i + 2])
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
