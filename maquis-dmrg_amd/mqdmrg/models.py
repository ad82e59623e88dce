"""Site bases, operators and MPOs of the benchmark Hamiltonians.

The engine only needs an MPO that represents the same Hamiltonian as the reference's (sigma
does not depend on the MPO gauge), so the MPOs are built here as finite-state machines with
explicit Hermitian pairing of bond states (the information the reference's tagged MPO maker
records, models/generate_mpo/tagged_mpo_maker_optim.hpp:288-308).

Conventions: an operator block maps physical sector ``in`` (ket) to ``out`` (bra); an entry
(row, col, coeff) means  <out, col| O |in, row> = coeff  (block_matrix/sparse_operator.h).
Fermionic operators carry their Jordan-Wigner signs explicitly: the local annihilators are
c_up = a_up and c_dn = (-1)^{n_up} a_dn, and F = (-1)^n is the local parity.
"""
import numpy as np

from .symmetry import Index, product_index, fuse, SYM_U1, SYM_TWOU1, SYM_SU2U1


class Op:
    """Block-sparse site operator: blocks[(in_charge, out_charge)] = list of (row, col, coeff, rspin, cspin)."""

    def __init__(self, blocks=None, spin=0, name=""):
        self.blocks = {}
        self.spin = spin
        self.name = name
        for k, v in (blocks or {}).items():
            self.blocks[k] = list(v)

    def add(self, cin, cout, row, col, coeff, rspin=0, cspin=0):
        self.blocks.setdefault((tuple(cin), tuple(cout)), []).append((row, col, coeff, rspin, cspin))

    def dense(self, phys):
        """Dense matrix D[out, in] over the flattened physical basis (sectors in Index order)."""
        off = np.cumsum([0] + phys.sizes)
        n = off[-1]
        D = np.zeros((n, n))
        for (ci, co), ents in self.blocks.items():
            for r, c, v, _, _ in ents:
                D[off[phys.position(co)] + c, off[phys.position(ci)] + r] += v
        return D

    @staticmethod
    def from_dense(D, phys, spin=0, name="", tol=0.0):
        off = np.cumsum([0] + phys.sizes)
        op = Op(spin=spin, name=name)
        for i, ci in enumerate(phys.charges):
            for o, co in enumerate(phys.charges):
                blk = D[off[o]:off[o + 1], off[i]:off[i + 1]]
                for c in range(blk.shape[0]):
                    for r in range(blk.shape[1]):
                        if abs(blk[c, r]) > tol:
                            op.add(ci, co, r, c, float(blk[c, r]))
        return op

    def kron(self, other, phys_a, phys_b, phys_ab, pb):
        """Two-site operator self (x) other on the fused basis phys_ab (ProductBasis pb)."""
        res = Op(spin=0, name=self.name + "*" + other.name)
        for (ai, ao), ea in self.blocks.items():
            for (bi, bo), eb in other.blocks.items():
                ci, co = fuse(ai, bi), fuse(ao, bo)
                oi, oo = pb(ai, bi), pb(ao, bo)
                dbi, dbo = phys_b.size_of(bi), phys_b.size_of(bo)
                for r1, c1, v1, _, _ in ea:
                    for r2, c2, v2, _, _ in eb:
                        res.add(ci, co, oi + r1 * dbi + r2, oo + c1 * dbo + c2, v1 * v2)
        return res


def _merge(op):
    """Sum duplicate (row, col) entries and drop zeros."""
    out = Op(spin=op.spin, name=op.name)
    for k, ents in op.blocks.items():
        acc = {}
        for r, c, v, rs, cs in ents:
            key = (r, c, rs, cs)
            acc[key] = acc.get(key, 0.0) + v
        for (r, c, rs, cs), v in acc.items():
            if v != 0.0:
                out.add(k[0], k[1], r, c, v, rs, cs)
    return out


class SiteBasis:
    def __init__(self, sym, phys, ops):
        self.sym = sym
        self.phys = phys
        self.ops = ops


def spin_half_u1():
    """Spin-1/2 with U1 charge 2Sz (models/coded/models_u1.hpp:56-57)."""
    up, dn = (1, 0), (-1, 0)
    phys = Index([up, dn], [1, 1])
    ops = {
        "id": Op({(up, up): [(0, 0, 1.0, 0, 0)], (dn, dn): [(0, 0, 1.0, 0, 0)]}, name="id"),
        "sz": Op({(up, up): [(0, 0, 0.5, 0, 0)], (dn, dn): [(0, 0, -0.5, 0, 0)]}, name="sz"),
        "splus": Op({(dn, up): [(0, 0, 1.0, 0, 0)]}, name="splus"),
        "sminus": Op({(up, dn): [(0, 0, 1.0, 0, 0)]}, name="sminus"),
    }
    return SiteBasis(SYM_U1, phys, ops)


def hubbard_twou1():
    """Spinful fermions with TwoU1 charge (n_up, n_down) (models/coded/models_2u1.hpp:62-65)."""
    e, u, d, ud = (0, 0), (1, 0), (0, 1), (1, 1)
    phys = Index([e, u, d, ud], [1, 1, 1, 1])
    # dense matrices in the Index order of phys: [ud, u, d, e] (descending)
    order = phys.charges
    pos = {c: i for i, c in enumerate(order)}

    def mat(entries):
        D = np.zeros((4, 4))
        for (cout, cin), v in entries.items():
            D[pos[cout], pos[cin]] = v
        return D

    # |ud> = c+_up c+_dn |0>
    cdag_up = mat({(u, e): 1.0, (ud, d): 1.0})
    cdag_dn = mat({(d, e): 1.0, (ud, u): -1.0})
    F = mat({(e, e): 1.0, (u, u): -1.0, (d, d): -1.0, (ud, ud): 1.0})
    n_up = mat({(u, u): 1.0, (ud, ud): 1.0})
    n_dn = mat({(d, d): 1.0, (ud, ud): 1.0})
    ops = {
        "id": np.eye(4), "F": F,
        "cdag_up": cdag_up, "c_up": cdag_up.T.copy(), "cdag_dn": cdag_dn, "c_dn": cdag_dn.T.copy(),
        "cdag_up_F": cdag_up @ F, "cdag_dn_F": cdag_dn @ F, "F_c_up": F @ cdag_up.T, "F_c_dn": F @ cdag_dn.T,
        "n_up": n_up, "n_dn": n_dn, "docc": n_up @ n_dn, "n": n_up + n_dn,
    }
    return SiteBasis(SYM_TWOU1, phys, {k: Op.from_dense(v, phys, name=k) for k, v in ops.items()})


class MPOTensor:
    """W(b1, b2) = sum of (scale * op); Hermitian pairing and SU2 spins of both bonds."""

    def __init__(self, row_dim, col_dim):
        self.row_dim, self.col_dim = row_dim, col_dim
        self.terms = {}
        self.left_spin = [0] * row_dim
        self.right_spin = [0] * col_dim
        self.left_herm = [None] * row_dim
        self.left_phase = [1] * row_dim
        self.right_herm = [None] * col_dim
        self.right_phase = [1] * col_dim

    def add(self, b1, b2, op, scale=1.0):
        self.terms.setdefault((b1, b2), []).append((op, float(scale)))

    def dense(self, phys):
        """W[b1, b2, out, in] dense (small systems only)."""
        n = phys.total()
        W = np.zeros((self.row_dim, self.col_dim, n, n))
        for (b1, b2), lst in self.terms.items():
            for op, sc in lst:
                W[b1, b2] += sc * op.dense(phys)
        return W


class SparseMPOTensor:
    """W = sum_o C_o (x) o: one sparse (row_dim x col_dim) coefficient matrix per distinct site
    operator.  Same interface as MPOTensor for the descriptor builder (sorted_terms) and the
    dense check; scales to the O(L^2) bond dimensions of ab initio Hamiltonians."""

    def __init__(self, row_dim, col_dim, ops, coeffs):
        import scipy.sparse as sp
        self.row_dim, self.col_dim = row_dim, col_dim
        self.ops = list(ops)
        self.coeffs = [sp.csr_matrix(C, shape=(row_dim, col_dim)) for C in coeffs]
        self.left_spin = [0] * row_dim
        self.right_spin = [0] * col_dim
        self.left_herm = [None] * row_dim
        self.left_phase = [1] * row_dim
        self.right_herm = [None] * col_dim
        self.right_phase = [1] * col_dim
        self.keep_zero = False   # keep explicit zero coefficients as terms (derivatives)

    def sorted_terms(self):
        """Arrays (b1, b2, op_index, scale) sorted by (b1, b2)."""
        b1s, b2s, oi, sc = [], [], [], []
        for k, C in enumerate(self.coeffs):
            Cc = C.tocoo()
            b1s.append(Cc.row)
            b2s.append(Cc.col)
            oi.append(np.full(Cc.nnz, k, dtype=np.int64))
            sc.append(Cc.data)
        if not b1s:
            z = np.zeros(0, np.int64)
            return z, z, z, np.zeros(0)
        b1, b2, oi, sc = (np.concatenate(x) for x in (b1s, b2s, oi, sc))
        if not getattr(self, "keep_zero", False):
            keep = sc != 0
            b1, b2, oi, sc = b1[keep], b2[keep], oi[keep], sc[keep]
        order = np.lexsort((oi, b2, b1))
        return b1[order], b2[order], oi[order], sc[order]

    @property
    def n_cells(self):
        return sum(C.nnz for C in self.coeffs)

    def dense(self, phys):
        n = phys.total()
        W = np.zeros((self.row_dim, self.col_dim, n, n))
        for op, C in zip(self.ops, self.coeffs):
            W += np.einsum("ab,oi->aboi", C.toarray(), op.dense(phys))
        return W

    @staticmethod
    def from_mpo_tensor(W):
        ops, idx, rows, cols, vals = [], {}, [], [], []
        for (b1, b2), lst in W.terms.items():
            for op, s in lst:
                if id(op) not in idx:
                    idx[id(op)] = len(ops)
                    ops.append(op)
                    rows.append([]), cols.append([]), vals.append([])
                k = idx[id(op)]
                rows[k].append(b1), cols[k].append(b2), vals[k].append(s)
        import scipy.sparse as sp
        C = [sp.coo_matrix((vals[k], (rows[k], cols[k])), shape=(W.row_dim, W.col_dim)) for k in range(len(ops))]
        T = SparseMPOTensor(W.row_dim, W.col_dim, ops, C)
        for a in ("left_spin", "right_spin", "left_herm", "left_phase", "right_herm", "right_phase"):
            setattr(T, a, list(getattr(W, a)))
        return T


def twosite_sparse(W1, W2, phys1, phys2):
    """Two-site MPO W12 = sum_{o1,o2} (C_o1 C_o2) (x) (o1 (x) o2) (mp_tensors/ts_ops.hpp:96-168)."""
    if isinstance(W1, MPOTensor):
        W1 = SparseMPOTensor.from_mpo_tensor(W1)
    if isinstance(W2, MPOTensor):
        W2 = SparseMPOTensor.from_mpo_tensor(W2)
    phys12, pb = product_index(phys1, phys2)
    ops, coeffs = [], []
    for o1, C1 in zip(W1.ops, W1.coeffs):
        for o2, C2 in zip(W2.ops, W2.coeffs):
            P = (C1 @ C2).tocsr()
            P.eliminate_zeros()
            if P.nnz == 0:
                continue
            k = _merge(o1.kron(o2, phys1, phys2, phys12, pb))
            if not k.blocks:
                continue
            ops.append(k)
            coeffs.append(P)
    T = SparseMPOTensor(W1.row_dim, W2.col_dim, ops, coeffs)
    T.left_spin, T.left_herm, T.left_phase = list(W1.left_spin), list(W1.left_herm), list(W1.left_phase)
    T.right_spin, T.right_herm, T.right_phase = list(W2.right_spin), list(W2.right_herm), list(W2.right_phase)
    return T, phys12, pb


def _set_pairs(herm, phase, pairs, self_adjoint):
    for a, b, pa, pb in pairs:
        herm[a], herm[b], phase[a], phase[b] = b, a, pa, pb
    for a in self_adjoint:
        herm[a] = a


def heisenberg_mpo(L, J=1.0, Jz=1.0):
    """H = sum_i J/2 (S+_i S-_{i+1} + S-_i S+_{i+1}) + Jz Sz_i Sz_{i+1}, open chain.
    Bond states: 0 start, 1 S+ placed, 2 S- placed, 3 Sz placed, 4 complete."""
    site = spin_half_u1()
    o = site.ops
    mpo = []
    for i in range(L):
        rows = 1 if i == 0 else 5
        cols = 1 if i == L - 1 else 5
        W = MPOTensor(rows, cols)
        r_start, r_done = 0, (0 if i == 0 else 4)
        c_start, c_done = (0 if i < L - 1 else None), (4 if i < L - 1 else 0)

        def C(b):  # column index on the last site collapses to 'complete'
            return b if i < L - 1 else 0
        if i < L - 1:
            W.add(r_start, 0, o["id"])
            W.add(r_start, 1, o["splus"], J / 2)
            W.add(r_start, 2, o["sminus"], J / 2)
            W.add(r_start, 3, o["sz"], Jz)
        if i > 0:
            W.add(1, C(4), o["sminus"])
            W.add(2, C(4), o["splus"])
            W.add(3, C(4), o["sz"])
            W.add(r_done, C(4), o["id"])
        if i == 0 and i == L - 1:
            pass
        mpo.append(W)
    _finish_bonds(mpo, pairs=[(1, 2, 1, 1)], self_adjoint=[0, 3, 4])
    return site, mpo


def hubbard_mpo(L, t=1.0, U=4.0):
    """1d Fermi-Hubbard, open chain: -t sum_{i,s} (c+_{i,s} c_{i+1,s} + h.c.) + U sum_i n_up n_dn
    (models/coded/models_2u1.hpp:125-159).  Bond states: 0 start, 1..4 hopping halves, 5 complete."""
    site = hubbard_twou1()
    o = site.ops
    mpo = []
    left_halves = ["cdag_up_F", "cdag_dn_F", "F_c_up", "F_c_dn"]
    right_halves = ["c_up", "c_dn", "cdag_up", "cdag_dn"]
    for i in range(L):
        rows = 1 if i == 0 else 6
        cols = 1 if i == L - 1 else 6
        W = MPOTensor(rows, cols)
        last = (i == L - 1)
        C = (lambda b: 0) if last else (lambda b: b)
        r_done = 0 if i == 0 else 5
        if not last:
            W.add(0, 0, o["id"])
            for k, name in enumerate(left_halves):
                W.add(0, 1 + k, o[name], -t)
        if i > 0:
            for k, name in enumerate(right_halves):
                W.add(1 + k, C(5), o[name])
            W.add(r_done, C(5), o["id"])
        # on-site term from start to complete
        W.add(0, C(5), o["docc"], U)
        mpo.append(W)
    _finish_bonds(mpo, pairs=[(1, 3, 1, 1), (2, 4, 1, 1)], self_adjoint=[0, 5])
    return site, mpo


def bosons_u1(nmax):
    """Soft-core bosons with U1 charge n, 0 <= n <= nmax (models/coded/models_u1.hpp:316-340)."""
    phys = Index([(n, 0) for n in range(nmax + 1)], [1] * (nmax + 1))
    ops = {"id": Op(name="id"), "create": Op(name="create"), "destroy": Op(name="destroy"),
           "count": Op(name="count"), "interaction": Op(name="interaction")}
    for n in range(nmax + 1):
        c = (n, 0)
        ops["id"].add(c, c, 0, 0, 1.0)
        if n > 0:
            ops["count"].add(c, c, 0, 0, float(n))
            ops["interaction"].add(c, c, 0, 0, float(n * n - n))
            ops["create"].add((n - 1, 0), c, 0, 0, float(np.sqrt(n)))
            ops["destroy"].add(c, (n - 1, 0), 0, 0, float(np.sqrt(n)))
    return SiteBasis(SYM_U1, phys, ops)


def boson_hubbard_mpo(L, nmax=2, t=1.0, U=8.0, V=0.0):
    """Coded boson Hubbard (models/coded/models_u1.hpp:359-392): U/2 n(n-1) on every site,
    -t (b+_p b_{p+1} + b_p b+_{p+1}) and V n_p n_{p+1} on every bond.
    Bond states: 0 start, 1 b+ placed, 2 b placed, 3 n placed, 4 complete."""
    site = bosons_u1(nmax)
    o = site.ops
    mpo = []
    for i in range(L):
        last = (i == L - 1)
        W = MPOTensor(1 if i == 0 else 5, 1 if last else 5)
        C = (lambda b: 0) if last else (lambda b: b)
        if not last:
            W.add(0, 0, o["id"])
            W.add(0, 1, o["create"], -t)
            W.add(0, 2, o["destroy"], -t)
            if V != 0.0:
                W.add(0, 3, o["count"], V)
        if i > 0:
            W.add(1, C(4), o["destroy"])
            W.add(2, C(4), o["create"])
            if V != 0.0:
                W.add(3, C(4), o["count"])
            W.add(4, C(4), o["id"])
        W.add(0, C(4), o["interaction"], U / 2)
        mpo.append(W)
    _finish_bonds(mpo, pairs=[(1, 2, 1, 1)], self_adjoint=[0, 3, 4])
    return site, mpo


def identity_mpo(phys, L):
    """The identity MPO (bond dimension 1) on L sites of the physical basis phys."""
    ident = Op.from_dense(np.eye(phys.total()), phys, name="id")
    out = []
    for _ in range(L):
        W = MPOTensor(1, 1)
        W.add(0, 0, ident)
        out.append(W)
    return out


def chain(model, L, **kw):
    """(site basis, MPO, total charge) of the benchmark chains."""
    if model == "heisenberg":
        basis, mpo = heisenberg_mpo(L, J=kw.get("J", 1.0), Jz=kw.get("Jz", kw.get("J", 1.0)))
        return basis, mpo, (kw.get("sz2", 0), 0)
    if model == "hubbard":
        basis, mpo = hubbard_mpo(L, t=kw.get("t", 1.0), U=kw.get("U", 4.0))
        return basis, mpo, (kw.get("nup", L // 2), kw.get("ndown", L // 2))
    if model == "bosehubbard":
        basis, mpo = boson_hubbard_mpo(L, nmax=kw.get("nmax", 2), t=kw.get("t", 1.0), U=kw.get("U", 8.0),
                                       V=kw.get("V", 0.0))
        return basis, mpo, (kw.get("N", L // 2), 0)
    if model == "qc":
        from . import qc
        if "fcidump" in kw:
            h, eri, ecore = qc.read_fcidump(kw["fcidump"])
        else:
            h, eri = qc.synthetic_integrals(L, seed=kw.get("integral_seed", 0))
            ecore = kw.get("ecore", 0.0)
        basis, mpo, _ = qc.qc_mpo(h, eri, ecore)
        N = kw.get("N", L)
        return basis, mpo, (N // 2, N - N // 2)
    if model == "qc_su2":
        from . import qc, su2
        if "fcidump" in kw:
            h, eri, ecore = qc.read_fcidump(kw["fcidump"])
        else:
            h, eri = qc.synthetic_integrals(L, seed=kw.get("integral_seed", 0))
            ecore = kw.get("ecore", 0.0)
        N = kw.get("N", L)
        basis, mpo, _ = su2.qc_su2_mpo(h, eri, N, ecore)
        # lowest total spin the electron count allows (singlet for even N, doublet for odd N)
        return basis, mpo, (N, kw.get("two_s", N % 2))
    raise ValueError(model)


def _finish_bonds(mpo, pairs, self_adjoint):
    """Register Hermitian partners on every internal bond (bulk bond numbering)."""
    L = len(mpo)
    for i in range(L - 1):
        W, Wn = mpo[i], mpo[i + 1]
        herm = [None] * W.col_dim
        phase = [1] * W.col_dim
        _set_pairs(herm, phase, pairs, self_adjoint)
        W.right_herm, W.right_phase = list(herm), list(phase)
        Wn.left_herm, Wn.left_phase = list(herm), list(phase)
    # edge bonds: the single start / complete state is self-adjoint
    mpo[0].left_herm = [0]
    mpo[-1].right_herm = [0]


def twosite_mpo(W1, W2, phys1, phys2):
    """make_twosite_mpo (mp_tensors/ts_ops.hpp:96-168): W12(b0, b2) = sum_b1 W1(b0,b1) (x) W2(b1,b2),
    summed into one operator per (b0, b2)."""
    phys12, pb = product_index(phys1, phys2)
    W = MPOTensor(W1.row_dim, W2.col_dim)
    acc = {}
    for (b0, b1), l1 in W1.terms.items():
        for (b1b, b2), l2 in W2.terms.items():
            if b1b != b1:
                continue
            for op1, s1 in l1:
                for op2, s2 in l2:
                    k = op1.kron(op2, phys1, phys2, phys12, pb)
                    tgt = acc.setdefault((b0, b2), Op(name="ts"))
                    for key, ents in k.blocks.items():
                        for r, c, v, rs, cs in ents:
                            tgt.add(key[0], key[1], r, c, v * s1 * s2, rs, cs)
    for key in sorted(acc):
        op = _merge(acc[key])
        if op.blocks:
            W.add(key[0], key[1], op, 1.0)
    W.left_spin, W.left_herm, W.left_phase = W1.left_spin, W1.left_herm, W1.left_phase
    W.right_spin, W.right_herm, W.right_phase = W2.right_spin, W2.right_herm, W2.right_phase
    return W, phys12, pb


def mpo_dense_chain(mpo, phys):
    """Full Hamiltonian matrix of a short chain (dense, for exact checks)."""
    Ws = [W.dense(phys) for W in mpo]
    # H = W0[0, :] * W1 * ... * W_{L-1}[:, 0]
    cur = Ws[0][0]          # (b, out, in)
    for Wk in Ws[1:]:
        # cur: (b, O, I), Wk: (b, c, o, i) -> (c, O*o, I*i)
        nb = np.einsum("bOI,bcoi->cOoIi", cur, Wk)
        c, O, o, I, i = nb.shape
        cur = nb.reshape(c, O * o, I * i)
    return cur[0]
