"""Ab initio (quantum chemistry) Hamiltonians for the benchmark configs.

* synthetic integrals exactly as the benchmark configurations specify them;
* FCIDUMP-style integral files in the reference parser's format (models/chem/parse_integrals.h:36-62:
  four header lines, then ``value i j k l`` with 1-based orbital indices, 0 for an absent index,
  core energy as ``value 0 0 0 0``);
* an MPO of H = sum h_pq c+_ps c_qs + 1/2 sum (pq|rs) c+_ps c+_rt c_st c_qs + E_core in the
  TwoU1 (n_up, n_down) site basis, built by a two-sided operator-string automaton: a term's bond
  state is its left operator string while that is the shorter side and its right operator string
  afterwards; the coefficient is applied once, at the switch.  Bond states with identical strings
  are shared, which gives a bond dimension O(L^2) (the reference obtains its O(L^2) bond
  dimension with the tagged MPO maker, models/generate_mpo/tagged_mpo_maker_optim.hpp).
* a determinant-space exact diagonalisation (FCI) for small checks.
"""
import itertools

import numpy as np

from . import models
from .models import MPOTensor, Op
from .symmetry import Index, SYM_TWOU1


# ---------------------------------------------------------------------------- integrals
def synthetic_integrals(L, seed=0, n_aux=48):
    """h_ii ~ U[-2,-1]; h_{i,i+-1} = -1; other h_ij symmetric N(0, 0.1);
    (ij|kl) = sum_P L^P_ij L^P_kl, L^P symmetric with L^P_ij ~ N(0, 0.3) exp(-|i-j|/4)."""
    rng = np.random.default_rng(seed)
    h = rng.normal(0.0, 0.1, (L, L))
    h = np.triu(h, 2)
    h = h + h.T
    for i in range(L):
        h[i, i] = rng.uniform(-2.0, -1.0)
    for i in range(L - 1):
        h[i, i + 1] = h[i + 1, i] = -1.0
    idx = np.arange(L)
    damp = np.exp(-np.abs(idx[:, None] - idx[None, :]) / 4.0)
    eri = np.zeros((L, L, L, L))
    for _ in range(n_aux):
        Lp = rng.normal(0.0, 0.3, (L, L))
        Lp = np.triu(Lp)
        Lp = Lp + np.triu(Lp, 1).T
        Lp *= damp
        eri += np.einsum("ij,kl->ijkl", Lp, Lp)
    return h, eri


def write_fcidump(path, h, eri, ecore=0.0, tol=1e-14):
    """Reference parser format: 4 header lines; unique 8-fold (ij|kl) and 2-fold h_ij elements."""
    L = h.shape[0]
    with open(path, "w") as f:
        f.write(" &FCI NORB=%d,\n  ORBSYM=%s\n  ISYM=1,\n &END\n" % (L, ",".join(["1"] * L) + ","))
        for i, j, k, l in itertools.product(range(L), repeat=4):
            if i < j or k < l or (i * L + j) < (k * L + l):
                continue
            v = eri[i, j, k, l]
            if abs(v) > tol:
                f.write("%23.16e %d %d %d %d\n" % (v, i + 1, j + 1, k + 1, l + 1))
        for i, j in itertools.product(range(L), repeat=2):
            if i < j:
                continue
            if abs(h[i, j]) > tol:
                f.write("%23.16e %d %d 0 0\n" % (h[i, j], i + 1, j + 1))
        f.write("%23.16e 0 0 0 0\n" % ecore)


def read_fcidump(path):
    """parse_integrals.h:36-62 semantics: skip 4 header lines, expand the 8-fold symmetry."""
    rows = []
    with open(path) as f:
        lines = f.readlines()[4:]
    for ln in lines:
        p = ln.split()
        if len(p) < 5:
            continue
        rows.append((float(p[0]), int(p[1]), int(p[2]), int(p[3]), int(p[4])))
    L = max(max(r[1:]) for r in rows)
    h = np.zeros((L, L))
    eri = np.zeros((L, L, L, L))
    ecore = 0.0
    for v, i, j, k, l in rows:
        if i == j == k == l == 0:
            ecore = v
        elif k == 0 and l == 0:
            h[i - 1, j - 1] = h[j - 1, i - 1] = v
        else:
            i, j, k, l = i - 1, j - 1, k - 1, l - 1
            for a, b, c, d in [(i, j, k, l), (j, i, k, l), (i, j, l, k), (j, i, l, k),
                               (k, l, i, j), (l, k, i, j), (k, l, j, i), (l, k, j, i)]:
                eri[a, b, c, d] = v
    return h, eri, ecore


# ---------------------------------------------------------------------------- site basis
def fermion_site():
    """TwoU1 site (n_up, n_down) with local ops; order of fermionic modes inside a site: up, down."""
    basis = models.hubbard_twou1()
    phys = basis.phys
    pos = {c: i for i, c in enumerate(phys.charges)}
    e, u, d, ud = (0, 0), (1, 0), (0, 1), (1, 1)

    def mat(entries):
        D = np.zeros((4, 4))
        for (cout, cin), v in entries.items():
            D[pos[cout], pos[cin]] = v
        return D
    cdag = [mat({(u, e): 1.0, (ud, d): 1.0}), mat({(d, e): 1.0, (ud, u): -1.0})]
    F = mat({(e, e): 1.0, (u, u): -1.0, (d, d): -1.0, (ud, ud): 1.0})
    return phys, cdag, [c.T.copy() for c in cdag], F


# ---------------------------------------------------------------------------- terms
def qc_terms(h, eri, tol=1e-14, with_source=False, h_tol=None):
    """Normal-ordered operator strings of H: list of (coeff, ((orb, spin, dagger), ...)).
    with_source: each entry also carries (kind, indices, factor) with coeff = factor * integral,
    kind "h" (p, q) or "eri" (p, q, r, s).  h_tol (default tol): threshold of the one-body entries;
    a negative value keeps every h_pq, zeros included, so that the MPO's sources cover the whole
    one-body density matrix dE/dh."""
    L = h.shape[0]
    terms = []
    h_tol = tol if h_tol is None else h_tol
    for p, q in itertools.product(range(L), repeat=2):
        if abs(h[p, q]) > h_tol:
            for s in range(2):
                t = (h[p, q], ((p, s, 1), (q, s, 0)))
                terms.append(t + (("h", (p, q), 1.0),) if with_source else t)
    nz = np.argwhere(np.abs(eri) > tol)
    for p, q, r, s_ in nz:
        v = 0.5 * eri[p, q, r, s_]
        for a in range(2):
            for b in range(2):
                if p == r and a == b:
                    continue        # c+_pa c+_pa = 0
                if s_ == q and a == b:
                    continue        # c_qa c_qa = 0
                t = (v, ((p, a, 1), (r, b, 1), (s_, b, 0), (q, a, 0)))
                terms.append(t + (("eri", (int(p), int(q), int(r), int(s_)), 0.5),) if with_source else t)
    return terms


class _OpTable:
    """Dedupe local 4x4 operators by content."""

    def __init__(self):
        self.mats, self.ids = [], {}

    def id(self, M):
        key = np.round(M, 14).tobytes()
        if key not in self.ids:
            self.ids[key] = len(self.mats)
            self.mats.append(M.copy())
        return self.ids[key]


def _site_order(ops):
    """Stable sort of a fermion operator string by orbital; returns (sign, sorted ops)."""
    ops = list(ops)
    sign = 1
    n = len(ops)
    for i in range(n):               # bubble sort keeps the relative order of equal orbitals
        for j in range(n - 1 - i):
            if ops[j][0] > ops[j + 1][0]:
                ops[j], ops[j + 1] = ops[j + 1], ops[j]
                sign = -sign
    return sign, ops


def qc_mpo(h, eri, ecore=0.0, tol=1e-14, sources=False, h_tol=None):
    """MPO of the ab initio Hamiltonian in the TwoU1 site basis (see module docstring).

    sources=True: also return, per site j, {(b1, b2, op_id): [(kind, indices, factor), ...]}: the
    integrals summed into each switch coefficient (coefficient = sum factor * integral); each
    tensor then carries op_ids (the table id of each of its ops) and keeps zero coefficients, so
    that the derivative with respect to every MPO coefficient (SigmaPlan.adjoint) maps back to the
    integrals (mpo_gradient_to_integrals)."""
    L = h.shape[0]
    phys, cdag, c, F = fermion_site()
    I4 = np.eye(4)
    table = _OpTable()
    idI, idF = table.id(I4), table.id(F)
    # ---- terms as site-local factors: T = coef * (x)_site Y_site, Y = X * F^(parity right of site)
    site_terms = {}
    local_cache = {}

    def local_op(pattern, odd_right):
        """op id of the product of (spin, dagger) factors at one site, times F if odd_right."""
        k = (pattern, odd_right)
        if k not in local_cache:
            X = I4
            for (s, dg) in pattern:
                X = X @ (cdag[s] if dg else c[s])
            if odd_right:
                X = X @ F
            local_cache[k] = table.id(X) if np.any(X) else None
        return local_cache[k]

    key_src = {}
    for item in qc_terms(h, eri, tol, with_source=sources, h_tol=h_tol):
        coef, ops = item[0], item[1]
        sign, sops = _site_order(ops)
        loc = []
        n = len(sops)
        i = 0
        while i < n:
            st = sops[i][0]
            j = i
            while j < n and sops[j][0] == st:
                j += 1
            oid = local_op(tuple((s, dg) for (_, s, dg) in sops[i:j]), (n - j) % 2)
            if oid is None:
                loc = None
                break
            loc.append((st, oid))
            i = j
        if loc is None:
            continue
        key = tuple(loc)
        site_terms[key] = site_terms.get(key, 0.0) + sign * coef
        if sources:
            kind, idx, fac = item[2]
            key_src.setdefault(key, []).append((kind, idx, sign * fac))
    if abs(ecore) > 0:
        site_terms[((0, idI),)] = site_terms.get(((0, idI),), 0.0) + ecore
        if sources:
            key_src.setdefault(((0, idI),), []).append(("ecore", (), 1.0))
    # fermionic parity of each distinct local operator
    parities = np.array([(a + b) % 2 for (a, b) in phys.charges])

    def op_parity(oid):
        nz = np.argwhere(np.abs(table.mats[oid]) > 0)
        return int(any(parities[i] != parities[j] for i, j in nz)) if len(nz) else 0
    op_par = [op_parity(k) for k in range(len(table.mats))]
    auto = operator_automaton(site_terms, L, op_par, lambda par, spin2: idF if par % 2 else idI,
                              sources=key_src if sources else None)
    ops = [Op.from_dense(M, phys, name="qc%d" % k) for k, M in enumerate(table.mats)]
    mpo = auto.tensors(ops, keep_zero=sources)
    basis = models.SiteBasis(SYM_TWOU1, phys, {})
    if sources:
        return basis, mpo, [len(b) for b in auto.bonds], auto.switch_src
    return basis, mpo, [len(b) for b in auto.bonds]


class operator_automaton:
    """Bond states and transitions of an MPO built from operator strings (site_terms: {((site, op),
    ...) ascending sites: coefficient}).  A term's bond state is its left operator string ("P")
    while that is the shorter side and its right operator string ("S") afterwards; the
    coefficient is applied once, at the switch; strings are shared between terms, which gives a
    bond dimension O(L^2) for ab initio Hamiltonians.  Between operators a string passes a site
    with pass_op(parity of the operators on its left, bond spin).  action[op] is the change of
    the doubled bond spin across op (0 for abelian groups); the bond spin of a P state is the sum
    of its actions, of an S state minus the sum.  sources: optional {key: list} accumulated per
    switch coefficient (switch_src)."""

    START, DONE = ("start",), ("done",)

    def __init__(self, site_terms, L, op_par, pass_op, action=None, sources=None, tol=1e-14):
        self.L = L
        self.action = action
        self.bonds = [dict() for _ in range(L + 1)]
        START, DONE = self.START, self.DONE
        for k in range(L + 1):
            if k < L:
                self._state(k, START)
            if k > 0:
                self._state(k, DONE)
        self.switch = [dict() for _ in range(L)]     # (a, b, op) -> summed coefficient
        self.fixed = [dict() for _ in range(L)]      # (a, b) -> op, transitions between shared states
        self.switch_src = [dict() for _ in range(L)]
        p_exit, s_entry = {}, {}                     # P key -> last bond alive; S key -> first bond alive
        mid = L // 2

        def key_parity(key):
            return sum(op_par[o] for _, o in key) % 2

        def key_spin(key):
            return sum(action[o] for _, o in key) if action is not None else 0

        def pass_of(state):
            par = key_parity(state[1])
            spin = key_spin(state[1]) * (1 if state[0] == "P" else -1)
            return pass_op(par, spin)

        cur = {"src": []}

        def trans(j, prev, nxt, oid, coef):
            a, b = self._state(j, prev), self._state(j + 1, nxt)
            if prev[0] in ("start", "P") and nxt[0] in ("S", "done"):
                k3 = (a, b, oid)
                self.switch[j][k3] = self.switch[j].get(k3, 0.0) + coef
                if sources is not None:
                    self.switch_src[j].setdefault(k3, []).extend(cur["src"])
            else:
                self.fixed[j][(a, b)] = oid

        for key, coef in site_terms.items():
            if abs(coef) < tol and sources is None:
                continue
            cur["src"] = sources.get(key, []) if sources is not None else []
            n = len(key)
            prev = START
            for t in range(n):
                s, oid = key[t]
                tl, tr = t + 1, n - t - 1
                if t == n - 1:
                    nxt = DONE
                elif tl < tr or (tl == tr and s + 1 <= mid):
                    nxt = ("P", key[:tl])
                else:
                    nxt = ("S", key[tl:])
                trans(s, prev, nxt, oid, coef)
                prev = nxt
                if t == n - 1:
                    break
                s_next = key[t + 1][0]
                if nxt[0] == "P":
                    if tl == tr and s_next > mid:
                        # tie gap straddling the middle: switch at pass-through site mid
                        p_exit[nxt] = max(p_exit.get(nxt, 0), mid)
                        snew = ("S", key[tl:])
                        trans(mid, nxt, snew, pass_of(nxt), coef)
                        s_entry[snew] = min(s_entry.get(snew, L + 1), mid + 1)
                        prev = snew
                    else:
                        p_exit[nxt] = max(p_exit.get(nxt, 0), s_next)
                else:
                    s_entry[nxt] = min(s_entry.get(nxt, L + 1), s + 1)
        # pass-through of shared states over sites without operators
        for key, ex in p_exit.items():
            first = key[1][-1][0] + 1
            oid = pass_of(key)
            for k in range(first, ex):
                self.fixed[k][(self._state(k, key), self._state(k + 1, key))] = oid
        for key, en in s_entry.items():
            last = key[1][0][0]
            oid = pass_of(key)
            for k in range(en, last):
                self.fixed[k][(self._state(k, key), self._state(k + 1, key))] = oid
        ident = pass_op(0, 0)
        for j in range(L):
            if j < L - 1:
                self.fixed[j][(self.bonds[j][START], self.bonds[j + 1][START])] = ident
            if j > 0:
                self.fixed[j][(self.bonds[j][DONE], self.bonds[j + 1][DONE])] = ident
        # doubled bond spin of every state
        self.spins = []
        for d in self.bonds:
            sp_ = [0] * len(d)
            for key, i in d.items():
                if key[0] in ("P", "S"):
                    sp_[i] = key_spin(key[1]) * (1 if key[0] == "P" else -1)
            self.spins.append(sp_)

    def _state(self, k, key):
        d = self.bonds[k]
        i = d.get(key)
        if i is None:
            i = d[key] = len(d)
        return i

    def tensors(self, ops, keep_zero=False):
        """Sparse MPO tensors: one coefficient matrix per distinct local operator of each site."""
        import scipy.sparse as sp
        mpo = []
        for j in range(self.L):
            rows, cols = len(self.bonds[j]), len(self.bonds[j + 1])
            per_op = {}
            for (a, b), oid in self.fixed[j].items():
                per_op.setdefault(oid, []).append((a, b, 1.0))
            for (a, b, oid), v in self.switch[j].items():
                per_op.setdefault(oid, []).append((a, b, v))
            used, coeffs, op_ids = [], [], []
            for oid, ents in sorted(per_op.items()):
                r, c_, v = zip(*ents)
                C = sp.coo_matrix((v, (r, c_)), shape=(rows, cols)).tocsr()
                C.sum_duplicates()
                used.append(ops[oid])
                coeffs.append(C)
                op_ids.append(oid)
            T = models.SparseMPOTensor(rows, cols, used, coeffs)
            T.left_spin, T.right_spin = list(self.spins[j]), list(self.spins[j + 1])
            if keep_zero:
                T.op_ids = op_ids
                T.keep_zero = True
            mpo.append(T)
        return mpo


def mpo_gradient_to_integrals(mpo, switch_src, term_grads, L):
    """dE/dh (L x L), dE/d(pq|rs) (L^4) and dE/decore from the derivatives with respect to the
    MPO term scales of every site (term_grads[j] in the order of mpo[j].sorted_terms()), through
    the integral sources of each switch coefficient (qc_mpo(..., sources=True)).  Coefficients
    that are 1 by construction (operator-string pass-through) carry no integral."""
    dh = np.zeros((L, L))
    deri = np.zeros((L, L, L, L))
    decore = 0.0
    for j, (W, g) in enumerate(zip(mpo, term_grads)):
        b1s, b2s, ois, _ = W.sorted_terms()
        assert len(g) == len(b1s)
        for t in range(len(b1s)):
            for kind, idx, fac in switch_src[j].get((int(b1s[t]), int(b2s[t]), W.op_ids[int(ois[t])]), ()):
                if kind == "h":
                    dh[idx] += fac * g[t]
                elif kind == "eri":
                    deri[idx] += fac * g[t]
                else:
                    decore += fac * g[t]
    return dh, deri, decore


# ---------------------------------------------------------------------------- exact check
def fci_energy(h, eri, n_up, n_down, ecore=0.0):
    """Lowest eigenvalue in the (n_up, n_down) sector from the explicit Jordan-Wigner matrix
    (2L spin orbitals, orbital-major, up before down).  Small L only."""
    L = h.shape[0]
    n = 2 * L
    dets = []
    for bits in range(1 << n):
        nu = sum((bits >> (2 * i)) & 1 for i in range(L))
        nd = sum((bits >> (2 * i + 1)) & 1 for i in range(L))
        if nu == n_up and nd == n_down:
            dets.append(bits)
    index = {b: k for k, b in enumerate(dets)}
    H = np.zeros((len(dets), len(dets)))

    def apply(ops, bits):
        sign = 1
        for (mode, dg) in reversed(ops):
            occ = (bits >> mode) & 1
            if dg and occ or (not dg and not occ):
                return 0, None
            sign *= (-1) ** bin(bits & ((1 << mode) - 1)).count("1")
            bits ^= (1 << mode)
        return sign, bits
    for coef, ops in qc_terms(h, eri):
        modes = [(2 * o + s, dg) for (o, s, dg) in ops]
        for k, b in enumerate(dets):
            sg, nb = apply(modes, b)
            if sg:
                H[index[nb], k] += coef * sg
    return float(np.linalg.eigvalsh(H)[0]) + ecore


def hermitian_partner(kind, idx):
    """The entry whose operator is the Hermitian conjugate of this entry's (real orbitals):
    h_pq c+_p c_q <-> h_qp, (pq|rs) c+_p c+_r c_s c_q <-> (qp|sr)."""
    if kind == "h":
        return (idx[1], idx[0])
    p, q, r, s = idx
    return (q, p, s, r)


def fci_energy_gradient_fd(h, eri, n_up, n_down, step=1e-5, entries=None):
    """Central finite differences of the FCI energy with respect to single integral entries
    (each h[p, q] and (pq|rs) entry an independent parameter).  A single off-diagonal entry alone
    makes H non-Hermitian, which has no variational lowest eigenvalue; to first order in the step
    its derivative is psi . (dH/dx) psi, the same for the entry and its Hermitian partner
    (hermitian_partner), so the entry and its partner are stepped together and the difference
    halved when they are distinct.  entries: list of ("h", (p, q)) / ("eri", (p, q, r, s));
    default all.  Returns {entry: dE/dx}."""
    L = h.shape[0]
    if entries is None:
        entries = [("h", ij) for ij in itertools.product(range(L), repeat=2)] + \
                  [("eri", ijkl) for ijkl in itertools.product(range(L), repeat=4)]
    out = {}
    for kind, idx in entries:
        idx = tuple(int(x) for x in idx)
        cls = {idx, hermitian_partner(kind, idx)}
        vals = []
        for sgn in (+1, -1):
            hh, ee = h.copy(), eri.copy()
            for x in cls:
                (hh if kind == "h" else ee)[x] += sgn * step
            vals.append(fci_energy(hh, ee, n_up, n_down))
        out[(kind, idx)] = (vals[0] - vals[1]) / (2 * step) / len(cls)
    return out
