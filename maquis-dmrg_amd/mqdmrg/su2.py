"""Spin-adapted (SU2U1) ab initio Hamiltonians, their two-site MPOs and two-site reduction.

The site basis, local operators and term list follow the reference's qc_su2 model
(models/chem/su2u1/model.hpp:33-580 with term_maker.h:34-463): charges are (N, 2S-shift), a site
holds |20>, |11>, |1-1>, |00> (the two singly occupied states are the two coupling paths of the
spin-1/2 multiplet), and each operator carries a SpinDescriptor (rank 2k, bond spin change).
The one-electron hopping is folded into two-electron terms with the particle number N as the
reference does (model.hpp:381-410), so the MPO represents H only in the N-electron sector.

The MPO is built by the operator-string automaton of qc.py (bond states = shared prefixes /
suffixes of the terms) with the reference's identity rules on the strings between operators
(tagged_mpo_maker_optim.hpp:500-522): fill where an odd number of fermion operators remains,
ident_full where the bond spin exceeds 1/2, ident otherwise.  Bond spins accumulate the
SpinDescriptor actions (block_matrix/symmetry/spin_descriptor.h:79-85, tagged maker :250-258).

The two-site MPO couples the site operators with 9j symbols and the bond spins with a 6j
(block_matrix/site_operator_algorithms.h:214-322 op_kron, mp_tensors/ts_ops.hpp:38-168
make_twosite_mpo), and the two-site tensor is recoupled by reduce_right / unreduce_right
(mp_tensors/ts_reduction.h:35-335) before and after the site problem."""
import itertools
import math
from fractions import Fraction
from functools import lru_cache

import numpy as np
import scipy.sparse as sp

from .models import Op, SparseMPOTensor, SiteBasis
from .symmetry import Index, SYM_SU2U1, fuse, product_index, SiteLayout

A_, B_, C_, D_ = (2, 0), (1, 1), (1, -1), (0, 0)


# ---------------------------------------------------------------------------- Wigner symbols
def _tri(a, b, c):
    """Triangle condition on doubled spins."""
    return abs(a - b) <= c <= a + b and (a + b + c) % 2 == 0


def _delta2(a, b, c):
    f = math.factorial
    return Fraction(f((a + b - c) // 2) * f((a - b + c) // 2) * f((-a + b + c) // 2), f((a + b + c) // 2 + 1))


@lru_cache(maxsize=None)
def sixj(j1, j2, j3, j4, j5, j6):
    """Wigner 6j symbol {j1 j2 j3; j4 j5 j6} of doubled spins (Racah's sum, exact rationals)."""
    if not (_tri(j1, j2, j3) and _tri(j1, j5, j6) and _tri(j4, j2, j6) and _tri(j4, j5, j3)):
        return 0.0
    f = math.factorial
    a = [(j1 + j2 + j3) // 2, (j1 + j5 + j6) // 2, (j4 + j2 + j6) // 2, (j4 + j5 + j3) // 2]
    b = [(j1 + j2 + j4 + j5) // 2, (j2 + j3 + j5 + j6) // 2, (j3 + j1 + j6 + j4) // 2]
    s = Fraction(0)
    for t in range(max(a), min(b) + 1):
        den = 1
        for x in a:
            den *= f(t - x)
        for x in b:
            den *= f(x - t)
        s += Fraction((-1) ** t * f(t + 1), den)
    d2 = _delta2(j1, j2, j3) * _delta2(j1, j5, j6) * _delta2(j4, j2, j6) * _delta2(j4, j5, j3)
    return float(s) * math.sqrt(float(d2))


@lru_cache(maxsize=None)
def ninej(a, b, c, d, e, f, g, h, i):
    """Wigner 9j symbol of doubled spins as a sum over products of three 6j symbols."""
    if not (_tri(a, b, c) and _tri(d, e, f) and _tri(g, h, i) and _tri(a, d, g) and _tri(b, e, h)
            and _tri(c, f, i)):
        return 0.0
    lo = max(abs(a - i), abs(d - h), abs(b - f))
    hi = min(a + i, d + h, b + f)
    s = 0.0
    for x in range(lo, hi + 1, 2):
        t = (x + 1) * sixj(a, b, c, f, i, x) * sixj(d, e, f, b, x, h) * sixj(g, h, i, x, a, d)
        s += -t if x % 2 else t
    return s


def mod_coupling(a, b, c, d, e, f, g, h, i):
    """sqrt((g+1)(h+1)(c+1)(f+1)) {9j} (block_matrix/symmetry/gsl_coupling.h:38-47)."""
    return math.sqrt((g + 1.) * (h + 1.) * (c + 1.) * (f + 1.)) * ninej(a, b, c, d, e, f, g, h, i)


def product_spin(a, b):
    """Spin of a two-site physical state from the spin shifts of its sites
    (block_matrix/symmetry/spin_descriptor.h:126-136)."""
    if a[1] == -1 and b[1] == 1:
        return 2
    return abs(a[1] + b[1])


# ---------------------------------------------------------------------------- site and operators
def _op(name, entries, spin=0, action=0, fermionic=False):
    """Op whose blocks map the input (ket) charge to the output (bra) charge; the entry spins are
    the |spin shift| of the block charges (site_operator.hpp:183-195)."""
    op = Op(spin=spin, name=name)
    for (cin, cout), v in entries.items():
        op.add(cin, cout, 0, 0, float(v), abs(cin[1]), abs(cout[1]))
    op.action = action
    op.fermionic = fermionic
    return op


def su2_site():
    """Site basis and operator table of qc_su2 (model.hpp:33-319): name -> Op."""
    phys = Index([A_, B_, C_, D_], [1, 1, 1, 1])
    s2, s32 = math.sqrt(2.0), math.sqrt(1.5)
    A, B, C, D = A_, B_, C_, D_
    ops = {}

    def reg(name, entries, spin=0, action=0, ferm=False):
        ops[name] = _op(name, entries, spin, action, ferm)

    reg("ident", {(A, A): 1, (B, B): 1, (C, C): 1, (D, D): 1})
    reg("ident_full", {(A, A): 1, (D, D): 1, (B, B): 1, (C, C): 1, (B, C): 1, (C, B): 1})
    reg("fill", {(A, A): 1, (D, D): 1, (B, B): -1, (C, C): -1, (B, C): -1, (C, B): -1})
    cf = {(B, A): s2, (C, A): s2, (D, B): 1, (D, C): 1}
    de = {(A, B): 1, (A, C): 1, (B, D): s2, (C, D): s2}
    df = {(A, B): 1, (A, C): 1, (B, D): -s2, (C, D): -s2}
    cr = {(B, A): s2, (C, A): s2, (D, B): -1, (D, C): -1}
    reg("create_fill", cf, 1, +1, True)
    reg("destroy", de, 1, -1, True)
    reg("destroy_fill", df, 1, +1, True)
    reg("create", cr, 1, -1, True)
    reg("create_fill_couple_down", cf, 1, -1, True)
    reg("destroy_fill_couple_down", df, 1, -1, True)
    reg("create_couple_up", cr, 1, +1, True)
    reg("destroy_couple_up", de, 1, +1, True)
    reg("create_fill_count", {(B, A): s2, (C, A): s2}, 1, +1, True)
    reg("destroy_count", {(A, B): 1, (A, C): 1}, 1, -1, True)
    reg("destroy_fill_count", {(A, B): 1, (A, C): 1}, 1, +1, True)
    reg("create_count", {(B, A): s2, (C, A): s2}, 1, -1, True)
    reg("count", {(A, A): 2, (B, B): 1, (C, C): 1})
    reg("docc", {(A, A): 1})
    reg("e2d", {(D, A): 1})
    reg("d2e", {(A, D): 1})
    reg("count_fill", {(A, A): 2, (B, B): -1, (C, C): -1, (B, C): -1, (C, B): -1})
    flip = {(B, B): s32, (C, C): s32, (B, C): s32, (C, B): s32}
    reg("flip_to_S2", flip, 2, +2)
    reg("flip_to_S0", flip, 2, -2)
    reg("flip_S0", flip, 2, 0)
    return SiteBasis(SYM_SU2U1, phys, ops)


class _Bundle:
    """OperatorBundle of term_maker.h:50-59 (operator names)."""

    def __init__(self, couple_up=None, couple_down=None, fill_couple_up=None, fill_couple_down=None,
                 no_couple=None, fill_no_couple=None):
        self.couple_up, self.couple_down = couple_up, couple_down
        self.fill_couple_up, self.fill_couple_down = fill_couple_up, fill_couple_down
        self.no_couple, self.fill_no_couple = no_couple, fill_no_couple


# OperatorCollection (model.hpp:285-318)
_OPS = {
    "create": _Bundle("create_couple_up", "create", "create_fill", "create_fill_couple_down"),
    "destroy": _Bundle("destroy_couple_up", "destroy", "destroy_fill", "destroy_fill_couple_down"),
    "create_count": _Bundle(couple_down="create_count", fill_couple_up="create_fill_count"),
    "destroy_count": _Bundle(couple_down="destroy_count", fill_couple_up="destroy_fill_count"),
    "count": _Bundle(no_couple="count", fill_no_couple="count_fill"),
    "flip": _Bundle("flip_to_S2", "flip_to_S0", no_couple="flip_S0"),
}


# ---------------------------------------------------------------------------- terms
def _sgn(i, j, k, l):
    """Odd permutation parity of (i, j, k, l) (term_maker.h:80-91)."""
    idx = (i, j, k, l)
    inv = sum(1 for a in range(3) for b in range(a + 1, 4) if idx[a] > idx[b])
    return inv % 2 != 0


def _align(i, j, k, l):
    """chem::align (models/chem/util.h:110-116)."""
    if i < j:
        i, j = j, i
    if k < l:
        k, l = l, k
    if i < k:
        i, k = k, i
        j, l = l, j
    if i == k and j < l:
        j, l = l, j
    return i, j, k, l


def _t2(scale, i, j, op1, op2):                       # TermMakerSU2::two_term
    return (scale, ((i, op1), (j, op2)))


def _pos2(sign, scale, i, j, op1, op1_fill, op2, op2_fill):   # positional_two_term
    op1_use = op1_fill if i < j else op2_fill
    op2_use = op2 if i < j else op1
    if j < i and sign:
        scale = -scale
    return (scale, ((min(i, j), op1_use), (max(i, j), op2_use)))


def _t3(scale, pb, p1, p2, boson, boson_fill, op1, op1_fill, op2, op2_fill):   # three_term
    op1_use = op1_fill if p1 < p2 else op2_fill
    op2_use = op2 if p1 < p2 else op1
    boson_use = boson_fill if (p1 < pb < p2 or p2 < pb < p1) else boson
    if p2 < p1:
        scale = -scale
    return (scale, tuple(sorted(((pb, boson_use), (min(p1, p2), op1_use), (max(p1, p2), op2_use)))))


def _t4(max2s, scale, i, j, k, l, op_i, op_k):        # TermMakerSU2::four_term
    if _sgn(i, j, k, l):
        scale = -scale
    st = sorted([(i, op_i), (j, op_i), (k, op_k), (l, op_k)], key=lambda x: x[0])
    if max2s == 2:
        names = [st[0][1].fill_couple_up, st[1][1].couple_up, st[2][1].fill_couple_down, st[3][1].couple_down]
    else:
        names = [st[0][1].fill_couple_up, st[1][1].couple_down, st[2][1].fill_couple_up, st[3][1].couple_down]
    return (scale, tuple((p, n) for (p, _), n in zip(st, names)))


def _ss_two(me, i, k, l, j):                           # SpinSumSU2::two_term
    cr, de, cnt, fl = _OPS["create"], _OPS["destroy"], _OPS["count"], _OPS["flip"]
    if i == j and k == l and j != k:
        return [_t2(me, i, k, "count", "count")]
    if i == k and j == l and j != k:
        return [_t2(2.0 * me, i, j, "e2d", "d2e")]
    if i == l and j == k and i != j:
        return [_pos2(False, math.sqrt(3.0) * me, i, j, fl.couple_down, fl.couple_up, fl.couple_down, fl.couple_up),
                _t2(-0.5 * me, i, j, "count", "count")]
    if (i == k == l) or (k == l == j) or (i == l == j) or (i == k == j):
        if i == k == l:
            s, p = i, j
        elif k == l == j:
            s, p = j, i
        elif i == l == j:
            s, p = i, k
        else:
            s, p = i, l
        if i == k:
            cc = _OPS["create_count"]
            return [_pos2(True, math.sqrt(2.0) * me, s, p, cc.couple_down, cc.fill_couple_up,
                          de.couple_down, de.fill_couple_up)]
        dc = _OPS["destroy_count"]
        return [_pos2(True, -math.sqrt(2.0) * me, s, p, dc.couple_down, dc.fill_couple_up,
                      cr.couple_down, cr.fill_couple_up)]
    raise ValueError("unexpected index arrangement for a two-site term")


def _ss_three(me, i, k, l, j):                         # SpinSumSU2::three_term
    cr, de, cnt, fl = _OPS["create"], _OPS["destroy"], _OPS["count"], _OPS["flip"]
    r2, r3 = math.sqrt(2.0), math.sqrt(3.0)
    if i == j or k == l:                                # three_termA
        if i == j:
            same = i
        if k == l:
            same, k, l = k, i, j
        return [_t3(r2 * me, same, k, l, cnt.no_couple, cnt.fill_no_couple,
                    cr.couple_down, cr.fill_couple_up, de.couple_down, de.fill_couple_up)]
    ret = []                                            # three_termB
    if i == k:
        p1, p2 = min(l, j), max(l, j)
        ret.append(_t3(-r2 * me, i, p1, p2, "e2d", "e2d", de.couple_down, de.fill_couple_up,
                       de.couple_down, de.fill_couple_up))
    if j == l:
        p1, p2 = min(i, k), max(i, k)
        ret.append(_t3(-r2 * me, j, p1, p2, "d2e", "d2e", cr.couple_down, cr.fill_couple_up,
                       cr.couple_down, cr.fill_couple_up))
    if j == k or i == l:
        if j == k:
            same, p1, p2 = j, l, i
        else:
            same, p1, p2 = i, j, k
        phase = -1.0 if _sgn(i, k, l, j) else 1.0
        if same < min(p1, p2):
            ret.append(_t3(phase * r3 * me, same, p1, p2, fl.couple_up, fl.couple_up,
                           cr.couple_down, cr.fill_couple_down, de.couple_down, de.fill_couple_down))
            ret.append(_t3(-0.5 * r2 * me, same, p1, p2, "count", "count",
                           cr.couple_down, cr.fill_couple_up, de.couple_down, de.fill_couple_up))
        elif same > max(p1, p2):
            ret.append(_t3(phase * r3 * me, same, p1, p2, fl.couple_down, fl.couple_down,
                           cr.couple_up, cr.fill_couple_up, de.couple_up, de.fill_couple_up))
            ret.append(_t3(-0.5 * r2 * me, same, p1, p2, "count", "count",
                           cr.couple_down, cr.fill_couple_up, de.couple_down, de.fill_couple_up))
        else:
            ret.append(_t3(phase * r3 * me, same, p1, p2, fl.no_couple, fl.no_couple,
                           cr.couple_down, cr.fill_couple_up, de.couple_down, de.fill_couple_up))
            ret.append(_t3(-0.5 * r2 * me, same, p1, p2, "count_fill", "count_fill",
                           cr.couple_down, cr.fill_couple_up, de.couple_down, de.fill_couple_up))
    return ret


def _ss_four(me, i, k, l, j):                          # SpinSumSU2::four_term
    cr, de = _OPS["create"], _OPS["destroy"]
    i_, j_, k_, l_ = _align(i, j, k, l)
    r3 = math.sqrt(3.0)
    if k_ > l_ > j_:
        return [_t4(2, -r3 * me, i, k, l, j, cr, de), _t4(1, me, i, k, l, j, cr, de)]
    if k_ > j_ > l_:
        loc = -me if _sgn(i, k, l, j) else me
        return [_t4(2, r3 * loc, i, k, l, j, cr, de), _t4(1, loc, i, k, l, j, cr, de)]
    if j_ > k_ > l_:
        return [_t4(1, 2.0 * me, i, k, l, j, cr, de)]
    raise ValueError("unexpected index arrangement for a four-site term")


def _ss_v(me, i, k, l, j):                             # SpinSumSU2::V_term
    n = len({i, k, l, j})
    if n == 4:
        return _ss_four(me, i, k, l, j)
    if n == 3:
        return _ss_three(me, i, k, l, j)
    if n == 2:
        return _ss_two(me, i, k, l, j)
    return [(2.0 * me, ((i, "docc"),))]


def integral_elements(h, eri, ecore=0.0, tol=1e-14):
    """(value, i, j, k, l) of the unique integrals with 0-based aligned indices, -1 where absent
    (parse_integrals.h:136-152 applied to the lines qc.write_fcidump writes)."""
    L = h.shape[0]
    out = []
    for i, j, k, l in itertools.product(range(L), repeat=4):
        if i < j or k < l or (i * L + j) < (k * L + l):
            continue
        if abs(eri[i, j, k, l]) > tol:
            out.append((float(eri[i, j, k, l]),) + _align(i, j, k, l))
    for i, j in itertools.product(range(L), repeat=2):
        if i >= j and abs(h[i, j]) > tol:
            out.append((float(h[i, j]),) + _align(i, j, -1, -1))
    out.append((float(ecore), -1, -1, -1, -1))
    return out


def qc_su2_terms(elements, L, N, sources=None):
    """Terms of qc_su2::create_terms (model.hpp:321-580): ({((pos, op), ...): coeff} with
    positions ascending, core energy).  Terms with identical operators and positions are summed
    (chem_helper.h:69-100).  sources: optional dict filled with key -> [(element index, factor)],
    coefficient = sum of factor * element value (every term is linear in one element)."""
    terms = {}
    ecore = 0.0
    cur = {"m": 0, "v": 0.0}

    def add(lst):
        # the builders run with a unit matrix element; the element scales the factor here
        for f, ops in lst:
            key = tuple(sorted(ops))
            terms[key] = terms.get(key, 0.0) + f * cur["v"]
            if sources is not None:
                sources.setdefault(key, []).append((cur["m"], f))
    cr, de = _OPS["create"], _OPS["destroy"]
    for m_, (v, i, j, k, l) in enumerate(elements):
        cur["m"], cur["v"] = m_, v
        v = 1.0
        if i == j == k == l == -1:
            ecore += cur["v"]
        elif i == j and k == -1 and l == -1:
            add([(v, ((i, "count"),))])
        elif k == -1 and l == -1:
            if N == 1:
                add([_pos2(True, math.sqrt(2.0) * v, j, i, "create", "create_fill", "destroy", "destroy_fill"),
                     _pos2(True, math.sqrt(2.0) * v, i, j, "create", "create_fill", "destroy", "destroy_fill")])
            else:
                f = v / (N - 1)
                for kk in range(L):
                    if kk == j or kk == i:
                        continue
                    add(_ss_three(f, i, kk, kk, j))
                    add(_ss_three(f, j, kk, kk, i))
                add(_ss_v(f, i, i, i, j) + _ss_v(f, j, i, i, i) + _ss_v(f, i, j, j, j) + _ss_v(f, j, j, j, i))
        elif i == j == k == l:
            add([(v, ((i, "docc"),))])
        elif (i == j == k and k != l) or (i != j and j == k == l):
            s, p = (i, l) if i == j else (l, i)
            add(_ss_two(v, s, s, s, p) + _ss_two(v, s, p, s, s))
        elif i == j and k == l and j != k:
            add(_ss_two(v, i, k, k, i))
        elif i == k and j == l and i != j:
            add(_ss_two(0.5 * v, i, i, j, j) + _ss_two(0.5 * v, j, j, i, i) + _ss_two(v, i, j, i, j))
        elif (i == j and j != k and k != l) or (k == l and i != j and j != k):
            if i == j:
                add(_ss_three(v, i, k, l, i) + _ss_three(v, i, l, k, i))
            else:
                add(_ss_three(v, i, k, k, j) + _ss_three(v, j, k, k, i))
        elif ((i == k and j != l) or j == k or (j == l and i != k)) and (i != j and k != l):
            add(_ss_three(v, i, k, l, j) + _ss_three(v, i, l, k, j) + _ss_three(v, j, k, l, i)
                + _ss_three(v, j, l, k, i))
        elif i != j and j != k and k != l and i != k and j != l:
            add(_ss_four(v, i, k, l, j) + _ss_four(v, i, l, k, j) + _ss_four(v, j, k, l, i)
                + _ss_four(v, j, l, k, i))
    return terms, ecore


def qc_su2_mpo(h, eri, N, ecore=0.0, tol=1e-14, sources=False):
    """MPO of the spin-adapted ab initio Hamiltonian in the N-electron sector.
    Returns (site basis, MPO tensors with bond spins, bond dimensions).

    sources=True: also returns the integral elements (integral_elements) and, per site j,
    {(b1, b2, op_id): [(element index, factor), ...]}: the elements summed into each switch
    coefficient; the tensors keep zero coefficients and carry op_ids, so that derivatives with
    respect to every MPO coefficient map back to the elements (mpo_gradient_to_elements)."""
    from . import qc
    L = h.shape[0]
    site = su2_site()
    names = list(site.ops)
    oid = {n: k for k, n in enumerate(names)}
    ops = [site.ops[n] for n in names]
    elements = integral_elements(h, eri, ecore, tol)
    src = {} if sources else None
    terms, ec = qc_su2_terms(elements, L, N, sources=src)
    site_terms, key_src = {}, {}
    for key, c in terms.items():
        if abs(c) < tol and not sources:
            continue
        k2 = tuple((p, oid[n]) for p, n in key)
        site_terms[k2] = site_terms.get(k2, 0.0) + c
        if sources:
            key_src.setdefault(k2, []).extend(src[key])
    if abs(ec) > 0 or sources:
        k2 = ((0, oid["ident"]),)
        site_terms[k2] = site_terms.get(k2, 0.0) + ec
        if sources:
            key_src.setdefault(k2, []).append((len(elements) - 1, 1.0))
    parity = [int(o.fermionic) for o in ops]
    action = [o.action for o in ops]
    i_id, i_full, i_fill = oid["ident"], oid["ident_full"], oid["fill"]

    def pass_op(par, spin2):
        if par % 2:
            return i_fill
        return i_full if spin2 > 1 else i_id
    auto = qc.operator_automaton(site_terms, L, parity, pass_op, action=action,
                                 sources=key_src if sources else None)
    mpo = auto.tensors(ops, keep_zero=sources)
    if sources:
        return site, mpo, [len(b) for b in auto.bonds], elements, auto.switch_src
    return site, mpo, [len(b) for b in auto.bonds]


def mpo_gradient_to_elements(mpo, switch_src, term_grads, n_elements):
    """dE/d(element value) for every integral element from the derivatives with respect to the MPO
    term scales of every site (term_grads[j] in mpo[j].sorted_terms() order)."""
    g = np.zeros(n_elements)
    for j, (W, tg) in enumerate(zip(mpo, term_grads)):
        b1s, b2s, ois, _ = W.sorted_terms()
        assert len(tg) == len(b1s)
        for t in range(len(b1s)):
            for m, f in switch_src[j].get((int(b1s[t]), int(b2s[t]), W.op_ids[int(ois[t])]), ()):
                g[m] += f * tg[t]
    return g


def element_gradient_to_tensors(elements, g, L):
    """Per-entry derivatives dE/dh[p, q] and dE/d(pq|rs) (each tensor entry a parameter of its
    own, as in qc.fci_energy_gradient_fd) from dE/d(element): an element stands for its whole
    permutation class, whose entries share the derivative equally.  Returns (dh, deri, decore)."""
    dh = np.zeros((L, L))
    deri = np.zeros((L, L, L, L))
    decore = 0.0
    for (v, i, j, k, l), gm in zip(elements, g):
        if i == j == k == l == -1:
            decore += gm
        elif k == -1:
            cls = {(i, j), (j, i)}
            for idx in cls:
                dh[idx] += gm / len(cls)
        else:
            cls = {(i, j, k, l), (j, i, k, l), (i, j, l, k), (j, i, l, k), (k, l, i, j), (l, k, i, j), (k, l, j, i),
                   (l, k, j, i)}
            for idx in cls:
                deri[idx] += gm / len(cls)
    return dh, deri, decore


# ---------------------------------------------------------------------------- two-site MPO
def _full_scalar(op, phys):
    """The spin-0 operator with the (|11>, |1-1>) coupling blocks set to 1 where it has none
    (the identity expansion of site_operator_algorithms.h:235-256)."""
    cb, cc = phys.charges[1], phys.charges[2]
    if (cb, cc) in op.blocks:
        return op
    out = Op(spin=op.spin, name=op.name + "+")
    for k, ents in op.blocks.items():
        out.blocks[k] = list(ents)
    out.add(cb, cc, 0, 0, 1.0, abs(cb[1]), abs(cc[1]))
    out.add(cc, cb, 0, 0, 1.0, abs(cc[1]), abs(cb[1]))
    return out


def op_kron_su2(o1, o2, k, phys, phys12, pb):
    """Two-site operator of rank k from site operators o1 (x) o2 with the per-entry 9j couplings
    (site_operator_algorithms.h:276-310); the bond-spin 6j factor is applied by the caller."""
    a, b = o1, o2
    if a.spin > 0 and b.spin == 0:
        b = _full_scalar(b, phys)
    if a.spin == 0 and b.spin > 0:
        a = _full_scalar(a, phys)
    k1, k2 = a.spin, b.spin
    res = Op(spin=k, name="%s*%s/%d" % (o1.name, o2.name, k))
    for (ia, oa), ea in a.blocks.items():
        for (ib, ob), eb in b.blocks.items():
            j1, j2, J = abs(ia[1]), abs(ib[1]), product_spin(ia, ib)
            j1p, j2p, Jp = abs(oa[1]), abs(ob[1]), product_spin(oa, ob)
            cpl = mod_coupling(j1, j2, J, k1, k2, k, j1p, j2p, Jp)
            if cpl == 0.0:
                continue
            for r1, c1, v1, _, _ in ea:
                for r2, c2, v2, _, _ in eb:
                    v = v1 * v2 * cpl
                    if v != 0.0:
                        res.add(fuse(ia, ib), fuse(oa, ob), pb(ia, ib), pb(oa, ob), v, J, Jp)
    return res


def twosite_su2(W1, W2, phys):
    """make_twosite_mpo for SU2 (ts_ops.hpp:38-168): W12(b1,b3) = sum_{b2, o1, o2, k} 6j-factor
    (x) op_kron_k(o1, o2), one operator per (o1, o2, k) with its coefficient matrix.

    The 6j factor of a term depends on the spins j (b1), jm (b2) and jp (b3) of the three bond
    states; per middle spin jm the product over b2 is one sparse product C1 P_jm C2 (P_jm: the
    projector on the b2 of spin jm), whose entry (b1, b3) is then scaled by the factor of
    (spin of b1, jm, spin of b3) and summed over jm in ascending order."""
    phys12, pb = product_index(phys, phys)
    ls = np.asarray(W1.left_spin)
    ms = np.asarray(W1.right_spin)
    rs = np.asarray(W2.right_spin)
    spins_l, spins_m, spins_r = sorted(set(ls.tolist())), sorted(set(ms.tolist())), sorted(set(rs.tolist()))
    li = np.searchsorted(np.asarray(spins_l), ls)   # position of each b1's spin in spins_l
    ri = np.searchsorted(np.asarray(spins_r), rs)
    Pm = {v: sp.diags((ms == v).astype(np.float64)) for v in spins_m}
    C1s = [{jm: (C @ Pm[jm]).tocsr() for jm in spins_m} for C in W1.coeffs]
    C2s = [{jm: (Pm[jm] @ C).tocsr() for jm in spins_m} for C in W2.coeffs]
    ftab = {}

    def factors(k1, k2, k, jm):
        key = (k1, k2, k, jm)
        F = ftab.get(key)
        if F is None:
            F = np.zeros((len(spins_l), len(spins_r)))
            for a, j in enumerate(spins_l):
                for c, jp in enumerate(spins_r):
                    f = math.sqrt((jm + 1.0) * (k + 1.0)) * sixj(j, jp, k, k2, k1, jm)
                    F[a, c] = -f if ((j + jp + k1 + k2) // 2) % 2 else f
            ftab[key] = F
        return F

    ops, coeffs = [], []
    for o1, c1 in zip(W1.ops, C1s):
        for o2, c2 in zip(W2.ops, C2s):
            k1, k2 = o1.spin, o2.spin
            prods = []
            for jm in spins_m:
                A, Bm = c1[jm], c2[jm]
                if A.nnz == 0 or Bm.nnz == 0:
                    continue
                X = (A @ Bm).tocoo()
                if X.nnz:
                    prods.append((jm, X))
            if not prods:
                continue
            for k in range(abs(k1 - k2), k1 + k2 + 1, 2):
                M = None
                for jm, X in prods:
                    f = factors(k1, k2, k, jm)[li[X.row], ri[X.col]]
                    P = sp.csr_matrix((X.data * f, (X.row, X.col)), shape=X.shape)
                    M = P if M is None else M + P
                M = M.tocsr()
                M.eliminate_zeros()
                if M.nnz == 0:
                    continue
                op = op_kron_su2(o1, o2, k, phys, phys12, pb)
                if not op.blocks:
                    continue
                ops.append(op)
                coeffs.append(M)
    T = SparseMPOTensor(W1.row_dim, W2.col_dim, ops, coeffs)
    T.left_spin, T.left_herm, T.left_phase = list(W1.left_spin), list(W1.left_herm), list(W1.left_phase)
    T.right_spin, T.right_herm, T.right_phase = list(W2.right_spin), list(W2.right_herm), list(W2.right_phase)
    return T, phys12, pb


# ---------------------------------------------------------------------------- two-site reduction
def _reduction_segments(left, phys, right, reduce):
    """Segments (out_offset, in_offset, length, coefficient) of reduce_right (reduce=True) or
    unreduce_right (ts_reduction.h:35-128, 232-333) on the right-paired padded layout of a
    two-site tensor with site basis phys on both sites.  Each (left sector, two-site physical
    sub-index, right sector) slab is a contiguous run of ls*rs elements."""
    phys12, pb = product_index(phys, phys)
    lay = SiteLayout(left, phys12, right)
    segs = []
    for lb, (lc, lsz) in enumerate(zip(left.charges, left.sizes)):
        base = lay.rp_offset[lb]
        for pc in phys12.charges:
            rc = fuse(lc, pc)
            if not right.has(rc):
                continue
            rsz = right.size_of(rc)
            col0 = lay.rpb(pc, rc)
            for c1 in phys.charges:
                for c2 in phys.charges:
                    if fuse(c1, c2) != pc:
                        continue
                    ip = pb(c1, c2)
                    S1, S2 = abs(c1[1]), abs(c2[1])
                    jl, jr = lc[1], rc[1]
                    if (jl == jr) and jl > 0 and S1 == 1 and S2 == 1:
                        if reduce:
                            jm = jl + c1[1]
                            b0 = ip if c1[1] == 1 else ip - 1
                            outs = [(b0 + j // 2, j, jm) for j in range(abs(S1 - S2), S1 + S2 + 1, 2)]
                        else:
                            j = 0 if c1[1] == 1 else 2
                            b0 = ip if j == 0 else ip - 1
                            outs = [((b0 + 1) if jm == jl - 1 else b0, j, jm) for jm in (jl - 1, jl + 1)]
                    else:
                        j = abs(c1[1] + c2[1])
                        jm = jl + c1[1]
                        if jm < 0:
                            continue
                        outs = [(ip, j, jm)]
                    for op_, j, jm in outs:
                        cf = math.sqrt((j + 1.0) * (jm + 1.0)) * sixj(jl, jr, j, S2, S1, jm)
                        if ((jl + jr + S1 + S2) // 2) % 2:
                            cf = -cf
                        if cf == 0.0:
                            continue
                        segs.append((base + (col0 + op_ * rsz) * lsz, base + (col0 + ip * rsz) * lsz, lsz * rsz, cf))
    return segs


def reduction_segments(left, phys, right):
    return _reduction_segments(left, phys, right, True)


def unreduction_segments(left, phys, right):
    return _reduction_segments(left, phys, right, False)


def apply_segments_np(segs, x, n_out):
    """out = sum over segments of coefficient * x[in : in + length] placed at out (host)."""
    out = np.zeros(n_out)
    for o, i, n, c in segs:
        out[o:o + n] += c * x[i:i + n]
    return out


# ---------------------------------------------------------------------------- exact checks
def fci_spin_energy(h, eri, N, two_s=0, ecore=0.0, nroots=1):
    """Lowest eigenvalues of H among states of total spin S = two_s / 2 with N electrons, from
    the explicit determinant matrices of H and S^2 in the Sz = S sector (small L only)."""
    from . import qc
    L = h.shape[0]
    n_up = (N + two_s) // 2
    n_dn = N - n_up
    dets = []
    for bits in range(1 << (2 * L)):
        nu = sum((bits >> (2 * i)) & 1 for i in range(L))
        nd = sum((bits >> (2 * i + 1)) & 1 for i in range(L))
        if nu == n_up and nd == n_dn:
            dets.append(bits)
    index = {b: k for k, b in enumerate(dets)}

    def matrix(terms):
        M = np.zeros((len(dets), len(dets)))
        for coef, ops in terms:
            modes = [(2 * o + s, dg) for (o, s, dg) in ops]
            for kk, b in enumerate(dets):
                sign, bits = 1, b
                ok = True
                for (mode, dg) in reversed(modes):
                    occ = (bits >> mode) & 1
                    if (dg and occ) or (not dg and not occ):
                        ok = False
                        break
                    sign *= (-1) ** bin(bits & ((1 << mode) - 1)).count("1")
                    bits ^= (1 << mode)
                if ok:
                    M[index[bits], kk] += coef * sign
        return M
    H = matrix(qc.qc_terms(h, eri))
    # S^2 = S- S+ + Sz^2 + Sz in the Sz sector
    s2_terms = []
    for i in range(L):
        for j in range(L):
            s2_terms.append((1.0, ((i, 1, 1), (i, 0, 0), (j, 0, 1), (j, 1, 0))))   # S-_i S+_j
    sz = 0.5 * (n_up - n_dn)
    S2 = matrix(s2_terms) + (sz * sz + sz) * np.eye(len(dets))
    w, V = np.linalg.eigh(S2)
    target = 0.25 * two_s * (two_s + 2)
    P = V[:, np.abs(w - target) < 1e-8]
    e = np.linalg.eigvalsh(P.T @ H @ P)
    return e[:nroots] + ecore if nroots > 1 else float(e[0]) + ecore
