"""ORACLE -- test infrastructure only (imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg).  ctypes wrapper of oracle/_build/libmqoracle.so, the CPU restatement of the
reference's sigma / boundary / Jacobi-Davidson path (see oracle/src/oracle.cpp for the per-function
reference citations), plus the numpy dense brute force that pins it on tiny systems."""
import ctypes as C
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# MQ_ORACLE_LIB: another build of the oracle (the sanitizer test's instrumented build)
LIB = os.environ.get("MQ_ORACLE_LIB") or os.path.join(HERE, "_build", "libmqoracle.so")
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "maquis-dmrg_amd"))

from mqdmrg import capi  # noqa: E402  (descriptor structs only)

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P = C.c_void_p
        L.mqo_last_error.restype = C.c_char_p
        L.mqo_wigner9j.restype = C.c_double
        L.mqo_wigner9j.argtypes = [C.c_int] * 9
        L.mqo_set_coupling.argtypes = [C.c_int] * 9 + [C.c_double, P]
        L.mqo_schedule_sigma.argtypes = [P, P, P, P, C.POINTER(C.c_void_p)]
        L.mqo_schedule_sigma_blocks.argtypes = [P, P, P, P, P, C.POINTER(C.c_void_p)]
        L.mqo_schedule_export.argtypes = [P, P, C.POINTER(C.c_uint64)]
        L.mqo_schedule_free.argtypes = [P]
        L.mqo_sigma_apply.argtypes = [P, P, P, P, P, P, P, P]
        L.mqo_left_step.argtypes = [P, P, P, P, P, P, C.c_int, C.POINTER(C.c_void_p)]
        L.mqo_right_step.argtypes = [P, P, P, P, P, P, C.c_int, C.POINTER(C.c_void_p)]
        L.mqo_boundary_index.argtypes = [P, P]
        L.mqo_boundary_data.restype = C.POINTER(C.c_double)
        L.mqo_boundary_data.argtypes = [P, C.c_uint32]
        L.mqo_boundary_free.argtypes = [P]
        L.mqo_subspace_min.argtypes = [P, C.c_int, P, P]
        L.mqo_jd.argtypes = [P, P, P, P, P, P, C.c_uint64, P, P, C.c_double, C.c_int, P, P]
        L.mqo_jd_ex.argtypes = [P, P, P, P, P, P, C.c_uint64, P, P, C.c_double, C.c_int, C.c_int, C.c_int, P, P, P]
        L.mqo_set_dgemm.argtypes = [P]
        _lib = L
    return _lib


def _chk(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().mqo_last_error().decode())


def _cohort_ptrs(cohorts):
    arr = (C.POINTER(C.c_double) * max(1, len(cohorts)))()
    for i, a in enumerate(cohorts):
        arr[i] = a.ctypes.data_as(C.POINTER(C.c_double)) if (a is not None and a.size) else None
    return arr


class HostBoundary:
    """A boundary on the host: BIndex + one numpy array per cohort."""

    def __init__(self, index, cohorts):
        self.index = index
        self.cohorts = [np.ascontiguousarray(c, np.float64) for c in cohorts]


def trivial_boundary(index, aux_dim):
    """Boundary(ud, ld, ad) with every element 1 (mp_tensors/boundary.h:258-283)."""
    n = len(index.charges)
    bs = [s * s for s in index.sizes]
    offs = [[k * bs[i] for k in range(aux_dim)] for i in range(n)]
    bi = capi.BIndex(index, index, aux_dim, list(range(n)), list(range(n)), [0] * n, [aux_dim] * n,
                     np.asarray(offs, np.int64).ravel(), np.ones(n * aux_dim), np.zeros(n * aux_dim, np.uint8))
    return HostBoundary(bi, [np.ones(aux_dim * b) for b in bs])


def schedule_sigma(site, mpo, left, right):
    keep = []
    h = C.c_void_p()
    _chk(lib().mqo_schedule_sigma(C.byref(site.desc), C.byref(mpo.desc), C.byref(left.index.desc(keep)),
                                  C.byref(right.index.desc(keep)), C.byref(h)))
    try:
        d = capi.mq_schedule_desc()
        f = C.c_uint64()
        _chk(lib().mqo_schedule_export(h, C.byref(d), C.byref(f)))
        return capi.schedule_to_dict(d), f.value
    finally:
        lib().mqo_schedule_free(h)


class SigmaOracle:
    """create_contraction_schedule + super_hamil_mv (CPU branch) on host data."""

    def __init__(self, site, mpo, left, right, only_blocks=None):
        """only_blocks: optional iterable of ket block (left sector) positions; sigma is then the
        contribution of those MPS blocks alone (a bounded sample for timing)."""
        self.site, self.left, self.right = site, left, right
        self.keep = []
        self.ld = left.index.desc(self.keep)
        self.rd = right.index.desc(self.keep)
        self.lp = _cohort_ptrs(left.cohorts)
        self.rp = _cohort_ptrs(right.cohorts)
        self.h = C.c_void_p()
        only = None
        if only_blocks is not None:
            mask = np.zeros(site.desc.left_i.n, np.uint8)
            mask[list(only_blocks)] = 1
            self.keep.append(mask)
            only = mask.ctypes.data
        _chk(lib().mqo_schedule_sigma_blocks(C.byref(site.desc), C.byref(mpo.desc), C.byref(self.ld),
                                             C.byref(self.rd), only, C.byref(self.h)))
        d = capi.mq_schedule_desc()
        f = C.c_uint64()
        _chk(lib().mqo_schedule_export(self.h, C.byref(d), C.byref(f)))
        self.flops = f.value

    def apply(self, psi):
        psi = np.ascontiguousarray(psi, np.float64)
        out = np.zeros_like(psi)
        _chk(lib().mqo_sigma_apply(self.h, C.byref(self.site.desc), C.byref(self.ld), self.lp, C.byref(self.rd),
                                   self.rp, psi.ctypes.data, out.ctypes.data))
        return out

    def jd(self, init, tol=1e-8, max_iter=8, gmres_iter=0, ortho=()):
        """Jacobi-Davidson lowest eigenpair (ietl/jacobi.h:386-474); gmres_iter > 0: GMRES correction
        (ietl/gmres.h), ortho: vectors the search space is kept orthogonal to (excited states)."""
        init = np.ascontiguousarray(init, np.float64)
        out = np.zeros_like(init)
        e = C.c_double()
        it = C.c_int()
        keep = [np.ascontiguousarray(o, np.float64) for o in ortho]
        ptrs = (C.c_void_p * max(1, len(keep)))(*[k.ctypes.data for k in keep])
        _chk(lib().mqo_jd_ex(self.h, C.byref(self.site.desc), C.byref(self.ld), self.lp, C.byref(self.rd), self.rp,
                             init.size, init.ctypes.data, out.ctypes.data, tol, max_iter, int(gmres_iter), len(keep),
                             ptrs, C.byref(e), C.byref(it)))
        return e.value, out, it.value

    def __del__(self):
        try:
            lib().mqo_schedule_free(self.h)
        except Exception:
            pass


class _TermMpo:
    """The MPO descriptor of `mpo` restricted to its term t (scale 1): H_eff of this MPO is the
    derivative of H_eff with respect to term_scale[t]."""

    def __init__(self, mpo, t):
        d = capi.mq_mpo_desc.from_buffer_copy(mpo.desc)
        assert 0 <= t < d.n_terms
        self.one = np.ones(1, np.float64)

        def at(ptr, ctype, size):
            return C.cast(C.cast(ptr, C.c_void_p).value + size * t, C.POINTER(ctype))
        d.n_terms = 1
        d.term_b1 = at(d.term_b1, C.c_uint32, 4)
        d.term_b2 = at(d.term_b2, C.c_uint32, 4)
        d.term_op = at(d.term_op, C.c_uint32, 4)
        d.term_scale = C.cast(self.one.ctypes.data, C.POINTER(C.c_double))
        self.parent = mpo
        self.desc = d


def sigma_term_gradient(site, mpo, left, right, psi):
    """grad[t] = psi . (dH_eff / d term_scale[t]) psi for every MPO term, by brute force: the
    oracle sigma of the single-term MPO (scale 1) dotted with psi.  H_eff is linear in each term
    scale (the reference's alpha = coefficient x coupling x scale, op_iterate.hpp:58-63,
    site_hamil_tasks.hpp:120-125), so this is the exact derivative of psi . H_eff psi."""
    psi = np.ascontiguousarray(psi, np.float64)
    n = mpo.desc.n_terms
    g = np.zeros(n)
    for t in range(n):
        g[t] = float(np.dot(psi, SigmaOracle(site, _TermMpo(mpo, t), left, right).apply(psi)))
    return g


def _take_boundary(h):
    d = capi.mq_bindex_desc()
    _chk(lib().mqo_boundary_index(h, C.byref(d)))
    bi = capi.BIndex.from_desc(d)
    cohorts = []
    for ci in range(bi.n_cohorts):
        n = bi.cohort_size(ci)
        p = lib().mqo_boundary_data(h, ci)
        cohorts.append(np.ctypeslib.as_array(p, shape=(n,)).copy() if n else np.zeros(0))
    lib().mqo_boundary_free(h)
    return HostBoundary(bi, cohorts)


def left_step(site, mpo, bra_lp, ket_rp, left, symmetric=True):
    keep = []
    ld = left.index.desc(keep)
    h = C.c_void_p()
    bra_lp = np.ascontiguousarray(bra_lp, np.float64)
    ket_rp = np.ascontiguousarray(ket_rp, np.float64)
    _chk(lib().mqo_left_step(C.byref(site.desc), C.byref(mpo.desc), bra_lp.ctypes.data, ket_rp.ctypes.data,
                             C.byref(ld), _cohort_ptrs(left.cohorts), int(symmetric), C.byref(h)))
    return _take_boundary(h)


def right_step(site, mpo, bra_rp, ket_rp, right, symmetric=True):
    keep = []
    rd = right.index.desc(keep)
    h = C.c_void_p()
    bra_rp = np.ascontiguousarray(bra_rp, np.float64)
    ket_rp = np.ascontiguousarray(ket_rp, np.float64)
    _chk(lib().mqo_right_step(C.byref(site.desc), C.byref(mpo.desc), bra_rp.ctypes.data, ket_rp.ctypes.data,
                              C.byref(rd), _cohort_ptrs(right.cohorts), int(symmetric), C.byref(h)))
    return _take_boundary(h)


def wigner9j(*args):
    return lib().mqo_wigner9j(*[int(a) for a in args])


def set_coupling(a, b, c, d, e, f, g, h, i, init=1.0):
    out = np.zeros(4)
    lib().mqo_set_coupling(a, b, c, d, e, f, g, h, i, float(init), out.ctypes.data)
    return out


def subspace_min(A):
    A = np.ascontiguousarray(A, np.float64)
    n = A.shape[0]
    v = np.zeros(n)
    val = C.c_double()
    _chk(lib().mqo_subspace_min(A.ctypes.data, n, v.ctypes.data, C.byref(val)))
    return val.value, v


def use_numpy_blas():
    """Route the oracle's dgemm through numpy's bundled OpenBLAS (ILP64) when it is found."""
    import glob
    import numpy
    libdir = os.path.join(os.path.dirname(os.path.dirname(numpy.__file__)), "numpy.libs")
    for path in glob.glob(os.path.join(libdir, "libscipy_openblas64_*.so")):
        blas = C.CDLL(path)
        for name in ("scipy_dgemm_64_", "dgemm_64_"):
            fn = getattr(blas, name, None)
            if fn is not None:
                for setter in ("scipy_openblas_set_num_threads64_", "openblas_set_num_threads64_"):
                    st = getattr(blas, setter, None)
                    if st is not None:
                        st(1)
                        break
                lib().mqo_set_dgemm(C.cast(fn, C.c_void_p))
                return path + ":" + name
    return None


# ---------------------------------------------------------------------------- dense brute force
def dense_boundary(hb, aux_dim=None):
    """Resolve a host boundary into dense matrices D[b][bra_full, ket_full], following the
    transposed / conjugate-scaled references of complement_transpose (mp_tensors/boundary.h:161-191)."""
    bi = hb.index
    nb_, nk_ = bi.bra.total(), bi.ket.total()
    offb = np.cumsum([0] + bi.bra.sizes)
    offk = np.cumsum([0] + bi.ket.sizes)
    D = np.zeros((bi.aux_dim, nb_, nk_))
    for ci in range(bi.n_cohorts):
        lb, rb = bi.bra_pos[ci], bi.ket_pos[ci]
        ls, rs = bi.bra.sizes[lb], bi.ket.sizes[rb]
        owner = ci
        if bi.tr[ci]:
            owner = bi.cohort_index(bi.ket.charges[rb], bi.bra.charges[lb])
        for b in range(bi.aux_dim):
            o = bi.offsets[ci, b]
            if o < 0:
                continue
            if bi.tr[ci] or bi.transposes[ci, b]:
                data = hb.cohorts[owner][o:o + ls * rs]
                blk = data.reshape(ls, rs)          # stored rs x ls column-major -> transpose
            else:
                data = hb.cohorts[ci][o:o + ls * rs]
                blk = data.reshape(rs, ls).T
            D[b, offb[lb]:offb[lb] + ls, offk[rb]:offk[rb] + rs] = bi.conj_scales[ci, b] * blk
    return D


def dense_site(layout, vec):
    """Right-paired padded vector -> dense array X[left_full, phys_full, right_full]."""
    L_, P_, R_ = layout.left, layout.phys, layout.right
    ol, op, orr = (np.cumsum([0] + I.sizes) for I in (L_, P_, R_))
    X = np.zeros((L_.total(), P_.total(), R_.total()))
    for lb, (lc, ls) in enumerate(zip(L_.charges, L_.sizes)):
        cols = layout.rp_cols[lb]
        blk = vec[layout.rp_offset[lb]:layout.rp_offset[lb] + ls * cols].reshape(cols, ls).T
        for s, (pc, d) in enumerate(zip(P_.charges, P_.sizes)):
            rc = (lc[0] + pc[0], lc[1] + pc[1])
            rb = R_.position(rc)
            if rb == len(R_):
                continue
            rs = R_.sizes[rb]
            c0 = layout.rpb(pc, rc)
            for ss in range(d):
                X[ol[lb]:ol[lb] + ls, op[s] + ss, orr[rb]:orr[rb] + rs] = blk[:, c0 + ss * rs:c0 + (ss + 1) * rs]
    return X


def dense_to_vec(layout, X):
    L_, P_, R_ = layout.left, layout.phys, layout.right
    ol, op, orr = (np.cumsum([0] + I.sizes) for I in (L_, P_, R_))
    vec = np.zeros(layout.rp_total)
    for lb, (lc, ls) in enumerate(zip(L_.charges, L_.sizes)):
        cols = layout.rp_cols[lb]
        blk = np.zeros((ls, cols))
        for s, (pc, d) in enumerate(zip(P_.charges, P_.sizes)):
            rc = (lc[0] + pc[0], lc[1] + pc[1])
            rb = R_.position(rc)
            if rb == len(R_):
                continue
            rs = R_.sizes[rb]
            c0 = layout.rpb(pc, rc)
            for ss in range(d):
                blk[:, c0 + ss * rs:c0 + (ss + 1) * rs] = X[ol[lb]:ol[lb] + ls, op[s] + ss, orr[rb]:orr[rb] + rs]
        vec[layout.rp_offset[lb]:layout.rp_offset[lb] + ls * cols] = blk.T.ravel()
    return vec


def dense_sigma(Ld, Wd, Rd, X):
    """sigma[a', s', c'] = sum L[b1, a', a] W[b1, b2, s', s] X[a, s, c] R[b2, c, c']."""
    T = np.einsum("xpa,asc->xpsc", Ld, X)
    T = np.einsum("xyqs,xpsc->ypqc", Wd, T)
    return np.einsum("ypqc,ycd->pqd", T, Rd)


def dense_left_step(Ld, Wd, bra, ket):
    """L'[b2, c', c] = sum bra[a', s', c'] L[b1, a', a] W[b1, b2, s', s] ket[a, s, c]."""
    T = np.einsum("xpa,asc->xpsc", Ld, ket)
    T = np.einsum("xyqs,xpsc->ypqc", Wd, T)
    return np.einsum("pqd,ypqc->ydc", bra, T)


def dense_right_step(Rd, Wd, bra, ket):
    """R'[b1, a, a'] = sum ket[a, s, c] W[b1, b2, s', s] R[b2, c, c'] bra[a', s', c']."""
    T = np.einsum("asc,ycd->yasd", ket, Rd)
    T = np.einsum("xyqs,yasd->xaqd", Wd, T)
    return np.einsum("xaqd,pqd->xap", T, bra)
