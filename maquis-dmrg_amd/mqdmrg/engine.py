"""Python handles over the engine's C-ABI: context, device boundaries, sigma plans, boundary
steps and the Jacobi-Davidson solver.  Device memory of site tensors is held in torch tensors on
the engine's own stream (torch is plumbing here: allocation and small dense algebra)."""
import ctypes as C
import weakref

import numpy as np
import torch

from . import capi
from .capi import lib, check
from .models import Op
from .symmetry import Index


class Engine:
    # open engines in creation order (weak references): torch's current stream is the newest open
    # engine's
    _open = []

    def __init__(self, device=0):
        self.device = device
        self.ctx = C.c_void_p()
        check(lib().mq_ctx_create(device, C.byref(self.ctx)))
        self.stream = torch.cuda.ExternalStream(lib().mq_ctx_stream(self.ctx), device="cuda:%d" % device)
        self.torch_device = torch.device("cuda:%d" % device)
        # every torch op of the caller is ordered with the engine's launches on one stream
        torch.cuda.set_stream(self.stream)
        Engine._open.append(weakref.ref(self))
        # boundaries and plans made on this context: freed before the context itself
        self._children = weakref.WeakSet()

    def adopt(self, obj):
        self._children.add(obj)
        return obj

    def sync(self):
        check(lib().mq_ctx_synchronize(self.ctx))

    def close(self):
        if self.ctx:
            for obj in list(self._children):
                obj.close()
            lib().mq_ctx_destroy(self.ctx)
            self.ctx = C.c_void_p()
            # torch must not keep issuing work to the destroyed stream: hand the current stream
            # back to the newest engine still open on the device, else to the default stream
            Engine._open = [r for r in Engine._open if r() is not None and r() is not self]
            if torch.cuda.current_stream(self.torch_device) == self.stream:
                rest = [r() for r in Engine._open if r() is not None and r().ctx and r().device == self.device]
                torch.cuda.set_stream(rest[-1].stream if rest else torch.cuda.default_stream(self.torch_device))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # vectors ------------------------------------------------------------------
    def dot(self, x, y):
        r = C.c_double()
        n = x.numel()
        check(lib().mq_vec_dot(self.ctx, n, dev_ptr(self, x, n), dev_ptr(self, y, n), C.byref(r)))
        return r.value

    def nrm2(self, x):
        r = C.c_double()
        check(lib().mq_vec_nrm2(self.ctx, x.numel(), dev_ptr(self, x, x.numel()), C.byref(r)))
        return r.value

    def multi_dot(self, V, k, ldv, x):
        """[V_j . x for j < k], V_j = V[j * ldv : j * ldv + n] (n = x.numel()), one batched reduction."""
        n = x.numel()
        out = np.zeros(k, np.float64)
        need = (k - 1) * ldv + n
        check(lib().mq_vec_multi_dot(self.ctx, n, int(k), dev_ptr(self, V, need), int(ldv), dev_ptr(self, x, n),
                                     out.ctypes.data_as(C.c_void_p)))
        return out

    def axpy(self, a, x, y):
        n = x.numel()
        check(lib().mq_vec_axpy(self.ctx, n, float(a), dev_ptr(self, x, n), dev_ptr(self, y, n)))

    def scal(self, a, x):
        check(lib().mq_vec_scal(self.ctx, x.numel(), float(a), dev_ptr(self, x, x.numel())))


def dev_ptr(engine, t, n):
    """Device address of a tensor the engine reads or writes n FP64 elements of.  The C-ABI takes
    raw pointers and cannot tell a host pointer, another device or a narrower dtype from a valid
    buffer, so every wrapper checks here, before any C call: an FP64, contiguous tensor on the
    engine's device with at least n elements."""
    if not isinstance(t, torch.Tensor):
        raise TypeError("expected a torch tensor, got %s" % type(t).__name__)
    if t.dtype != torch.float64:
        raise ValueError("engine buffers are float64, got %s" % t.dtype)
    if not t.is_cuda or t.device != engine.torch_device:
        raise ValueError("tensor on %s, the engine runs on %s" % (t.device, engine.torch_device))
    if not t.is_contiguous():
        raise ValueError("engine buffers must be contiguous")
    if t.numel() < n:
        raise ValueError("buffer of %d elements, %d needed" % (t.numel(), n))
    return t.data_ptr()


# ---------------------------------------------------------------------------- descriptors
class MpoDesc:
    """mq_mpo_desc of an MPOTensor (arrays kept alive by this object)."""

    def __init__(self, W):
        self.keep = []
        if hasattr(W, "sorted_terms"):
            ops = W.ops
            b1s, b2s, tops, scales = W.sorted_terms()
        else:
            ops, op_id = [], {}
            b1s, b2s, tops, scales = [], [], [], []
            for (b1, b2) in sorted(W.terms):
                for op, sc in W.terms[(b1, b2)]:
                    if id(op) not in op_id:
                        op_id[id(op)] = len(ops)
                        ops.append(op)
                    b1s.append(b1)
                    b2s.append(b2)
                    tops.append(op_id[id(op)])
                    scales.append(sc)
        opd = (capi.mq_op_desc * max(1, len(ops)))()
        for k, op in enumerate(ops):
            keys = list(op.blocks.keys())
            cin = capi.charges_array([k_[0] for k_ in keys])
            cout = capi.charges_array([k_[1] for k_ in keys])
            eb = np.zeros(len(keys) + 1, dtype=np.uint32)
            rows, cols, rsp, csp, coef = [], [], [], [], []
            for i, key in enumerate(keys):
                ents = op.blocks[key]
                eb[i + 1] = eb[i] + len(ents)
                for r, c, v, rs, cs in ents:
                    rows.append(r)
                    cols.append(c)
                    coef.append(v)
                    rsp.append(rs)
                    csp.append(cs)
            arrs = [np.asarray(rows, np.uint32), np.asarray(cols, np.uint32), np.asarray(rsp, np.uint8),
                    np.asarray(csp, np.uint8), np.asarray(coef, np.float64)]
            self.keep += [cin, cout, eb] + arrs
            d = opd[k]
            d.spin = op.spin
            d.n_blocks = len(keys)
            d.in_charge = C.cast(cin, C.POINTER(capi.mq_charge))
            d.out_charge = C.cast(cout, C.POINTER(capi.mq_charge))
            d.entry_begin = capi.ptr(eb, C.c_uint32)
            d.row = capi.ptr(arrs[0], C.c_uint32)
            d.col = capi.ptr(arrs[1], C.c_uint32)
            d.row_spin = capi.ptr(arrs[2], C.c_uint8)
            d.col_spin = capi.ptr(arrs[3], C.c_uint8)
            d.coeff = capi.ptr(arrs[4], C.c_double)
        self.keep.append(opd)
        herm = lambda h: np.asarray([capi.MQ_NONE if x is None else x for x in h], np.uint32)
        a = [np.asarray(W.left_spin, np.int32), np.asarray(W.right_spin, np.int32), herm(W.left_herm),
             np.asarray(W.left_phase, np.int32), herm(W.right_herm), np.asarray(W.right_phase, np.int32),
             np.asarray(b1s, np.uint32), np.asarray(b2s, np.uint32), np.asarray(tops, np.uint32),
             np.asarray(scales, np.float64)]
        self.keep += a
        d = capi.mq_mpo_desc()
        d.row_dim, d.col_dim = W.row_dim, W.col_dim
        d.left_spin = capi.ptr(a[0], C.c_int32)
        d.right_spin = capi.ptr(a[1], C.c_int32)
        d.left_herm = capi.ptr(a[2], C.c_uint32)
        d.left_phase = capi.ptr(a[3], C.c_int32)
        d.right_herm = capi.ptr(a[4], C.c_uint32)
        d.right_phase = capi.ptr(a[5], C.c_int32)
        d.n_ops = len(ops)
        d.ops = C.cast(opd, C.POINTER(capi.mq_op_desc))
        d.n_terms = len(b1s)
        d.term_b1 = capi.ptr(a[6], C.c_uint32)
        d.term_b2 = capi.ptr(a[7], C.c_uint32)
        d.term_op = capi.ptr(a[8], C.c_uint32)
        d.term_scale = capi.ptr(a[9], C.c_double)
        self.desc = d
        self.term_b1, self.term_b2, self.term_op, self.term_scale = a[6], a[7], a[8], a[9]

    def with_scales(self, scales):
        """The same MPO with other term scales (same order): a view sharing the op tables."""
        new = MpoDesc.__new__(MpoDesc)
        sc = np.ascontiguousarray(scales, np.float64)
        assert sc.shape == self.term_scale.shape
        new.keep = self.keep + [sc]
        new.term_b1, new.term_b2, new.term_op, new.term_scale = self.term_b1, self.term_b2, self.term_op, sc
        d = capi.mq_mpo_desc.from_buffer_copy(self.desc)
        d.term_scale = capi.ptr(sc, C.c_double)
        new.desc = d
        return new


class SiteDesc:
    def __init__(self, sym, left, phys, right):
        from .symmetry import SiteLayout
        self.keep = []
        self.desc = capi.site_desc(sym, left, phys, right, self.keep)
        lay = SiteLayout(left, phys, right)
        self.layout = lay
        self.rp_total, self.lp_total = lay.rp_total, lay.lp_total


# ---------------------------------------------------------------------------- boundaries
class DeviceBoundary:
    def __init__(self, engine, handle):
        self.engine = engine
        self.h = handle
        engine.adopt(self)
        d = capi.mq_bindex_desc()
        check(lib().mq_boundary_index(self.h, C.byref(d)))
        self.index = capi.BIndex.from_desc(d)

    @staticmethod
    def trivial(engine, index, aux_dim):
        keep = []
        d = capi.index_desc(index, keep)
        h = C.c_void_p()
        check(lib().mq_boundary_create_trivial(engine.ctx, C.byref(d), aux_dim, C.byref(h)))
        return DeviceBoundary(engine, h)

    @staticmethod
    def upload(engine, bindex, cohorts):
        keep = []
        d = bindex.desc(keep)
        arrs = [None if c is None else np.ascontiguousarray(c, dtype=np.float64) for c in cohorts]
        ptrs = (C.POINTER(C.c_double) * max(1, len(arrs)))()
        for i, a in enumerate(arrs):
            ptrs[i] = capi.ptr(a, C.c_double) if a is not None and a.size else None
        h = C.c_void_p()
        check(lib().mq_boundary_create(engine.ctx, C.byref(d), ptrs, C.byref(h)))
        return DeviceBoundary(engine, h)

    def download(self):
        out = []
        for ci in range(self.index.n_cohorts):
            n = self.index.cohort_size(ci)
            a = np.zeros(n, dtype=np.float64)
            if n:
                check(lib().mq_boundary_download(self.h, ci, a.ctypes.data))
            out.append(a)
        return out

    def cohort(self, ci):
        """Device copy of cohort ci (its n_blocks blocks of ls x rs, column-major, back to back) as
        a torch tensor, in stream order with the engine's launches."""
        n = self.index.cohort_size(ci)
        out = torch.empty(max(n, 1), dtype=torch.float64, device=self.engine.torch_device)[:n]
        if n:
            p = C.c_void_p()
            check(lib().mq_boundary_cohort_ptr(self.h, ci, C.byref(p)))
            check(lib().mq_vec_copy(self.engine.ctx, n, p, out.data_ptr()))
        return out

    # residency (mq_boundary_evict / _prefetch / _residency) ---------------------------------
    DEVICE, HOST, DISK = 0, 1, 2

    def evict(self, path=None):
        """Free the device copy (keeping a pinned host copy, or with path a spill file)."""
        check(lib().mq_boundary_evict(self.engine.ctx, self.h, None if path is None else str(path).encode()))

    def prefetch(self):
        """Start bringing an evicted boundary back to the device (asynchronous copy stream)."""
        check(lib().mq_boundary_prefetch(self.engine.ctx, self.h))

    def residency(self):
        """(where: DEVICE / HOST / DISK, device bytes, pinned host bytes)."""
        w, d, hb = C.c_int(), C.c_uint64(), C.c_uint64()
        check(lib().mq_boundary_residency(self.h, C.byref(w), C.byref(d), C.byref(hb)))
        return w.value, d.value, hb.value

    def traces(self):
        a = np.zeros(self.index.aux_dim, dtype=np.float64)
        check(lib().mq_boundary_traces(self.h, a.ctypes.data))
        return a

    def close(self):
        if self.h:
            lib().mq_boundary_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------- schedules
def build_schedule(kind, site, mpo_desc, bindex_l=None, bindex_r=None, symmetric=True):
    """Host-only schedule builders; returns (dict of arrays, flops, out_index or None)."""
    keep = []
    h = C.c_void_p()
    if kind == "sigma":
        check(lib().mq_schedule_build_sigma(C.byref(site.desc), C.byref(mpo_desc.desc),
                                            C.byref(bindex_l.desc(keep)), C.byref(bindex_r.desc(keep)), C.byref(h)))
    elif kind == "left":
        check(lib().mq_schedule_build_left(C.byref(site.desc), C.byref(mpo_desc.desc),
                                           C.byref(bindex_l.desc(keep)), int(symmetric), C.byref(h)))
    else:
        check(lib().mq_schedule_build_right(C.byref(site.desc), C.byref(mpo_desc.desc),
                                            C.byref(bindex_r.desc(keep)), int(symmetric), C.byref(h)))
    try:
        return _schedule_result(kind, h)
    finally:
        lib().mq_schedule_destroy(h)


def step_index(kind, site, mpo_desc, bindex, symmetric=True):
    """Index of the boundary a left / right step produces, without building the step's schedule."""
    keep = []
    h = C.c_void_p()
    fn = lib().mq_schedule_build_left_index if kind == "left" else lib().mq_schedule_build_right_index
    check(fn(C.byref(site.desc), C.byref(mpo_desc.desc), C.byref(bindex.desc(keep)), int(symmetric), C.byref(h)))
    try:
        bd = capi.mq_bindex_desc()
        check(lib().mq_schedule_out_index(h, C.byref(bd)))
        return capi.BIndex.from_desc(bd)
    finally:
        lib().mq_schedule_destroy(h)


def _schedule_result(kind, h):
    d = capi.mq_schedule_desc()
    check(lib().mq_schedule_export(h, C.byref(d)))
    s = capi.schedule_to_dict(d)
    f = C.c_uint64()
    check(lib().mq_schedule_flops(h, C.byref(f)))
    out_idx = None
    if kind != "sigma":
        bd = capi.mq_bindex_desc()
        check(lib().mq_schedule_out_index(h, C.byref(bd)))
        out_idx = capi.BIndex.from_desc(bd)
    return s, f.value, out_idx


# ---------------------------------------------------------------------------- sigma
SIGMA_AUTO, SIGMA_SCHEDULE, SIGMA_GROUPED = 0, 1, 2
SIGMA_ADJOINT = 0x100   # OR-ed into the algorithm: the plan also supports SigmaPlan.adjoint
AUTOTUNE_LIBRARY_ONLY = 0x10   # OR-ed into the precision of mq_plan_autotune
# arithmetic of the sigma contraction (mq_precision)
PREC_F64, PREC_F32, PREC_BF16X3 = 0, 1, 2
PRECISIONS = {"f64": PREC_F64, "f32": PREC_F32, "bf16x3": PREC_BF16X3}


class SigmaPlan:
    def __init__(self, engine, site, mpo_desc, left, right, algorithm=SIGMA_AUTO, adjoint=False):
        """adjoint: keep the MPO term map for SigmaPlan.adjoint (GROUPED plans)."""
        self.engine = engine
        self.h = None
        engine.adopt(self)
        self.left, self.right = left, right   # keep boundaries alive
        self.n_terms = mpo_desc.desc.n_terms
        self.h = C.c_void_p()
        alg = int(algorithm) | (SIGMA_ADJOINT if adjoint else 0)
        check(lib().mq_sigma_plan_create_ex(engine.ctx, C.byref(site.desc), C.byref(mpo_desc.desc), left.h, right.h,
                                            alg, C.byref(self.h)))
        n = C.c_uint64()
        check(lib().mq_plan_vector_size(self.h, C.byref(n)))
        self.n = n.value
        self._flops = None
        b = C.c_uint64()
        check(lib().mq_plan_workspace_bytes(self.h, C.byref(b)))
        self.workspace_bytes = b.value

    @property
    def flops(self):
        """The reference's flop model of one application (the S-stage part of a grouped plan is
        counted on first access: it scans the dense MPO matrices once per sector pair)."""
        if self._flops is None:
            f = C.c_uint64()
            check(lib().mq_plan_flops(self.h, C.byref(f)))
            self._flops = f.value
        return self._flops

    @staticmethod
    def from_schedule(engine, sched, ket_rows, ket_cols, left, right):
        self = SigmaPlan.__new__(SigmaPlan)
        self.engine, self.left, self.right = engine, left, right
        self.h = None
        engine.adopt(self)
        keep = []
        d = capi.schedule_from_dict(sched, keep)
        rows = np.ascontiguousarray(ket_rows, np.uint32)
        cols = np.ascontiguousarray(ket_cols, np.uint32)
        self.h = C.c_void_p()
        check(lib().mq_sigma_plan_create_from_schedule(engine.ctx, C.byref(d), len(rows), capi.ptr(rows, C.c_uint32),
                                                       capi.ptr(cols, C.c_uint32), left.h, right.h, C.byref(self.h)))
        n = C.c_uint64()
        check(lib().mq_plan_vector_size(self.h, C.byref(n)))
        self.n = n.value
        self._flops = None
        self.workspace_bytes = 0
        return self

    def apply(self, psi, sigma, precision=PREC_F64):
        e = self.engine
        check(lib().mq_sigma_apply_ex(self.h, dev_ptr(e, psi, self.n), dev_ptr(e, sigma, self.n), int(precision)))

    def adjoint(self, psi):
        """d(psi . H_eff psi) / d term_scale[t] for every term of the plan's MPO descriptor (numpy,
        FP64): for a normalized eigenvector, dE/d scale_t."""
        g = np.zeros(max(1, self.n_terms), np.float64)
        check(lib().mq_sigma_adjoint(self.h, C.c_void_p(dev_ptr(self.engine, psi, self.n)), g.ctypes.data_as(C.c_void_p)))
        return g[:self.n_terms]

    def autotune(self, precision=PREC_F64, library_only=False):
        """Time the tile configurations of the T and final stages and keep the fastest per row-size
        class (cached per engine by the stage's block-size histogram); in FP32 / BF16x3 also the
        S-stage library GEMMs (cached by shape and by shape class, whose solution later GEMMs of the
        class adopt).  library_only: only the library GEMMs that no cache covers.  Returns the
        number of timings made (0 when every decision came from a cache)."""
        n = C.c_int()
        flag = AUTOTUNE_LIBRARY_ONLY if library_only else 0
        check(lib().mq_plan_autotune(self.h, int(precision) | flag, C.byref(n)))
        return n.value

    def set_graphs(self, enable):
        """Replay captured HIP graphs of whole applications (default on)."""
        check(lib().mq_plan_set_graphs(self.h, int(bool(enable))))

    def profile(self, psi, sigma, reps=5, precision=PREC_F64):
        """Average ms per stage [fill, T GEMM, gather, final GEMM] and the algorithmic flops of
        the three compute stages and bytes of the gather."""
        ms = (C.c_double * 4)()
        fl = (C.c_uint64 * 3)()
        gb = C.c_uint64()
        e = self.engine
        check(lib().mq_plan_profile(self.h, dev_ptr(e, psi, self.n), dev_ptr(e, sigma, self.n), int(reps),
                                    int(precision), ms, fl, C.byref(gb)))
        return list(ms), list(fl), gb.value

    def jd(self, init, out, tol=1e-8, max_iter=8, gmres_iter=0, ortho=()):
        """Jacobi-Davidson lowest eigenpair (mq_jd_solve_ex): gmres_iter > 0 solves the correction
        equation with GMRES; ortho: device vectors the solution is kept orthogonal to."""
        e = C.c_double()
        it = C.c_int()
        g = self.engine
        ptrs = (C.c_void_p * max(1, len(ortho)))(*[dev_ptr(g, o, self.n) for o in ortho])
        check(lib().mq_jd_solve_ex(self.h, dev_ptr(g, init, self.n), dev_ptr(g, out, self.n), float(tol), int(max_iter),
                                   int(gmres_iter), len(ortho), ptrs, C.byref(e), C.byref(it)))
        return e.value, it.value

    def jd_refined(self, init, out, tol=1e-8, max_iter=8, precision=PREC_F32, max_outer=4):
        """Jacobi-Davidson with sigma in `precision` and an FP64 residual each outer step.
        Returns (energy, {"reduced": n, "fp64": n, "outer": n}, final FP64 residual norm)."""
        e = C.c_double()
        cnt = (C.c_int * 3)()
        r = C.c_double()
        g = self.engine
        check(lib().mq_jd_solve_refined(self.h, dev_ptr(g, init, self.n), dev_ptr(g, out, self.n), float(tol),
                                        int(max_iter), int(precision), int(max_outer), C.byref(e), cnt, C.byref(r)))
        return e.value, {"reduced": cnt[0], "fp64": cnt[1], "outer": cnt[2]}, r.value

    def close(self):
        if self.h:
            lib().mq_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def apply_batched(engine, plans, psis, sigmas, precision=PREC_F64):
    """sigma_i = H_i psi_i for independent plans in one batch (concurrent side streams, captured
    into one HIP graph from the second call with the same plans on)."""
    n = len(plans)
    assert n == len(psis) == len(sigmas) and n > 0
    hp = (C.c_void_p * n)(*[p.h for p in plans])
    xp = (C.c_void_p * n)(*[dev_ptr(engine, x, p.n) for p, x in zip(plans, psis)])
    yp = (C.c_void_p * n)(*[dev_ptr(engine, y, p.n) for p, y in zip(plans, sigmas)])
    check(lib().mq_sigma_apply_batched(engine.ctx, n, hp, xp, yp, int(precision)))


def left_step(engine, site, mpo_desc, bra_lp, ket_rp, left, symmetric=True, algorithm=SIGMA_AUTO):
    """overlap_mpo_left_step; algorithm SIGMA_AUTO (= grouped), SIGMA_SCHEDULE or SIGMA_GROUPED."""
    h = C.c_void_p()
    check(lib().mq_boundary_left_step_ex(engine.ctx, C.byref(site.desc), C.byref(mpo_desc.desc),
                                         dev_ptr(engine, bra_lp, site.lp_total), dev_ptr(engine, ket_rp, site.rp_total),
                                         left.h, int(symmetric), int(algorithm), C.byref(h)))
    return DeviceBoundary(engine, h)


def right_step(engine, site, mpo_desc, bra_rp, ket_rp, right, symmetric=True, algorithm=SIGMA_AUTO):
    """overlap_mpo_right_step; algorithm as for left_step."""
    h = C.c_void_p()
    check(lib().mq_boundary_right_step_ex(engine.ctx, C.byref(site.desc), C.byref(mpo_desc.desc),
                                          dev_ptr(engine, bra_rp, site.rp_total), dev_ptr(engine, ket_rp, site.rp_total),
                                          right.h, int(symmetric), int(algorithm), C.byref(h)))
    return DeviceBoundary(engine, h)


def reshape_right_to_left(engine, site, src, dst):
    check(lib().mq_reshape_right_to_left(engine.ctx, C.byref(site.desc), dev_ptr(engine, src, site.rp_total),
                                         dev_ptr(engine, dst, site.lp_total)))


def reshape_left_to_right(engine, site, src, dst):
    check(lib().mq_reshape_left_to_right(engine.ctx, C.byref(site.desc), dev_ptr(engine, src, site.lp_total),
                                         dev_ptr(engine, dst, site.rp_total)))
