// Host planner: object model construction and the three schedule builders.
//
// The builders produce the reference's ScheduleNew plan (solver/tasks.h:50-348) so that a
// caller that keeps its own create_contraction_schedule can hand the same tables to the
// device engine.  The enumeration order of T tuples, cohorts, units, lines and tasks follows
// the reference exactly, because cohort numbering of new boundaries depends on it:
//   sigma : right_tasks.hpp:39-78 (rshtm_t_tasks) + site_hamil_tasks.hpp:41-137 (shtm_tasks)
//   left  : left_tasks.hpp:38-75 (lshtm_t_tasks) + left_tasks.hpp:77-172 (lshtm_tasks)
//   right : right_tasks.hpp:39-78 + right_tasks.hpp:80-167 (rshtm_tasks)
// T tuples are looked up through a hash map instead of the reference's linear get_ti scan
// (tasks.cpp:766-775); MPS blocks are built in parallel and concatenated in block order.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <map>
#include <mutex>
#include <numeric>

#include "host_types.hpp"

namespace mq {

// ============================================================================ SiteLayout
SiteLayout::SiteLayout(const mq_site_desc& d) : left_i(d.left_i), phys_i(d.phys_i), right_i(d.right_i) {
    symm.kind = d.sym;
    MQ_REQUIRE(d.sym == MQ_SYM_U1 || d.sym == MQ_SYM_TWOU1 || d.sym == MQ_SYM_SU2U1, "unknown symmetry");
    ProductBasis rpb(phys_i, right_i, -1);
    ProductBasis lpb(phys_i, left_i, +1);
    rp_cols.resize(left_i.size());
    rp_offset.resize(left_i.size());
    for (uint32_t lb = 0; lb < left_i.size(); ++lb) {
        uint64_t cols = 0;
        for (uint32_t s = 0; s < phys_i.size(); ++s) {
            uint32_t rb = right_i.position(left_i.charges[lb] + phys_i.charges[s]);
            if (rb < right_i.size()) cols += uint64_t(phys_i.sizes[s]) * right_i.sizes[rb];
        }
        rp_cols[lb] = uint32_t(cols);
        rp_offset[lb] = rp_total;
        rp_total += round_up(uint64_t(left_i.sizes[lb]) * cols, MQ_BUFFER_ALIGNMENT);
    }
    lp_rows.resize(right_i.size());
    lp_offset.resize(right_i.size());
    for (uint32_t rb = 0; rb < right_i.size(); ++rb) {
        uint64_t rows = 0;
        for (uint32_t s = 0; s < phys_i.size(); ++s) {
            uint32_t lb = left_i.position(right_i.charges[rb] - phys_i.charges[s]);
            if (lb < left_i.size()) rows += uint64_t(phys_i.sizes[s]) * left_i.sizes[lb];
        }
        lp_rows[rb] = uint32_t(rows);
        lp_offset[rb] = lp_total;
        lp_total += round_up(rows * right_i.sizes[rb], MQ_BUFFER_ALIGNMENT);
    }
}

// ============================================================================ Mpo
Mpo::Mpo(const mq_mpo_desc& d) : row_dim(d.row_dim), col_dim(d.col_dim) {
    MQ_REQUIRE(row_dim > 0 && col_dim > 0, "mpo with zero bond dimension");
    left_spin.assign(row_dim, 0);
    right_spin.assign(col_dim, 0);
    if (d.left_spin) left_spin.assign(d.left_spin, d.left_spin + row_dim);
    if (d.right_spin) right_spin.assign(d.right_spin, d.right_spin + col_dim);
    left_herm.herm.assign(row_dim, MQ_NONE);
    left_herm.phase.assign(row_dim, 1);
    right_herm.herm.assign(col_dim, MQ_NONE);
    right_herm.phase.assign(col_dim, 1);
    if (d.left_herm) left_herm.herm.assign(d.left_herm, d.left_herm + row_dim);
    if (d.left_phase) left_herm.phase.assign(d.left_phase, d.left_phase + row_dim);
    if (d.right_herm) right_herm.herm.assign(d.right_herm, d.right_herm + col_dim);
    if (d.right_phase) right_herm.phase.assign(d.right_phase, d.right_phase + col_dim);
    for (uint32_t b = 0; b < row_dim; ++b)
        MQ_REQUIRE(left_herm.herm[b] == MQ_NONE || left_herm.herm[b] < row_dim, "left hermitian index out of range");
    for (uint32_t b = 0; b < col_dim; ++b)
        MQ_REQUIRE(right_herm.herm[b] == MQ_NONE || right_herm.herm[b] < col_dim, "right hermitian index out of range");

    MQ_REQUIRE(d.n_ops == 0 || d.ops, "mpo without op table");
    ops.resize(d.n_ops);
    for (uint32_t k = 0; k < d.n_ops; ++k) {
        const mq_op_desc& od = d.ops[k];
        Op& op = ops[k];
        op.spin = od.spin;
        MQ_REQUIRE(od.n_blocks == 0 || (od.in_charge && od.out_charge && od.entry_begin), "op descriptor");
        uint32_t ne = od.n_blocks ? od.entry_begin[od.n_blocks] : 0;
        MQ_REQUIRE(ne == 0 || (od.row && od.col && od.coeff), "op entries");
        for (uint32_t w = 0; w < od.n_blocks; ++w) {
            MQ_REQUIRE(od.entry_begin[w] <= od.entry_begin[w + 1], "op entry ranges");
            op.blocks.push_back({from_c(od.in_charge[w]), from_c(od.out_charge[w]), od.entry_begin[w],
                                 od.entry_begin[w + 1]});
        }
        op.row.assign(od.row, od.row + ne);
        op.col.assign(od.col, od.col + ne);
        op.coeff.assign(od.coeff, od.coeff + ne);
        op.row_spin.assign(ne, 0);
        op.col_spin.assign(ne, 0);
        if (od.row_spin) op.row_spin.assign(od.row_spin, od.row_spin + ne);
        if (od.col_spin) op.col_spin.assign(od.col_spin, od.col_spin + ne);
    }
    terms.resize(d.n_terms);
    rows.assign(row_dim, {});
    cols.assign(col_dim, {});
    for (uint32_t t = 0; t < d.n_terms; ++t) {
        uint32_t b1 = d.term_b1[t], b2 = d.term_b2[t];
        MQ_REQUIRE(b1 < row_dim && b2 < col_dim && d.term_op[t] < d.n_ops, "mpo term out of range");
        if (t > 0) {
            uint32_t p1 = d.term_b1[t - 1], p2 = d.term_b2[t - 1];
            MQ_REQUIRE(p1 < b1 || (p1 == b1 && p2 <= b2), "mpo terms must be sorted by (b1, b2)");
        }
        terms[t] = {d.term_op[t], d.term_scale[t]};
        if (!rows[b1].empty() && rows[b1].back().other == b2) {
            rows[b1].back().t1 = t + 1;
        } else {
            rows[b1].push_back({b2, t, t + 1});
        }
    }
    // columns: b1 in increasing order for each b2
    for (uint32_t b1 = 0; b1 < row_dim; ++b1)
        for (auto& c : rows[b1]) cols[c.other].push_back({b1, c.t0, c.t1});
}

// ============================================================================ BoundaryIndex
BoundaryIndex::BoundaryIndex(const mq_bindex_desc& d) : bra(d.bra), ket(d.ket), aux_dim(d.aux_dim) {
    uint32_t n = d.n_cohorts;
    MQ_REQUIRE(n == 0 || (d.bra_pos && d.ket_pos && d.tr && d.n_blocks && d.offsets && d.conj_scales),
               "boundary index descriptor with null arrays");
    bra_pos.assign(d.bra_pos, d.bra_pos + n);
    ket_pos.assign(d.ket_pos, d.ket_pos + n);
    tr.assign(d.tr, d.tr + n);
    n_blocks.assign(d.n_blocks, d.n_blocks + n);
    offsets.assign(d.offsets, d.offsets + size_t(n) * aux_dim);
    conj_scales.assign(d.conj_scales, d.conj_scales + size_t(n) * aux_dim);
    transposes.assign(size_t(n) * aux_dim, 0);
    if (d.transposes) transposes.assign(d.transposes, d.transposes + size_t(n) * aux_dim);
    for (uint32_t ci = 0; ci < n; ++ci) {
        MQ_REQUIRE(bra_pos[ci] < bra.size() && ket_pos[ci] < ket.size(), "cohort position out of range");
        for (uint32_t b = 0; b < aux_dim; ++b) {
            int64_t o = offset(ci, b);
            MQ_REQUIRE(o >= -1, "negative offset");
        }
    }
    rebuild_lookup();
    MQ_REQUIRE(lbrb_.size() == n, "duplicate cohort (bra, ket) pair");
}

void BoundaryIndex::rebuild_lookup() {
    lbrb_.clear();
    for (uint32_t ci = 0; ci < n_cohorts(); ++ci) lbrb_[uint64_t(bra_pos[ci]) << 32 | ket_pos[ci]] = ci;
}

// BoundaryIndex::add_cohort (boundary.h:138-159)
uint32_t BoundaryIndex::add_cohort(uint32_t lb, uint32_t rb, const std::vector<int64_t>& off) {
    uint32_t ci = n_cohorts();
    bra_pos.push_back(lb);
    ket_pos.push_back(rb);
    tr.push_back(0);
    uint32_t nb = 0;
    for (auto o : off) nb += (o > -1);
    n_blocks.push_back(nb);
    offsets.insert(offsets.end(), off.begin(), off.end());
    conj_scales.insert(conj_scales.end(), off.size(), 1.0);
    transposes.insert(transposes.end(), off.size(), 0);
    lbrb_[uint64_t(lb) << 32 | rb] = ci;
    return ci;
}

// SU2 correction of a transposed boundary block (boundary.h:43-62); 1 for abelian groups.
static double conjugate_correction(const Symm& sy, const Charge& lc, const Charge& rc) {
    if (!sy.su2()) return 1.0;
    int sl = sy.spin(lc), sr = sy.spin(rc);
    double S = std::min(sl, sr);
    switch (sr - sl) {
        case 0: return 1.0;
        case 1: return -std::sqrt((S + 1.) / (S + 2.));
        case -1: return std::sqrt((S + 2.) / (S + 1.));
        case 2: return -std::sqrt((S + 1.) / (S + 3.));
        case -2: return -std::sqrt((S + 3.) / (S + 1.));
        default: fail(MQ_ERR_INVALID, "hermitian conjugate only implemented up to rank 1");
    }
}

// BoundaryIndex::complement_transpose (boundary.h:161-191): Hermitian partners of skipped
// blocks become transposed references into the partner cohort.  Bra and ket indices are
// identical in every sweep step, so partner positions are found by charge.
void BoundaryIndex::complement_transpose(const Hermitian& herm, bool forward, const Symm& sy) {
    MQ_REQUIRE(herm.size() == aux_dim, "hermitian size != aux_dim");
    for (uint32_t rb = 0; rb < ket.size(); ++rb)
        for (uint32_t lb = 0; lb < bra.size(); ++lb) {
            auto it = lbrb_.find(uint64_t(lb) << 32 | rb);
            if (it == lbrb_.end()) continue;
            uint32_t ci_A = it->second;
            Charge kc = ket.charges[rb], bc = bra.charges[lb];
            uint32_t ci_B = cohort_index(kc, bc);
            if (ci_B == n_cohorts()) {
                uint32_t lb2 = bra.position(kc), rb2 = ket.position(bc);
                MQ_REQUIRE(lb2 < bra.size() && rb2 < ket.size(), "complement_transpose needs bra == ket indices");
                ci_B = add_cohort(lb2, rb2, std::vector<int64_t>(aux_dim, -1));
                tr[ci_B] = 1;
            }
            for (uint32_t b = 0; b < aux_dim; ++b) {
                if (herm.skip(b, kc, bc)) {
                    size_t iB = size_t(ci_B) * aux_dim + b;
                    MQ_REQUIRE(offsets[iB] == -1, "complement_transpose: block already present");
                    offsets[iB] = offsets[size_t(ci_A) * aux_dim + herm.conj(b)];
                    conj_scales[iB] = conjugate_correction(sy, kc, bc) *
                                      double(herm.phase[forward ? b : herm.conj(b)]);
                    transposes[iB] = 1;
                }
            }
        }
}

void BoundaryIndex::fill_desc(mq_bindex_desc* d, std::vector<mq_charge>& bra_c, std::vector<mq_charge>& ket_c) const {
    bra_c = bra.c_charges();
    ket_c = ket.c_charges();
    d->bra = {bra.size(), bra_c.data(), bra.sizes.data()};
    d->ket = {ket.size(), ket_c.data(), ket.sizes.data()};
    d->n_cohorts = n_cohorts();
    d->aux_dim = aux_dim;
    d->bra_pos = bra_pos.data();
    d->ket_pos = ket_pos.data();
    d->tr = tr.data();
    d->n_blocks = n_blocks.data();
    d->offsets = offsets.data();
    d->conj_scales = conj_scales.data();
    d->transposes = transposes.data();
}

// ============================================================================ SU2 couplings
// Racah formula for the 6j symbol and the 6j-sum for the 9j symbol; all spins are passed
// doubled (2j) as in gsl_sf_coupling_9j.
static long double fact(int n) {
    static std::vector<long double> f = [] {
        std::vector<long double> v(400);
        v[0] = 1;
        for (int i = 1; i < 400; ++i) v[i] = v[i - 1] * i;
        return v;
    }();
    if (n < 0 || n >= 400) fail(MQ_ERR_CAPACITY, "factorial argument out of range");
    return f[n];
}
static bool tri2(int a, int b, int c) {  // doubled spins
    return c >= std::abs(a - b) && c <= a + b && ((a + b + c) % 2 == 0);
}
static long double delta2(int a, int b, int c) {
    return std::sqrt(fact((a + b - c) / 2) * fact((a - b + c) / 2) * fact((-a + b + c) / 2) / fact((a + b + c) / 2 + 1));
}
static long double sixj2(int j1, int j2, int j3, int j4, int j5, int j6) {
    if (!tri2(j1, j2, j3) || !tri2(j1, j5, j6) || !tri2(j4, j2, j6) || !tri2(j4, j5, j3)) return 0;
    int a1 = (j1 + j2 + j3) / 2, a2 = (j1 + j5 + j6) / 2, a3 = (j4 + j2 + j6) / 2, a4 = (j4 + j5 + j3) / 2;
    int b1 = (j1 + j2 + j4 + j5) / 2, b2 = (j2 + j3 + j5 + j6) / 2, b3 = (j3 + j1 + j6 + j4) / 2;
    int tmin = std::max(std::max(a1, a2), std::max(a3, a4)), tmax = std::min(b1, std::min(b2, b3));
    long double sum = 0;
    for (int t = tmin; t <= tmax; ++t) {
        long double term = fact(t + 1) / (fact(t - a1) * fact(t - a2) * fact(t - a3) * fact(t - a4) *
                                          fact(b1 - t) * fact(b2 - t) * fact(b3 - t));
        sum += (t % 2 ? -term : term);
    }
    return sum * delta2(j1, j2, j3) * delta2(j1, j5, j6) * delta2(j4, j2, j6) * delta2(j4, j5, j3);
}

double wigner9j(int a, int b, int c, int d, int e, int f, int g, int h, int i) {
    if (!tri2(a, b, c) || !tri2(d, e, f) || !tri2(g, h, i) || !tri2(a, d, g) || !tri2(b, e, h) || !tri2(c, f, i))
        return 0.0;
    int xmin = std::max(std::abs(a - i), std::max(std::abs(d - h), std::abs(b - f)));
    int xmax = std::min(a + i, std::min(d + h, b + f));
    long double sum = 0;
    for (int x = xmin; x <= xmax; x += 2) {
        long double t = (x + 1) * sixj2(a, b, c, f, i, x) * sixj2(d, e, f, b, x, h) * sixj2(g, h, i, x, a, d);
        sum += (x % 2 ? -t : t);  // (-1)^(2x) with x doubled
    }
    return double(sum);
}

static inline double mod_coupling(int a, int b, int c, int d, int e, int f, int g, int h, int i) {
    return std::sqrt((g + 1.) * (h + 1.) * (c + 1.) * (f + 1.)) * wigner9j(a, b, c, d, e, f, g, h, i);
}
static inline bool triangle(int a, int b, int c) { return std::abs(a - b) <= c && c <= a + b; }

// SU2::set_coupling (gsl_coupling.h:72-92)
void su2_set_coupling(int a, int b, int c, int d, int e, int f, int g, int h, int i, double init, double cpl[4]) {
    cpl[0] = cpl[1] = cpl[2] = cpl[3] = 0.0;
    double pref = std::sqrt((i + 1.) * (a + 1.) / ((g + 1.) * (c + 1.))) * init;
    if (triangle(a, b, c)) {
        cpl[0] = pref * mod_coupling(a, b, c, d, e, f, g, h, i);
        cpl[2] = pref * mod_coupling(a, b, c, d, e, f, g, 2, i);
    }
    if (triangle(a, 2, c)) {
        cpl[1] = pref * mod_coupling(a, 2, c, d, e, f, g, h, i);
        cpl[3] = pref * mod_coupling(a, 2, c, d, e, f, g, 2, i);
    }
}


// ============================================================================ per-block builder
namespace {

struct UnitB {
    uint32_t offset = 0, ms = 0, pending = 0;
    std::vector<uint32_t> line_b1, line_count, ti, col;
    std::vector<double> alpha;
};

struct CohB {
    uint32_t lb, rb, ls, rs, ci, ci_eff, nSrows, stripe = 0;
    std::vector<UnitB> units;
    std::vector<int64_t> mpo_offsets;   // counts per MPO row, then block offsets after finalize
    const std::vector<uint32_t>* sfold;
    CohB(const std::vector<uint32_t>& sf, uint32_t nunits, uint32_t lb_, uint32_t rb_, uint32_t ls_,
         uint32_t rs_, uint32_t ci_, uint32_t cie, uint32_t mpodim)
        : lb(lb_), rb(rb_), ls(ls_), rs(rs_), ci(ci_), ci_eff(cie), nSrows(mpodim), units(nunits),
          mpo_offsets(mpodim, 0), sfold(&sf) {}
    void add_unit(uint32_t s, uint32_t d, uint32_t m, uint32_t off) {
        for (uint32_t i = 0; i < d; ++i) {
            UnitB& u = units[(*sfold)[s] + i];
            u.offset = i * m + off;
            u.ms = m;
            stripe += m;
        }
    }
    void push(uint32_t s, uint32_t ss2, double scale, uint32_t ti, uint32_t col) {
        UnitB& u = units[(*sfold)[s] + ss2];
        u.alpha.push_back(scale);
        u.ti.push_back(ti);
        u.col.push_back(col);
        ++u.pending;
    }
    void add_line(uint32_t b) {
        for (auto& u : units)
            if (u.pending) {
                u.line_b1.push_back(b);
                u.line_count.push_back(u.pending);
                mpo_offsets[b] += u.pending;
                u.pending = 0;
            }
    }
    // Cohort::compute_mpo_offsets (tasks.cpp:506-526)
    void finalize() {
        std::vector<int64_t> rank(mpo_offsets.size());
        uint32_t n = 0;
        for (size_t b = 0; b < mpo_offsets.size(); ++b) rank[b] = mpo_offsets[b] ? n++ : 0;
        nSrows = n;
        for (auto& u : units)
            for (auto& b : u.line_b1) b = uint32_t(rank[b]);
        uint64_t bs = uint64_t(ls) * rs;
        uint32_t cnt = 0;
        for (auto& o : mpo_offsets) o = o ? int64_t(bs * cnt++) : -1;
    }
    size_t n_tasks() const {
        size_t n = 0;
        for (auto& u : units) n += u.alpha.size();
        return n;
    }
};

struct BlockB {
    std::vector<uint32_t> t_mps_offset, t_ci, t_ci_eff, t_lb_ket;
    std::vector<uint64_t> t_sza;
    std::unordered_map<uint64_t, uint32_t> t_lookup;
    std::vector<CohB> cohorts;
    void push_t(uint32_t off, uint32_t ci, uint32_t cie, uint32_t lbk, uint64_t sza) {
        t_lookup[uint64_t(off) << 32 | ci] = uint32_t(t_mps_offset.size());
        t_mps_offset.push_back(off);
        t_ci.push_back(ci);
        t_ci_eff.push_back(cie);
        t_lb_ket.push_back(lbk);
        t_sza.push_back(sza);
    }
    uint32_t get_ti(uint32_t off, uint32_t ci) const {
        auto it = t_lookup.find(uint64_t(off) << 32 | ci);
        return it == t_lookup.end() ? MQ_NONE : it->second;
    }
};

// detail::op_iterate (op_iterate.hpp:37-69)
inline void op_iterate(const Op& W, const OpBlock& wb, const double cpl[4], CohB& coh, uint32_t s,
                       uint32_t m2_size, const BlockB& blk, uint32_t mps_offset, uint32_t ci,
                       uint32_t boundary_col) {
    for (uint32_t e = wb.e0; e < wb.e1; ++e) {
        unsigned rspin = W.row_spin[e], cspin = W.col_spin[e], casenr = 0;
        if (rspin == 2 && cspin == 2) casenr = 3;
        else if (rspin == 2) casenr = 1;
        else if (cspin == 2) casenr = 2;
        double scale = W.coeff[e] * cpl[casenr];
        uint32_t ti = blk.get_ti(mps_offset + W.row[e] * m2_size, ci);
        if (ti != MQ_NONE) coh.push(s, W.col[e], scale, ti, boundary_col);
    }
}

std::vector<uint32_t> make_sfold(const Index& phys, uint32_t* total) {
    std::vector<uint32_t> sf(phys.size());
    uint32_t sum = 0;
    for (uint32_t s = 0; s < phys.size(); ++s) sf[s] = sum, sum += phys.sizes[s];
    *total = sum;
    return sf;
}

// T tuples of the sigma / right step (right_tasks.hpp:39-78), for ket left sector lb_ket.
void right_t_tasks(const BoundaryIndex& right, const Index& left_i, const Index& right_i, const Index& phys_i,
                   const ProductBasis& right_pb, uint32_t lb_ket, BlockB& blk) {
    Charge lc_ket = left_i.charges[lb_ket];
    for (uint32_t s = 0; s < phys_i.size(); ++s) {
        Charge rc_ket = lc_ket + phys_i.charges[s];
        uint32_t rb_ket = right_i.position(rc_ket);
        if (rb_ket == right_i.size()) continue;
        uint32_t rs_ket = right_i.sizes[rb_ket];
        uint32_t ket_offset = uint32_t(right_pb(s, rb_ket));
        for (uint32_t rb_bra = 0; rb_bra < right_i.size(); ++rb_bra) {
            uint32_t ci = right.cohort_index(rc_ket, right_i.charges[rb_bra]);
            if (ci == right.n_cohorts()) continue;
            uint32_t ci_eff = right.tr[ci] ? right.cohort_index(right_i.charges[rb_bra], rc_ket) : ci;
            uint64_t sz = uint64_t(left_i.sizes[lb_ket]) * right_i.sizes[rb_bra] * right.n_blocks[ci_eff];
            uint64_t sza = round_up(sz, MQ_BUFFER_ALIGNMENT);
            for (uint32_t ss = 0; ss < phys_i.sizes[s]; ++ss) blk.push_t(ket_offset + ss * rs_ket, ci, ci_eff, lb_ket, sza);
        }
    }
}

// T tuples of the left step (left_tasks.hpp:38-75), for ket right sector rb_ket.
void left_t_tasks(const BoundaryIndex& left, const Index& left_i, const Index& right_i, const Index& phys_i,
                  const ProductBasis& right_pb, uint32_t rb_ket, BlockB& blk) {
    Charge rc_ket = right_i.charges[rb_ket];
    uint32_t rs_ket = right_i.sizes[rb_ket];
    for (uint32_t s = 0; s < phys_i.size(); ++s) {
        Charge lc_ket = rc_ket - phys_i.charges[s];
        uint32_t lb_ket = left_i.position(lc_ket);
        if (lb_ket == left_i.size()) continue;
        uint32_t ket_offset = uint32_t(right_pb(s, rb_ket));
        for (uint32_t lb_bra = 0; lb_bra < left_i.size(); ++lb_bra) {
            uint32_t ci = left.cohort_index(left_i.charges[lb_bra], lc_ket);
            if (ci == left.n_cohorts()) continue;
            uint32_t ci_eff = left.tr[ci] ? left.cohort_index(lc_ket, left_i.charges[lb_bra]) : ci;
            uint64_t sz = uint64_t(left_i.sizes[lb_bra]) * rs_ket * left.n_blocks[ci_eff];
            uint64_t sza = round_up(sz, MQ_BUFFER_ALIGNMENT);
            for (uint32_t ss = 0; ss < phys_i.sizes[s]; ++ss) blk.push_t(ket_offset + ss * rs_ket, ci, ci_eff, lb_ket, sza);
        }
    }
}

void flatten(std::vector<BlockB>& blocks, Schedule& S) {
    S.n_blocks = uint32_t(blocks.size());
    for (auto& blk : blocks) {
        S.t_mps_offset.insert(S.t_mps_offset.end(), blk.t_mps_offset.begin(), blk.t_mps_offset.end());
        S.t_ci.insert(S.t_ci.end(), blk.t_ci.begin(), blk.t_ci.end());
        S.t_ci_eff.insert(S.t_ci_eff.end(), blk.t_ci_eff.begin(), blk.t_ci_eff.end());
        S.t_lb_ket.insert(S.t_lb_ket.end(), blk.t_lb_ket.begin(), blk.t_lb_ket.end());
        S.t_sza.insert(S.t_sza.end(), blk.t_sza.begin(), blk.t_sza.end());
        S.t_begin.push_back(S.t_mps_offset.size());
        for (auto& c : blk.cohorts) {
            S.coh_lb.push_back(c.lb);
            S.coh_rb.push_back(c.rb);
            S.coh_ls.push_back(c.ls);
            S.coh_rs.push_back(c.rs);
            S.coh_ci.push_back(c.ci);
            S.coh_ci_eff.push_back(c.ci_eff);
            S.coh_nSrows.push_back(c.nSrows);
            S.coh_stripe.push_back(c.stripe);
            if (S.kind != Schedule::SIGMA) {
                MQ_REQUIRE(c.mpo_offsets.size() == S.aux_dim, "cohort offsets size");
                S.coh_mpo_offsets.insert(S.coh_mpo_offsets.end(), c.mpo_offsets.begin(), c.mpo_offsets.end());
            }
            for (auto& u : c.units) {
                S.unit_offset.push_back(u.offset);
                S.unit_ms.push_back(u.ms);
                S.line_b1.insert(S.line_b1.end(), u.line_b1.begin(), u.line_b1.end());
                S.line_count.insert(S.line_count.end(), u.line_count.begin(), u.line_count.end());
                S.unit_line_begin.push_back(S.line_b1.size());
                S.task_ti.insert(S.task_ti.end(), u.ti.begin(), u.ti.end());
                S.task_col.insert(S.task_col.end(), u.col.begin(), u.col.end());
                S.task_alpha.insert(S.task_alpha.end(), u.alpha.begin(), u.alpha.end());
                S.unit_task_begin.push_back(S.task_ti.size());
            }
            S.coh_unit_begin.push_back(S.unit_offset.size());
        }
        S.coh_begin.push_back(S.coh_lb.size());
        BlockB().t_lookup.swap(blk.t_lookup);
    }
}

}  // namespace

// ============================================================================ sigma
// create_contraction_schedule (site_hamil_schedule.hpp:40-93): per ket left sector lb_in,
// rshtm_t_tasks then shtm_tasks (site_hamil_tasks.hpp:41-137).
Schedule build_sigma_schedule(const SiteLayout& site, const Mpo& mpo, const BoundaryIndex& left,
                              const BoundaryIndex& right) {
    const Index &left_i = site.left_i, &right_i = site.right_i, &phys_i = site.phys_i;
    MQ_REQUIRE(left.aux_dim == mpo.row_dim, "left boundary aux_dim != mpo row_dim");
    MQ_REQUIRE(right.aux_dim == mpo.col_dim, "right boundary aux_dim != mpo col_dim");
    require_bond(left, left_i, "left");
    require_bond(right, right_i, "right");
    const Symm sy = site.symm;
    ProductBasis right_pb(phys_i, right_i, -1);
    uint32_t nunits;
    std::vector<uint32_t> sfold = make_sfold(phys_i, &nunits);
    std::vector<BlockB> blocks(left_i.size());
    std::vector<std::string> errors;
    std::mutex emu;

#pragma omp parallel for schedule(dynamic, 1)
    for (int mb_ = 0; mb_ < int(left_i.size()); ++mb_) {
        try {
            uint32_t lb_in = uint32_t(mb_);
            BlockB& blk = blocks[lb_in];
            CouplingCache ccache;
            right_t_tasks(right, left_i, right_i, phys_i, right_pb, lb_in, blk);
            Charge lc_in = left_i.charges[lb_in];
            uint32_t ls_in = left_i.sizes[lb_in];
            for (uint32_t lb_out = 0; lb_out < left_i.size(); ++lb_out) {
                Charge lc_out = left_i.charges[lb_out];
                uint32_t ls_out = left_i.sizes[lb_out];
                uint32_t ci = left.cohort_index(lc_out, lc_in);
                if (ci == left.n_cohorts()) continue;
                uint32_t ci_eff = left.tr[ci] ? left.cohort_index(lc_in, lc_out) : ci;
                MQ_REQUIRE(ci_eff < left.n_cohorts(), "transposed left cohort without partner");
                CohB coh(sfold, nunits, lb_in, lb_out, ls_in, ls_out, ci, ci_eff, left.n_blocks[ci_eff]);
                for (uint32_t s = 0; s < phys_i.size(); ++s) {
                    Charge phys_out = phys_i.charges[s];
                    Charge rc_out = lc_out + phys_out;
                    uint32_t rb_out = right_i.position(rc_out);
                    if (rb_out == right_i.size()) continue;
                    uint32_t rs_out = right_i.sizes[rb_out];
                    uint32_t out_offset = uint32_t(right_pb(s, rb_out));
                    coh.add_unit(s, phys_i.sizes[s], rs_out, out_offset);
                    // Wigner9jCache(lc_out, lc_in, rc_out): J = S(lc_in), I = S(lc_out), I' = S(rc_out)
                    int J = sy.spin(lc_in), I = sy.spin(lc_out), Ip = sy.spin(rc_out);
                    for (uint32_t b1 = 0; b1 < mpo.row_dim; ++b1) {
                        if (!left.has_block(ci, b1)) continue;
                        uint32_t left_idx = uint32_t(left.offset(ci, b1) / (int64_t(ls_in) * ls_out));
                        int A = mpo.left_spin[b1];
                        if (sy.su2() && !triangle(sy.spin(lc_in), A, sy.spin(lc_out))) continue;
                        for (const auto& cell : mpo.rows[b1]) {
                            uint32_t b2 = cell.other;
                            for (uint32_t t = cell.t0; t < cell.t1; ++t) {
                                const Op& W = mpo.ops[mpo.terms[t].op];
                                int K = W.spin, Ap = mpo.right_spin[b2];
                                for (const OpBlock& wb : W.blocks) {
                                    if (phys_out != wb.out) continue;
                                    Charge phys_in = wb.in;
                                    Charge rc_in = lc_in + phys_in;
                                    uint32_t ci_right = right.cohort_index(rc_in, rc_out);
                                    if (!right.has_block(ci_right, b2)) continue;
                                    uint32_t rb_in = right_i.position(rc_in);
                                    if (rb_in == right_i.size()) continue;
                                    uint32_t rs_in = right_i.sizes[rb_in];
                                    uint32_t s_in = phys_i.position(phys_in);
                                    MQ_REQUIRE(s_in < phys_i.size(), "operator charge not in physical index");
                                    uint32_t in_offset = uint32_t(right_pb(s_in, rb_in));
                                    int64_t right_offset = right.offset(ci_right, b2);
                                    double scale = right.conjugate_scale(ci_right, b2) * mpo.terms[t].scale *
                                                   left.conjugate_scale(ci, b1);
                                    double cpl[4];
                                    if (sy.su2()) {
                                        int Jp = sy.spin(rc_in);
                                        const double* u = ccache.get(J, std::abs(J - Jp), Jp, A, K, Ap, I, std::abs(I - Ip), Ip);
                                        for (int k = 0; k < 4; ++k) cpl[k] = u[k] * scale;
                                    } else {
                                        cpl[0] = cpl[1] = cpl[2] = cpl[3] = scale;
                                    }
                                    op_iterate(W, wb, cpl, coh, s, rs_in, blk, in_offset, ci_right,
                                               uint32_t(right_offset / rs_in));
                                }
                            }
                        }
                        coh.add_line(left_idx);
                    }
                }
                if (coh.n_tasks()) blk.cohorts.push_back(std::move(coh));
            }
        } catch (const std::exception& e) {
            std::lock_guard<std::mutex> lk(emu);
            errors.push_back(e.what());
        }
    }
    if (!errors.empty()) fail(MQ_ERR_INVALID, errors[0]);
    Schedule S;
    S.kind = Schedule::SIGMA;
    flatten(blocks, S);
    return S;
}

// ============================================================================ left step
// overlap_mpo_left_step schedule (move_boundary.hpp:186-209): per ket right sector rb_ket,
// lshtm_t_tasks and lshtm_tasks; new index from the cohort offsets + complement_transpose.
Schedule build_left_schedule(const SiteLayout& site, const Mpo& mpo, const BoundaryIndex& left, bool symmetric) {
    const Index &left_i = site.left_i, &right_i = site.right_i, &phys_i = site.phys_i;
    MQ_REQUIRE(left.aux_dim == mpo.row_dim, "left boundary aux_dim != mpo row_dim");
    require_bond(left, left_i, "left");
    const Symm sy = site.symm;
    ProductBasis ket_right_pb(phys_i, right_i, -1);
    ProductBasis bra_left_pb(phys_i, left_i, +1);
    uint32_t nunits;
    std::vector<uint32_t> sfold = make_sfold(phys_i, &nunits);
    std::vector<BlockB> blocks(right_i.size());
    std::vector<std::string> errors;
    std::mutex emu;
    const int site_basis_max_diff = 2;

#pragma omp parallel for schedule(dynamic, 1)
    for (int rb_ = 0; rb_ < int(right_i.size()); ++rb_) {
        try {
            uint32_t rb_ket = uint32_t(rb_);
            BlockB& blk = blocks[rb_ket];
            CouplingCache ccache;
            left_t_tasks(left, left_i, right_i, phys_i, ket_right_pb, rb_ket, blk);
            Charge rc_ket = right_i.charges[rb_ket];
            uint32_t rs_ket = right_i.sizes[rb_ket];
            for (uint32_t rb_bra = 0; rb_bra < right_i.size(); ++rb_bra) {
                Charge rc_bra = right_i.charges[rb_bra];
                if (std::abs(sy.particle_number(rc_bra) - sy.particle_number(rc_ket)) > site_basis_max_diff) continue;
                uint32_t rs_bra = right_i.sizes[rb_bra];
                CohB coh(sfold, nunits, rb_bra, rb_ket, rs_bra, rs_ket, 0, 0, mpo.col_dim);
                for (uint32_t s = 0; s < phys_i.size(); ++s) {
                    Charge phys_out = phys_i.charges[s];
                    Charge lc_bra = rc_bra - phys_out;
                    uint32_t lb_bra = left_i.position(lc_bra);
                    if (lb_bra == left_i.size()) continue;
                    uint32_t ls_bra = left_i.sizes[lb_bra];
                    uint32_t bra_offset = uint32_t(bra_left_pb(s, lb_bra));
                    coh.add_unit(s, phys_i.sizes[s], ls_bra, bra_offset);
                    for (uint32_t b2 = 0; b2 < mpo.col_dim; ++b2) {
                        if (symmetric && mpo.right_herm.skip(b2, rc_bra, rc_ket)) continue;
                        int Ap = mpo.right_spin[b2];
                        if (sy.su2() && !triangle(sy.spin(rc_ket), Ap, sy.spin(rc_bra))) continue;
                        for (const auto& cell : mpo.cols[b2]) {
                            uint32_t b1 = cell.other;
                            for (uint32_t t = cell.t0; t < cell.t1; ++t) {
                                const Op& W = mpo.ops[mpo.terms[t].op];
                                int K = W.spin, A = mpo.left_spin[b1];
                                for (const OpBlock& wb : W.blocks) {
                                    if (phys_out != wb.out) continue;
                                    Charge phys_in = wb.in;
                                    Charge lc_ket = rc_ket - phys_in;
                                    uint32_t lb_ket = left_i.position(lc_ket);
                                    if (lb_ket == left_i.size()) continue;
                                    uint32_t ci = left.cohort_index(lc_bra, lc_ket);
                                    if (!left.has_block(ci, b1)) continue;
                                    uint32_t ls_ket = left_i.sizes[lb_ket];
                                    uint32_t s_in = phys_i.position(phys_in);
                                    MQ_REQUIRE(s_in < phys_i.size(), "operator charge not in physical index");
                                    uint32_t ket_offset = uint32_t(ket_right_pb(s_in, rb_ket));
                                    int64_t left_offset = left.offset(ci, b1);
                                    double scale = left.conjugate_scale(ci, b1) * mpo.terms[t].scale;
                                    double cpl[4];
                                    if (sy.su2()) {
                                        const double* u = ccache.get(
                                            sy.spin(lc_ket), std::abs(sy.spin(rc_ket - lc_ket)), sy.spin(rc_ket), A, K, Ap,
                                            sy.spin(lc_bra), std::abs(sy.spin(rc_bra - lc_bra)), sy.spin(rc_bra));
                                        for (int k = 0; k < 4; ++k) cpl[k] = u[k] * scale;
                                    } else {
                                        cpl[0] = cpl[1] = cpl[2] = cpl[3] = scale;
                                    }
                                    op_iterate(W, wb, cpl, coh, s, rs_ket, blk, ket_offset, ci,
                                               uint32_t(left_offset / ls_ket));
                                }
                            }
                        }
                        coh.add_line(b2);
                    }
                }
                coh.finalize();
                if (coh.n_tasks()) blk.cohorts.push_back(std::move(coh));
            }
        } catch (const std::exception& e) {
            std::lock_guard<std::mutex> lk(emu);
            errors.push_back(e.what());
        }
    }
    if (!errors.empty()) fail(MQ_ERR_INVALID, errors[0]);
    Schedule S;
    S.kind = Schedule::LEFT;
    S.aux_dim = mpo.col_dim;
    flatten(blocks, S);
    // new BoundaryIndex(bra_right_i, ket_right_i) (move_boundary.hpp:202-205)
    S.out_index = BoundaryIndex(right_i, right_i, mpo.col_dim);
    for (uint32_t c = 0; c < S.n_cohorts(); ++c) {
        std::vector<int64_t> off(S.coh_mpo_offsets.begin() + size_t(c) * S.aux_dim,
                                 S.coh_mpo_offsets.begin() + size_t(c + 1) * S.aux_dim);
        S.out_index.add_cohort(S.coh_lb[c], S.coh_rb[c], off);
    }
    if (symmetric) {
        S.out_index.complement_transpose(mpo.right_herm, true, sy);
    }
    return S;
}

// ============================================================================ right step
// overlap_mpo_right_step schedule (move_boundary.hpp:311-329): per ket left sector lb_ket,
// rshtm_t_tasks and rshtm_tasks (right_tasks.hpp:80-167).
Schedule build_right_schedule(const SiteLayout& site, const Mpo& mpo, const BoundaryIndex& right, bool symmetric) {
    const Index &left_i = site.left_i, &right_i = site.right_i, &phys_i = site.phys_i;
    MQ_REQUIRE(right.aux_dim == mpo.col_dim, "right boundary aux_dim != mpo col_dim");
    require_bond(right, right_i, "right");
    const Symm sy = site.symm;
    ProductBasis bra_right_pb(phys_i, right_i, -1);
    uint32_t nunits;
    std::vector<uint32_t> sfold = make_sfold(phys_i, &nunits);
    std::vector<BlockB> blocks(left_i.size());
    std::vector<std::string> errors;
    std::mutex emu;
    const int site_basis_max_diff = 2;

#pragma omp parallel for schedule(dynamic, 1)
    for (int lb_ = 0; lb_ < int(left_i.size()); ++lb_) {
        try {
            uint32_t lb_ket = uint32_t(lb_);
            BlockB& blk = blocks[lb_ket];
            CouplingCache ccache;
            right_t_tasks(right, left_i, right_i, phys_i, bra_right_pb, lb_ket, blk);
            Charge lc_ket = left_i.charges[lb_ket];
            uint32_t ls_ket = left_i.sizes[lb_ket];
            for (uint32_t lb_bra = 0; lb_bra < left_i.size(); ++lb_bra) {
                Charge lc_bra = left_i.charges[lb_bra];
                if (std::abs(sy.particle_number(lc_ket) - sy.particle_number(lc_bra)) > site_basis_max_diff) continue;
                uint32_t ls_bra = left_i.sizes[lb_bra];
                CohB coh(sfold, nunits, lb_ket, lb_bra, ls_ket, ls_bra, 0, 0, mpo.row_dim);
                for (uint32_t s = 0; s < phys_i.size(); ++s) {
                    Charge phys_out = phys_i.charges[s];
                    Charge rc_bra = lc_bra + phys_out;
                    uint32_t rb_bra = right_i.position(rc_bra);
                    if (rb_bra == right_i.size()) continue;
                    uint32_t rs_bra = right_i.sizes[rb_bra];
                    uint32_t bra_offset = uint32_t(bra_right_pb(s, rb_bra));
                    coh.add_unit(s, phys_i.sizes[s], rs_bra, bra_offset);
                    // Wigner9jCache(lc_bra, lc_ket, rc_bra): J = S(lc_ket), I = S(lc_bra), I' = S(rc_bra)
                    int J = sy.spin(lc_ket), I = sy.spin(lc_bra), Ip = sy.spin(rc_bra);
                    for (uint32_t b1 = 0; b1 < mpo.row_dim; ++b1) {
                        if (symmetric && mpo.left_herm.skip(b1, lc_ket, lc_bra)) continue;
                        int A = mpo.left_spin[b1];
                        if (sy.su2() && !triangle(sy.spin(lc_ket), A, sy.spin(lc_bra))) continue;
                        for (const auto& cell : mpo.rows[b1]) {
                            uint32_t b2 = cell.other;
                            for (uint32_t t = cell.t0; t < cell.t1; ++t) {
                                const Op& W = mpo.ops[mpo.terms[t].op];
                                int K = W.spin, Ap = mpo.right_spin[b2];
                                for (const OpBlock& wb : W.blocks) {
                                    if (phys_out != wb.out) continue;
                                    Charge phys_in = wb.in;
                                    Charge rc_ket = lc_ket + phys_in;
                                    uint32_t ci = right.cohort_index(rc_ket, rc_bra);
                                    if (!right.has_block(ci, b2)) continue;
                                    uint32_t rb_ket = right_i.position(rc_ket);
                                    if (rb_ket == right_i.size()) continue;
                                    uint32_t rs_ket = right_i.sizes[rb_ket];
                                    uint32_t s_in = phys_i.position(phys_in);
                                    MQ_REQUIRE(s_in < phys_i.size(), "operator charge not in physical index");
                                    uint32_t ket_offset = uint32_t(bra_right_pb(s_in, rb_ket));
                                    int64_t right_offset = right.offset(ci, b2);
                                    double scale = right.conjugate_scale(ci, b2) * mpo.terms[t].scale;
                                    double cpl[4];
                                    if (sy.su2()) {
                                        int Jp = sy.spin(rc_ket);
                                        const double* u = ccache.get(J, std::abs(J - Jp), Jp, A, K, Ap, I, std::abs(I - Ip), Ip);
                                        for (int k = 0; k < 4; ++k) cpl[k] = u[k] * scale;
                                    } else {
                                        cpl[0] = cpl[1] = cpl[2] = cpl[3] = scale;
                                    }
                                    op_iterate(W, wb, cpl, coh, s, rs_ket, blk, ket_offset, ci,
                                               uint32_t(right_offset / rs_ket));
                                }
                            }
                        }
                        coh.add_line(b1);
                    }
                }
                coh.finalize();
                if (coh.n_tasks()) blk.cohorts.push_back(std::move(coh));
            }
        } catch (const std::exception& e) {
            std::lock_guard<std::mutex> lk(emu);
            errors.push_back(e.what());
        }
    }
    if (!errors.empty()) fail(MQ_ERR_INVALID, errors[0]);
    Schedule S;
    S.kind = Schedule::RIGHT;
    S.aux_dim = mpo.row_dim;
    flatten(blocks, S);
    // new BoundaryIndex(ket_left_i, bra_left_i) (move_boundary.hpp:331-336); always complemented
    S.out_index = BoundaryIndex(left_i, left_i, mpo.row_dim);
    for (uint32_t c = 0; c < S.n_cohorts(); ++c) {
        std::vector<int64_t> off(S.coh_mpo_offsets.begin() + size_t(c) * S.aux_dim,
                                 S.coh_mpo_offsets.begin() + size_t(c + 1) * S.aux_dim);
        S.out_index.add_cohort(S.coh_lb[c], S.coh_rb[c], off);
    }
    S.out_index.complement_transpose(mpo.left_herm, false, sy);
    return S;
}

bool trace_enabled() {
    static const bool on = std::getenv("MQ_TRACE") != nullptr;
    return on;
}

void trace_phase(const char* name) {
    static thread_local std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    if (!trace_enabled()) return;
    const auto t = std::chrono::steady_clock::now();
    std::fprintf(stderr, "[mq] %-28s %9.3f ms\n", name, std::chrono::duration<double, std::milli>(t - t0).count());
    t0 = t;
}

// ============================================================================ index-only steps
// The BoundaryIndex that build_left_schedule / build_right_schedule produce, without building their
// task lists.  A line b of a cohort is present iff some sparse-operator entry reaches it: entries
// always find their T tuple, so presence does not depend on the entries' values.  Lookups go
// through dense (sector, sector) and (sector, physical sector) tables.
namespace {

struct Flow {
    uint32_t b, s_out, s_in;   // the other MPO index, physical sectors of the bra and ket side
};

constexpr uint32_t MAX_PHYS_PAIRS = 1024;   // (bra, ket) physical sector pairs of one site

// per MPO index (col b2 for the left step, row b1 for the right step) the distinct flows of
// non-empty operator blocks
std::vector<std::vector<Flow>> mpo_flows(const Mpo& mpo, const Index& phys_i, bool by_col) {
    const uint32_t n = by_col ? mpo.col_dim : mpo.row_dim, n_other = by_col ? mpo.row_dim : mpo.col_dim;
    const uint32_t np = phys_i.size();
    const auto& cells = by_col ? mpo.cols : mpo.rows;
    // operator -> its flows (so, si) of non-empty blocks, once
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> op_flows(mpo.ops.size());
    for (size_t o = 0; o < mpo.ops.size(); ++o)
        for (const OpBlock& wb : mpo.ops[o].blocks) {
            if (wb.e1 <= wb.e0) continue;
            uint32_t so = phys_i.position(wb.out), si = phys_i.position(wb.in);
            MQ_REQUIRE(so < np && si < np, "operator charge not in physical index");
            op_flows[o].push_back({so, si});
        }
    // first-seen order per b, duplicates found through a stamp per (other, so, si)
    std::vector<uint32_t> stamp(size_t(n_other) * np * np, 0);
    std::vector<std::vector<Flow>> out(n);
    for (uint32_t b = 0; b < n; ++b)
        for (const auto& cell : cells[b])
            for (uint32_t t = cell.t0; t < cell.t1; ++t)
                for (const auto& f : op_flows[mpo.terms[t].op]) {
                    uint32_t& st = stamp[(size_t(cell.other) * np + f.first) * np + f.second];
                    if (st == b + 1) continue;
                    st = b + 1;
                    out[b].push_back({cell.other, f.first, f.second});
                }
    return out;
}

// table[x * nphys + s] = position of charges[x] + sign * phys[s] in target (target.size() if absent)
std::vector<uint32_t> shift_table(const Index& from, const Index& phys, int sign, const Index& target) {
    std::vector<uint32_t> t(size_t(from.size()) * phys.size());
    for (uint32_t x = 0; x < from.size(); ++x)
        for (uint32_t s = 0; s < phys.size(); ++s)
            t[size_t(x) * phys.size() + s] =
                target.position(sign > 0 ? from.charges[x] + phys.charges[s] : from.charges[x] - phys.charges[s]);
    return t;
}

std::vector<uint32_t> cohort_table(const BoundaryIndex& B, const Index& bra, const Index& ket) {
    std::vector<uint32_t> t(size_t(bra.size()) * ket.size());
    for (uint32_t i = 0; i < bra.size(); ++i)
        for (uint32_t j = 0; j < ket.size(); ++j) t[size_t(i) * ket.size() + j] = B.cohort_index(bra.charges[i], ket.charges[j]);
    return t;
}

// The flows inverted: for every physical sector pair (so, si) and MPO index b of the old boundary,
// the new MPO indices whose flows start there (CSR over (so * np + si) * aux + b).
struct FlowEdges {
    uint32_t aux = 0;
    std::vector<uint32_t> off, to;
};
FlowEdges invert_flows(const std::vector<std::vector<Flow>>& flows, uint32_t np, uint32_t aux) {
    FlowEdges E;
    E.aux = aux;
    E.off.assign(size_t(np) * np * aux + 1, 0);
    for (const auto& fl : flows)
        for (const Flow& f : fl) ++E.off[(size_t(f.s_out) * np + f.s_in) * aux + f.b + 1];
    for (size_t i = 1; i < E.off.size(); ++i) E.off[i] += E.off[i - 1];
    E.to.resize(E.off.back());
    std::vector<uint32_t> fill(E.off.begin(), E.off.end() - 1);
    for (uint32_t nb = 0; nb < flows.size(); ++nb)
        for (const Flow& f : flows[nb]) E.to[fill[(size_t(f.s_out) * np + f.s_in) * aux + f.b]++] = nb;
    return E;
}

// the stored blocks (MPO indices) of every cohort of B, ascending
std::vector<std::vector<uint32_t>> present_blocks(const BoundaryIndex& B) {
    std::vector<std::vector<uint32_t>> out(B.n_cohorts());
#pragma omp parallel for schedule(static)
    for (int c = 0; c < int(B.n_cohorts()); ++c)
        for (uint32_t b = 0; b < B.aux_dim; ++b)
            if (B.offset(uint32_t(c), b) > -1) out[c].push_back(b);
    return out;
}

// lines of one (ket, bra) sector pair of the new boundary: ln[nb] = 1 for every new MPO index nb
// that a flow (b, so, si) reaches from a stored block b of the old cohort of (so, si); coh[so * np
// + si] is that cohort (>= n_cohorts when absent)
void mark_lines(uint8_t* ln, const uint32_t* coh, uint32_t n_cohorts, uint32_t np, const FlowEdges& E,
                const std::vector<std::vector<uint32_t>>& present) {
    for (uint32_t q = 0; q < np * np; ++q) {
        if (coh[q] >= n_cohorts) continue;
        const uint32_t* eo = &E.off[size_t(q) * E.aux];
        for (uint32_t b : present[coh[q]])
            for (uint32_t k = eo[b]; k < eo[b + 1]; ++k) ln[E.to[k]] = 1;
    }
}

}  // namespace

BoundaryIndex left_step_index(const SiteLayout& site, const Mpo& mpo, const BoundaryIndex& left, bool symmetric) {
    const Index &left_i = site.left_i, &right_i = site.right_i, &phys_i = site.phys_i;
    MQ_REQUIRE(left.aux_dim == mpo.row_dim, "left boundary aux_dim != mpo row_dim");
    require_bond(left, left_i, "left");
    const Symm sy = site.symm;
    const uint32_t nr = right_i.size(), np = phys_i.size(), w = mpo.col_dim;
    MQ_REQUIRE(np * np <= MAX_PHYS_PAIRS, "too many physical sectors");
    auto flows = mpo_flows(mpo, phys_i, true);
    auto down = shift_table(right_i, phys_i, -1, left_i);   // rc - phys -> left sector
    auto coh = cohort_table(left, left_i, left_i);
    const FlowEdges edges = invert_flows(flows, np, left.aux_dim);
    const auto present = present_blocks(left);
    trace_phase("left index: flows");
    // lines[(rb_ket * nr + rb_bra) * w + b2]
    std::vector<uint8_t> lines(size_t(nr) * nr * w, 0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int rbk = 0; rbk < int(nr); ++rbk) {
        Charge rc_ket = right_i.charges[rbk];
        for (uint32_t rbb = 0; rbb < nr; ++rbb) {
            Charge rc_bra = right_i.charges[rbb];
            if (std::abs(sy.particle_number(rc_bra) - sy.particle_number(rc_ket)) > 2) continue;
            uint8_t* ln = &lines[(size_t(rbk) * nr + rbb) * w];
            // the left cohort of every flow sector pair (so, si)
            uint32_t cq[MAX_PHYS_PAIRS];
            for (uint32_t so = 0; so < np; ++so)
                for (uint32_t si = 0; si < np; ++si) {
                    uint32_t lbb = down[size_t(rbb) * np + so], lbk = down[size_t(rbk) * np + si];
                    cq[so * np + si] = (lbb == left_i.size() || lbk == left_i.size()) ? left.n_cohorts()
                                                                                       : coh[size_t(lbb) * left_i.size() + lbk];
                }
            mark_lines(ln, cq, left.n_cohorts(), np, edges, present);
            for (uint32_t b2 = 0; b2 < w; ++b2)
                if ((symmetric && mpo.right_herm.skip(b2, rc_bra, rc_ket)) ||
                    (sy.su2() && !triangle(sy.spin(rc_ket), mpo.right_spin[b2], sy.spin(rc_bra))))
                    ln[b2] = 0;
        }
    }
    trace_phase("left index: lines");
    BoundaryIndex out(right_i, right_i, w);
    std::vector<int64_t> off(w);
    for (uint32_t rbk = 0; rbk < nr; ++rbk)
        for (uint32_t rbb = 0; rbb < nr; ++rbb) {
            const uint8_t* ln = &lines[(size_t(rbk) * nr + rbb) * w];
            uint64_t bs = uint64_t(right_i.sizes[rbb]) * right_i.sizes[rbk];
            uint32_t cnt = 0;
            for (uint32_t b = 0; b < w; ++b) off[b] = ln[b] ? int64_t(bs * cnt++) : -1;
            if (cnt) out.add_cohort(rbb, rbk, off);
        }
    trace_phase("left index: cohorts");
    if (symmetric) out.complement_transpose(mpo.right_herm, true, sy);
    trace_phase("left index: transposes");
    return out;
}

BoundaryIndex right_step_index(const SiteLayout& site, const Mpo& mpo, const BoundaryIndex& right, bool symmetric) {
    const Index &left_i = site.left_i, &right_i = site.right_i, &phys_i = site.phys_i;
    MQ_REQUIRE(right.aux_dim == mpo.col_dim, "right boundary aux_dim != mpo col_dim");
    require_bond(right, right_i, "right");
    const Symm sy = site.symm;
    const uint32_t nl = left_i.size(), np = phys_i.size(), w = mpo.row_dim;
    MQ_REQUIRE(np * np <= MAX_PHYS_PAIRS, "too many physical sectors");
    auto flows = mpo_flows(mpo, phys_i, false);
    auto up = shift_table(left_i, phys_i, +1, right_i);   // lc + phys -> right sector
    auto coh = cohort_table(right, right_i, right_i);     // (ket sector, bra sector) of the right boundary
    const FlowEdges edges = invert_flows(flows, np, right.aux_dim);
    const auto present = present_blocks(right);
    std::vector<uint8_t> lines(size_t(nl) * nl * w, 0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int lbk = 0; lbk < int(nl); ++lbk) {
        Charge lc_ket = left_i.charges[lbk];
        for (uint32_t lbb = 0; lbb < nl; ++lbb) {
            Charge lc_bra = left_i.charges[lbb];
            if (std::abs(sy.particle_number(lc_ket) - sy.particle_number(lc_bra)) > 2) continue;
            uint8_t* ln = &lines[(size_t(lbk) * nl + lbb) * w];
            // the right cohort of every flow sector pair (so, si)
            uint32_t cq[MAX_PHYS_PAIRS];
            for (uint32_t so = 0; so < np; ++so)
                for (uint32_t si = 0; si < np; ++si) {
                    uint32_t rbb = up[size_t(lbb) * np + so], rbk = up[size_t(lbk) * np + si];
                    cq[so * np + si] = (rbb == right_i.size() || rbk == right_i.size()) ? right.n_cohorts()
                                                                                         : coh[size_t(rbk) * right_i.size() + rbb];
                }
            mark_lines(ln, cq, right.n_cohorts(), np, edges, present);
            for (uint32_t b1 = 0; b1 < w; ++b1)
                if ((symmetric && mpo.left_herm.skip(b1, lc_ket, lc_bra)) ||
                    (sy.su2() && !triangle(sy.spin(lc_ket), mpo.left_spin[b1], sy.spin(lc_bra))))
                    ln[b1] = 0;
        }
    }
    BoundaryIndex out(left_i, left_i, w);
    std::vector<int64_t> off(w);
    for (uint32_t lbk = 0; lbk < nl; ++lbk)
        for (uint32_t lbb = 0; lbb < nl; ++lbb) {
            const uint8_t* ln = &lines[(size_t(lbk) * nl + lbb) * w];
            uint64_t bs = uint64_t(left_i.sizes[lbk]) * left_i.sizes[lbb];
            uint32_t cnt = 0;
            for (uint32_t b = 0; b < w; ++b) off[b] = ln[b] ? int64_t(bs * cnt++) : -1;
            if (cnt) out.add_cohort(lbk, lbb, off);
        }
    out.complement_transpose(mpo.left_herm, false, sy);
    return out;
}

// ============================================================================ flops
// MPSBlock::n_flops + Cohort::n_flops (tasks.cpp:388-398, 777-796).  t_index is the boundary
// the T stage multiplies with; lr_ket_sizes the row count of each ket block used in T.
uint64_t schedule_flops(const Schedule& S, const BoundaryIndex& t_index, const std::vector<uint32_t>& ket_rows) {
    uint64_t f = 0;
    for (uint32_t m = 0; m < S.n_blocks; ++m)
        for (uint64_t ti = S.t_begin[m]; ti < S.t_begin[m + 1]; ++ti) {
            uint32_t ci = S.t_ci[ti], cie = S.t_ci_eff[ti];
            if (S.kind == Schedule::LEFT)  // nb products (bls x ls_ket) * (ls_ket x rs_ket); ket_rows = rs of block m
                f += 2ull * t_index.left_size(ci) * t_index.right_size(ci) * t_index.n_blocks[cie] * ket_rows[m];
            else                           // (ls_ket x rs_ket) * (rs_ket x nb*rs_bra)
                f += 2ull * ket_rows[S.t_lb_ket[ti]] * t_index.cohort_size(cie);
        }
    for (uint32_t c = 0; c < S.n_cohorts(); ++c) {
        uint64_t ls = S.coh_ls[c], rs = S.coh_rs[c];
        f += 2ull * rs * ls * S.coh_nSrows[c] * S.coh_stripe[c];
        for (uint64_t u = S.coh_unit_begin[c]; u < S.coh_unit_begin[c + 1]; ++u) {
            uint64_t na = S.unit_task_begin[u + 1] - S.unit_task_begin[u];
            uint64_t slab = (S.kind == Schedule::LEFT) ? rs : ls;
            f += 2ull * slab * S.unit_ms[u] * na;
        }
    }
    return f;
}

// ============================================================================ export / import
void Schedule::export_desc(mq_schedule_desc* d) const {
    d->n_blocks = n_blocks;
    d->t_begin = t_begin.data();
    d->t_mps_offset = t_mps_offset.data();
    d->t_ci = t_ci.data();
    d->t_ci_eff = t_ci_eff.data();
    d->t_lb_ket = t_lb_ket.data();
    d->coh_begin = coh_begin.data();
    d->coh_lb = coh_lb.data();
    d->coh_rb = coh_rb.data();
    d->coh_ls = coh_ls.data();
    d->coh_rs = coh_rs.data();
    d->coh_ci = coh_ci.data();
    d->coh_ci_eff = coh_ci_eff.data();
    d->coh_nSrows = coh_nSrows.data();
    d->coh_stripe = coh_stripe.data();
    d->coh_unit_begin = coh_unit_begin.data();
    d->unit_offset = unit_offset.data();
    d->unit_ms = unit_ms.data();
    d->unit_line_begin = unit_line_begin.data();
    d->line_b1 = line_b1.data();
    d->line_count = line_count.data();
    d->unit_task_begin = unit_task_begin.data();
    d->task_ti = task_ti.data();
    d->task_col = task_col.data();
    d->task_alpha = task_alpha.data();
    d->aux_dim = aux_dim;
    d->coh_mpo_offsets = coh_mpo_offsets.empty() ? nullptr : coh_mpo_offsets.data();
}

void Schedule::import_desc(const mq_schedule_desc& d, uint32_t nc, uint32_t nu, uint64_t nl, uint64_t nt) {
    n_blocks = d.n_blocks;
    uint64_t ntt = d.t_begin[d.n_blocks];
    t_begin.assign(d.t_begin, d.t_begin + d.n_blocks + 1);
    t_mps_offset.assign(d.t_mps_offset, d.t_mps_offset + ntt);
    t_ci.assign(d.t_ci, d.t_ci + ntt);
    t_ci_eff.assign(d.t_ci_eff, d.t_ci_eff + ntt);
    t_lb_ket.assign(d.t_lb_ket, d.t_lb_ket + ntt);
    coh_begin.assign(d.coh_begin, d.coh_begin + d.n_blocks + 1);
    coh_lb.assign(d.coh_lb, d.coh_lb + nc);
    coh_rb.assign(d.coh_rb, d.coh_rb + nc);
    coh_ls.assign(d.coh_ls, d.coh_ls + nc);
    coh_rs.assign(d.coh_rs, d.coh_rs + nc);
    coh_ci.assign(d.coh_ci, d.coh_ci + nc);
    coh_ci_eff.assign(d.coh_ci_eff, d.coh_ci_eff + nc);
    coh_nSrows.assign(d.coh_nSrows, d.coh_nSrows + nc);
    coh_stripe.assign(d.coh_stripe, d.coh_stripe + nc);
    coh_unit_begin.assign(d.coh_unit_begin, d.coh_unit_begin + nc + 1);
    unit_offset.assign(d.unit_offset, d.unit_offset + nu);
    unit_ms.assign(d.unit_ms, d.unit_ms + nu);
    unit_line_begin.assign(d.unit_line_begin, d.unit_line_begin + nu + 1);
    line_b1.assign(d.line_b1, d.line_b1 + nl);
    line_count.assign(d.line_count, d.line_count + nl);
    unit_task_begin.assign(d.unit_task_begin, d.unit_task_begin + nu + 1);
    task_ti.assign(d.task_ti, d.task_ti + nt);
    task_col.assign(d.task_col, d.task_col + nt);
    task_alpha.assign(d.task_alpha, d.task_alpha + nt);
    aux_dim = d.aux_dim;
}

}  // namespace mq
