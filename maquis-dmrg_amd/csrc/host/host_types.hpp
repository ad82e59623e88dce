// Host-side object model of the sigma / boundary engine: charges, indices, MPO view,
// boundary index and the flattened schedule.  Everything here is plain C++17.
#pragma once

#include <algorithm>
#include <array>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "mq_sigma.h"

namespace mq {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
[[noreturn]] inline void fail(int code, const std::string& msg) { throw Error(code, msg); }
#define MQ_REQUIRE(cond, msg) do { if (!(cond)) ::mq::fail(MQ_ERR_INVALID, msg); } while (0)

inline uint64_t round_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

// ---------------------------------------------------------------------------- charges
// Charge arithmetic for the three groups.  U1 uses q[0] only; TwoU1 and SU2U1 use both.
// fuse is component-wise addition for all three (block_matrix/symmetry/su2u1.h:49-52,
// nu1_tpl.h); ordering is lexicographic and Index keeps descending order (indexing_stable.hpp:61-66).
struct Charge {
    int32_t a = 0, b = 0;
    bool operator==(const Charge& o) const { return a == o.a && b == o.b; }
    bool operator!=(const Charge& o) const { return !(*this == o); }
    bool operator<(const Charge& o) const { return a < o.a || (a == o.a && b < o.b); }
    bool operator>(const Charge& o) const { return o < *this; }
    Charge operator+(const Charge& o) const { return {a + o.a, b + o.b}; }
    Charge operator-(const Charge& o) const { return {a - o.a, b - o.b}; }
    Charge operator-() const { return {-a, -b}; }
    uint64_t key() const { return (uint64_t(uint32_t(a)) << 32) | uint32_t(b); }
};
inline Charge from_c(const mq_charge& c) { return {c.q[0], c.q[1]}; }

struct Symm {
    int kind = MQ_SYM_U1;
    bool su2() const { return kind == MQ_SYM_SU2U1; }
    // SymmGroup::particleNumber: U1 -> q (u1.h:49); TwoU1 -> sum (nu1_tpl.h:277); SU2U1 -> N (su2u1.h:43)
    int particle_number(const Charge& c) const {
        if (kind == MQ_SYM_TWOU1) return c.a + c.b;
        return c.a;
    }
    int spin(const Charge& c) const { return su2() ? c.b : 0; }
};

// ---------------------------------------------------------------------------- Index
class Index {
public:
    std::vector<Charge> charges;
    std::vector<uint32_t> sizes;
    Index() = default;
    explicit Index(const mq_index_desc& d) {
        MQ_REQUIRE(d.n == 0 || (d.charges && d.sizes), "index descriptor with null arrays");
        charges.resize(d.n);
        sizes.assign(d.sizes, d.sizes + d.n);
        for (uint32_t i = 0; i < d.n; ++i) charges[i] = from_c(d.charges[i]);
        build();
    }
    void build() {
        pos_.clear();
        pos_.reserve(charges.size() * 2 + 1);
        for (uint32_t i = 0; i < charges.size(); ++i) pos_[charges[i].key()] = i;
    }
    uint32_t size() const { return uint32_t(charges.size()); }
    // position of charge c, size() if absent
    uint32_t position(const Charge& c) const {
        auto it = pos_.find(c.key());
        return it == pos_.end() ? size() : it->second;
    }
    std::vector<mq_charge> c_charges() const {
        std::vector<mq_charge> r(charges.size());
        for (size_t i = 0; i < charges.size(); ++i) r[i].q[0] = charges[i].a, r[i].q[1] = charges[i].b;
        return r;
    }
private:
    std::unordered_map<uint64_t, uint32_t> pos_;
};

// ProductBasis(a, b, f) (indexing_stable.hpp:310-378): offset of (a-sector, b-sector) inside the
// fused sector f(a, b), enumerating a outer, b inner.
class ProductBasis {
public:
    // sign = -1 gives fuse(-a, b) (right-paired columns), +1 gives fuse(a, b) (left-paired rows)
    ProductBasis(const Index& a, const Index& b, int sign) : na_(a.size()), nb_(b.size()) {
        off_.assign(size_t(na_) * nb_, 0);
        std::unordered_map<uint64_t, uint64_t> acc;
        for (uint32_t i = 0; i < na_; ++i)
            for (uint32_t j = 0; j < nb_; ++j) {
                Charge pc = sign < 0 ? b.charges[j] - a.charges[i] : a.charges[i] + b.charges[j];
                uint64_t& s = acc[pc.key()];
                off_[size_t(i) * nb_ + j] = s;
                s += uint64_t(a.sizes[i]) * b.sizes[j];
            }
        for (auto& kv : acc) total_[kv.first] = kv.second;
    }
    uint64_t operator()(uint32_t ia, uint32_t ib) const { return off_[size_t(ia) * nb_ + ib]; }
    uint64_t size(const Charge& pc) const {
        auto it = total_.find(pc.key());
        return it == total_.end() ? 0 : it->second;
    }
private:
    uint32_t na_, nb_;
    std::vector<uint64_t> off_;
    std::unordered_map<uint64_t, uint64_t> total_;
};

// ---------------------------------------------------------------------------- site layouts
// Right-paired block lb: ls(lb) x cols(lb); left-paired block rb: rows(rb) x rs(rb).
struct SiteLayout {
    Symm symm;
    Index left_i, phys_i, right_i;
    std::vector<uint32_t> rp_cols;       // per left sector
    std::vector<uint64_t> rp_offset;     // padded Davidson-vector offsets
    uint64_t rp_total = 0;
    std::vector<uint32_t> lp_rows;       // per right sector
    std::vector<uint64_t> lp_offset;
    uint64_t lp_total = 0;
    explicit SiteLayout(const mq_site_desc& d);
};

// ---------------------------------------------------------------------------- MPO
struct OpBlock {
    Charge in, out;
    uint32_t e0, e1;
};
struct Op {
    int spin = 0;
    std::vector<OpBlock> blocks;
    std::vector<uint32_t> row, col;
    std::vector<uint8_t> row_spin, col_spin;
    std::vector<double> coeff;
};
struct TermRef { uint32_t op; double scale; };

// Hermitian pairing of bond indices (mp_tensors/mpotensor_detail.h:141-187).
struct Hermitian {
    std::vector<uint32_t> herm;
    std::vector<int> phase;
    // skip(b, l, r): l < r and b has a registered conjugate (mpotensor_detail.h:147-157)
    bool skip(uint32_t b, const Charge& l, const Charge& r) const {
        return l < r && herm[b] != MQ_NONE;
    }
    uint32_t conj(uint32_t b) const { return herm[b]; }
    uint32_t size() const { return uint32_t(herm.size()); }
};

class Mpo {
public:
    uint32_t row_dim = 0, col_dim = 0;
    std::vector<int> left_spin, right_spin;
    Hermitian left_herm, right_herm;
    std::vector<Op> ops;
    // row(b1) -> list of (b2, first term, last term) and column(b2) -> (b1, ...)
    struct Cell { uint32_t other; uint32_t t0, t1; };
    std::vector<std::vector<Cell>> rows, cols;
    std::vector<TermRef> terms;   // sorted by (b1, b2)
    explicit Mpo(const mq_mpo_desc& d);
};

// ---------------------------------------------------------------------------- BoundaryIndex
class BoundaryIndex {
public:
    Index bra, ket;
    uint32_t aux_dim = 0;
    std::vector<uint32_t> bra_pos, ket_pos, n_blocks;
    std::vector<uint8_t> tr;
    std::vector<int64_t> offsets;     // n_cohorts * aux_dim
    std::vector<double> conj_scales;
    std::vector<uint8_t> transposes;

    BoundaryIndex() = default;
    BoundaryIndex(const Index& b, const Index& k, uint32_t aux) : bra(b), ket(k), aux_dim(aux) {}
    explicit BoundaryIndex(const mq_bindex_desc& d);

    uint32_t n_cohorts() const { return uint32_t(bra_pos.size()); }
    int64_t offset(uint32_t ci, uint32_t b) const { return offsets[size_t(ci) * aux_dim + b]; }
    bool has_block(uint32_t ci, uint32_t b) const { return ci < n_cohorts() && offset(ci, b) > -1; }
    double conjugate_scale(uint32_t ci, uint32_t b) const { return conj_scales[size_t(ci) * aux_dim + b]; }
    uint64_t left_size(uint32_t ci) const { return bra.sizes[bra_pos[ci]]; }
    uint64_t right_size(uint32_t ci) const { return ket.sizes[ket_pos[ci]]; }
    uint64_t block_size(uint32_t ci) const { return left_size(ci) * right_size(ci); }
    uint64_t cohort_size(uint32_t ci) const { return uint64_t(n_blocks[ci]) * block_size(ci); }
    uint32_t cohort_index(const Charge& lc, const Charge& rc) const {
        uint32_t lb = bra.position(lc), rb = ket.position(rc);
        if (lb == bra.size() || rb == ket.size()) return n_cohorts();
        auto it = lbrb_.find(uint64_t(lb) << 32 | rb);
        return it == lbrb_.end() ? n_cohorts() : it->second;
    }
    uint32_t add_cohort(uint32_t lb, uint32_t rb, const std::vector<int64_t>& off);
    void complement_transpose(const Hermitian& herm, bool forward, const Symm& sy);
    void rebuild_lookup();
    void fill_desc(mq_bindex_desc* d, std::vector<mq_charge>& bra_c, std::vector<mq_charge>& ket_c) const;
private:
    std::unordered_map<uint64_t, uint32_t> lbrb_;
};

// SU2 couplings (block_matrix/symmetry/gsl_coupling.h): 9j-based factors for the four
// (row-spin, col-spin) cases of a sparse operator entry.
double wigner9j(int a, int b, int c, int d, int e, int f, int g, int h, int i);
void su2_set_coupling(int a, int b, int c, int d, int e, int f, int g, int h, int i,
                      double init, double couplings[4]);

// Memoised unit couplings keyed on all nine arguments (the reference caches per (J, I, I'),
// Wigner9jCacheI in gsl_coupling.h:129-204; the values are the same).
class CouplingCache {
public:
    const double* get(int a, int b, int c, int d, int e, int f, int g, int h, int i) {
        uint64_t key = 0;
        int v[9] = {a, b, c, d, e, f, g, h, i};
        for (int k = 0; k < 9; ++k) key = key * 131 + uint64_t(v[k] + 3);
        auto it = map_.find(key);
        if (it != map_.end()) return it->second.data();
        std::array<double, 4> r;
        su2_set_coupling(a, b, c, d, e, f, g, h, i, 1.0, r.data());
        return map_.emplace(key, r).first->second.data();
    }
private:
    std::unordered_map<uint64_t, std::array<double, 4>> map_;
};


// ---------------------------------------------------------------------------- Schedule
// Flattened ScheduleNew: per MPS block the T tuples and cohorts; per cohort its units;
// per unit its lines and (ti, col, alpha) tasks.
struct Schedule {
    enum Kind { SIGMA = 0, LEFT = 1, RIGHT = 2 } kind = SIGMA;
    uint32_t n_blocks = 0;
    std::vector<uint64_t> t_begin{0};
    std::vector<uint32_t> t_mps_offset, t_ci, t_ci_eff, t_lb_ket;
    std::vector<uint64_t> t_sza;
    std::vector<uint64_t> coh_begin{0};
    std::vector<uint32_t> coh_lb, coh_rb, coh_ls, coh_rs, coh_ci, coh_ci_eff, coh_nSrows, coh_stripe;
    std::vector<uint64_t> coh_unit_begin{0};
    std::vector<uint32_t> unit_offset, unit_ms;
    std::vector<uint64_t> unit_line_begin{0};
    std::vector<uint32_t> line_b1, line_count;
    std::vector<uint64_t> unit_task_begin{0};
    std::vector<uint32_t> task_ti, task_col;
    std::vector<double> task_alpha;
    uint32_t aux_dim = 0;
    std::vector<int64_t> coh_mpo_offsets;     // boundary steps: n_cohorts * aux_dim
    BoundaryIndex out_index;                  // boundary steps: index of the new boundary
    uint64_t flops = 0;

    uint32_t n_cohorts() const { return uint32_t(coh_lb.size()); }
    void export_desc(mq_schedule_desc* d) const;
    void import_desc(const mq_schedule_desc& d, uint32_t n_cohorts, uint32_t n_units, uint64_t n_lines,
                     uint64_t n_tasks);
};

// Boundary indices must describe the bond they are contracted with: both indices of a left
// boundary equal the site's left index, both of a right boundary its right index.
inline bool same_index(const Index& a, const Index& b) { return a.charges == b.charges && a.sizes == b.sizes; }
inline void require_bond(const BoundaryIndex& B, const Index& bond, const char* what) {
    if (!same_index(B.bra, bond) || !same_index(B.ket, bond))
        fail(MQ_ERR_INVALID, std::string(what) + " boundary index does not match the site's bond index");
}

// Builders: reference create_contraction_schedule / left / right schedules.
Schedule build_sigma_schedule(const SiteLayout& site, const Mpo& mpo, const BoundaryIndex& left,
                              const BoundaryIndex& right);
Schedule build_left_schedule(const SiteLayout& site, const Mpo& mpo, const BoundaryIndex& left,
                             bool symmetric);
Schedule build_right_schedule(const SiteLayout& site, const Mpo& mpo, const BoundaryIndex& right,
                              bool symmetric);
// The boundary index those two step schedules produce, without their task lists.
// Host-side phase timings of plan construction, printed to stderr when MQ_TRACE is set in the
// environment: trace_phase(name) reports the time since the previous call on this thread.
bool trace_enabled();
void trace_phase(const char* name);

BoundaryIndex left_step_index(const SiteLayout& site, const Mpo& mpo, const BoundaryIndex& left, bool symmetric);
BoundaryIndex right_step_index(const SiteLayout& site, const Mpo& mpo, const BoundaryIndex& right, bool symmetric);
uint64_t schedule_flops(const Schedule& s, const BoundaryIndex& t_index, const std::vector<uint32_t>& lr_ket_sizes);

}  // namespace mq
