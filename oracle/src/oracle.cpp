// ORACLE -- test infrastructure only.
//
// CPU restatement of the reference's sigma-vector / boundary-propagation path of QCMaquis
// (sekelle/maquis-dmrg), used as the parity checker of the gfx950 engine and as the timed CPU
// baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
// It shares nothing with the product except the plain descriptor structs of include/mq_sigma.h.
//
// Every function names the reference code it restates (paths relative to
// dmrg/framework/dmrg/ of the reference):
//   Cohort / SUnit / MPSBlock / ScheduleNew        solver/tasks.h:50-348, solver/tasks.cpp:72-1019
//   shtm_tasks, rshtm_t_tasks, rshtm_tasks          mp_tensors/contractions/create_schedule/*.hpp
//   lshtm_t_tasks, lshtm_tasks, op_iterate
//   BoundaryIndex::add_cohort/complement_transpose  mp_tensors/boundary.h:138-191
//   overlap_mpo_left_step / right_step (CPU branch) mp_tensors/contractions/boundary_ops/move_boundary.hpp:147-404
//   super_hamil_mv (CPU branch)                     solver/super_hamil_mv.hpp:121-132
//   jacobi_davidson::calculate_eigenvalue           ietl/jacobi.h:386-474
//   SU2::set_coupling / Wigner9jCacheI              block_matrix/symmetry/gsl_coupling.h:72-204
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <functional>
#include <vector>

#include "mq_sigma.h"

namespace oracle {

static thread_local std::string last_error;
static const unsigned NPOS = 0xffffffffu;

// ------------------------------------------------------------------ dense kernels
// Column-major C = alpha*op(A)*op(B) + beta*C (BLAS dgemm semantics).  An external Fortran
// dgemm (ILP64) may be installed with mqo_set_dgemm; otherwise a cache-blocked loop is used.
typedef void (*dgemm64_t)(const char*, const char*, const int64_t*, const int64_t*, const int64_t*, const double*,
                          const double*, const int64_t*, const double*, const int64_t*, const double*, double*,
                          const int64_t*, size_t, size_t);
static dgemm64_t ext_dgemm = nullptr;

static void gemm(char ta, char tb, int M, int N, int K, double alpha, const double* A, int lda, const double* B,
                 int ldb, double beta, double* C, int ldc) {
    if (M <= 0 || N <= 0) return;
    if (ext_dgemm) {
        int64_t m = M, n = N, k = K, la = lda, lb = ldb, lc = ldc;
        ext_dgemm(&ta, &tb, &m, &n, &k, &alpha, A, &la, B, &lb, &beta, C, &lc, 1, 1);
        return;
    }
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < M; ++i) C[i + size_t(j) * ldc] *= beta;
    if (K <= 0 || alpha == 0.0) return;
    const bool tA = (ta == 'T'), tB = (tb == 'T');
    const int KB = 128;
    for (int k0 = 0; k0 < K; k0 += KB) {
        int k1 = std::min(K, k0 + KB);
        for (int j = 0; j < N; ++j) {
            double* c = C + size_t(j) * ldc;
            for (int k = k0; k < k1; ++k) {
                double b = alpha * (tB ? B[j + size_t(k) * ldb] : B[k + size_t(j) * ldb]);
                if (b == 0.0) continue;
                if (!tA) {
                    const double* a = A + size_t(k) * lda;
                    for (int i = 0; i < M; ++i) c[i] += a[i] * b;
                } else {
                    for (int i = 0; i < M; ++i) c[i] += A[k + size_t(i) * lda] * b;
                }
            }
        }
    }
}

static inline void axpy(size_t n, double a, const double* x, double* y) {
    for (size_t i = 0; i < n; ++i) y[i] += a * x[i];
}

static inline size_t round_up128(size_t x) { return (x + 127) / 128 * 128; }

// ------------------------------------------------------------------ charges and indices
struct Chg {
    int a, b;
    bool operator==(const Chg& o) const { return a == o.a && b == o.b; }
    bool operator!=(const Chg& o) const { return !(*this == o); }
    bool operator<(const Chg& o) const { return std::tie(a, b) < std::tie(o.a, o.b); }
};
static inline Chg fuse(Chg x, Chg y) { return {x.a + y.a, x.b + y.b}; }
static inline Chg neg(Chg x) { return {-x.a, -x.b}; }

struct Idx {
    std::vector<std::pair<Chg, size_t>> v;
    Idx() {}
    Idx(const mq_index_desc& d) {
        for (uint32_t i = 0; i < d.n; ++i) v.push_back({{d.charges[i].q[0], d.charges[i].q[1]}, d.sizes[i]});
    }
    size_t size() const { return v.size(); }
    // Index::position (indexing_stable.hpp:129-141), linear search in the stored order
    size_t position(Chg c) const {
        for (size_t i = 0; i < v.size(); ++i)
            if (v[i].first == c) return i;
        return v.size();
    }
    std::vector<size_t> sizes() const {
        std::vector<size_t> r;
        for (auto& p : v) r.push_back(p.second);
        return r;
    }
};

// ProductBasis(a, b, f) (indexing_stable.hpp:310-378)
struct PB {
    std::map<std::pair<Chg, Chg>, size_t> kv;
    std::map<Chg, size_t> sz;
    PB(const Idx& A, const Idx& B, bool neg_first) {
        for (auto& x : A.v)
            for (auto& y : B.v) {
                Chg pc = neg_first ? fuse(neg(x.first), y.first) : fuse(x.first, y.first);
                kv[{x.first, y.first}] = sz[pc];
                sz[pc] += x.second * y.second;
            }
    }
    size_t operator()(Chg a, Chg b) const { return kv.at({a, b}); }
};

struct Sym {
    int kind;
    bool su2() const { return kind == MQ_SYM_SU2U1; }
    int spin(Chg c) const { return su2() ? c.b : 0; }
    int particleNumber(Chg c) const { return kind == MQ_SYM_TWOU1 ? c.a + c.b : c.a; }
};

// ------------------------------------------------------------------ SU2 couplings
// 9j from the 3j-free Racah/6j route is avoided on purpose: here the 9j is evaluated as a
// sum over products of three 6j symbols, each 6j by the Racah single-sum formula in exact
// integer-ratio arithmetic carried in long double.
static long double lfact(int n) {
    long double r = 1;
    for (int i = 2; i <= n; ++i) r *= i;
    return r;
}
static bool tri(int a, int b, int c) { return c >= std::abs(a - b) && c <= a + b && (a + b + c) % 2 == 0; }
static long double sixj(int a, int b, int c, int d, int e, int f) {  // {a b c; d e f}, doubled spins
    if (!tri(a, b, c) || !tri(a, e, f) || !tri(d, b, f) || !tri(d, e, c)) return 0;
    auto D = [](int x, int y, int z) {
        return std::sqrt(lfact((x + y - z) / 2) * lfact((x - y + z) / 2) * lfact((-x + y + z) / 2) /
                         lfact((x + y + z) / 2 + 1));
    };
    int t1 = (a + b + c) / 2, t2 = (a + e + f) / 2, t3 = (d + b + f) / 2, t4 = (d + e + c) / 2;
    int u1 = (a + b + d + e) / 2, u2 = (b + c + e + f) / 2, u3 = (a + c + d + f) / 2;
    long double s = 0;
    for (int z = std::max({t1, t2, t3, t4}); z <= std::min({u1, u2, u3}); ++z)
        s += ((z & 1) ? -1.0L : 1.0L) * lfact(z + 1) /
             (lfact(z - t1) * lfact(z - t2) * lfact(z - t3) * lfact(z - t4) * lfact(u1 - z) * lfact(u2 - z) *
              lfact(u3 - z));
    return s * D(a, b, c) * D(a, e, f) * D(d, b, f) * D(d, e, c);
}
double ninej(int a, int b, int c, int d, int e, int f, int g, int h, int i) {
    if (!tri(a, b, c) || !tri(d, e, f) || !tri(g, h, i) || !tri(a, d, g) || !tri(b, e, h) || !tri(c, f, i)) return 0;
    long double s = 0;
    int lo = std::max({std::abs(a - i), std::abs(d - h), std::abs(b - f)});
    int hi = std::min({a + i, d + h, b + f});
    for (int x = lo; x <= hi; x += 2)
        s += ((x & 1) ? -1.0L : 1.0L) * (x + 1) * sixj(a, b, c, f, i, x) * sixj(d, e, f, b, x, h) *
             sixj(g, h, i, x, a, d);
    return double(s);
}
static bool triangle(int a, int b, int c) { return std::abs(a - b) <= c && c <= a + b; }
static double mod_coupling(int a, int b, int c, int d, int e, int f, int g, int h, int i) {
    return std::sqrt((g + 1.) * (h + 1.) * (c + 1.) * (f + 1.)) * ninej(a, b, c, d, e, f, g, h, i);
}
// SU2::set_coupling (gsl_coupling.h:72-92)
static void set_coupling(int a, int b, int c, int d, int e, int f, int g, int h, int i, double init, double cp[4]) {
    cp[0] = cp[1] = cp[2] = cp[3] = 0;
    double pre = std::sqrt((i + 1.) * (a + 1.) / ((g + 1.) * (c + 1.))) * init;
    if (triangle(a, b, c)) {
        cp[0] = pre * mod_coupling(a, b, c, d, e, f, g, h, i);
        cp[2] = pre * mod_coupling(a, b, c, d, e, f, g, 2, i);
    }
    if (triangle(a, 2, c)) {
        cp[1] = pre * mod_coupling(a, 2, c, d, e, f, g, h, i);
        cp[3] = pre * mod_coupling(a, 2, c, d, e, f, g, 2, i);
    }
}
// Wigner9jCacheI (gsl_coupling.h:129-204): table over J' in [J-2, J+2] and 12 (A, K, A') cases.
struct W9j {
    int Jpmin = 0;
    std::vector<double> coeff;
    bool active = false;
    W9j(const Sym& s, Chg lc, Chg mc, Chg rc) {
        if (!s.su2()) return;
        active = true;
        int J = s.spin(mc), I = s.spin(lc), Ip = s.spin(rc);
        int two_sp = std::abs(I - Ip);
        int Jpmax = J + 2;
        Jpmin = std::max(J - 2, 0);
        coeff.assign(48 * (Jpmax - Jpmin + 1), 0.0);
        static const int K[12] = {0, 1, 2, 0, 0, 1, 2, 1, 0, 1, 2, 0};
        static const int Ap[12] = {0, 1, 2, 0, 1, 0, 1, 2, 2, 1, 0, 0};
        for (int Jp = Jpmin; Jp <= Jpmax; ++Jp)
            for (int q = 0; q < 12; ++q)
                set_coupling(J, std::abs(J - Jp), Jp, q / 4, K[q], Ap[q], I, two_sp, Ip, 1.0,
                             &coeff[4 * q + 48 * (Jp - Jpmin)]);
    }
    static int hash(int a, int b, int c) { return 4 * a + b + ((a == 1 && b == 1) ? c : 0); }
    void set_scale(int A, int K, int Ap, int Jp, double scale, double cp[4]) const {
        if (!active) {
            cp[0] = cp[1] = cp[2] = cp[3] = scale;
            return;
        }
        const double* c = &coeff[4 * hash(A, K, Ap) + 48 * (Jp - Jpmin)];
        for (int k = 0; k < 4; ++k) cp[k] = c[k] * scale;
    }
};

// ------------------------------------------------------------------ MPO
struct SparseEntry { unsigned row, col, rs, cs; double coeff; };
struct OpB { Chg in, out; std::vector<SparseEntry> e; };
struct Oper { int spin; std::vector<OpB> blocks; };
struct Herm {
    std::vector<unsigned> H;
    std::vector<int> P;
    bool skip(unsigned b, Chg l, Chg r) const { return (l < r) && H[b] != NPOS; }
};
struct MPO {
    unsigned rows, cols;
    std::vector<int> lspin, rspin;
    Herm lh, rh;
    std::vector<Oper> ops;
    // at(b1, b2): list of (op, scale)
    std::map<std::pair<unsigned, unsigned>, std::vector<std::pair<unsigned, double>>> at;
    std::vector<std::vector<unsigned>> row_b2, col_b1;
    MPO(const mq_mpo_desc& d) : rows(d.row_dim), cols(d.col_dim) {
        lspin.assign(rows, 0);
        rspin.assign(cols, 0);
        if (d.left_spin) lspin.assign(d.left_spin, d.left_spin + rows);
        if (d.right_spin) rspin.assign(d.right_spin, d.right_spin + cols);
        lh.H.assign(rows, NPOS);
        lh.P.assign(rows, 1);
        rh.H.assign(cols, NPOS);
        rh.P.assign(cols, 1);
        if (d.left_herm) lh.H.assign(d.left_herm, d.left_herm + rows);
        if (d.left_phase) lh.P.assign(d.left_phase, d.left_phase + rows);
        if (d.right_herm) rh.H.assign(d.right_herm, d.right_herm + cols);
        if (d.right_phase) rh.P.assign(d.right_phase, d.right_phase + cols);
        for (unsigned k = 0; k < d.n_ops; ++k) {
            const mq_op_desc& o = d.ops[k];
            Oper op;
            op.spin = o.spin;
            for (unsigned w = 0; w < o.n_blocks; ++w) {
                OpB b;
                b.in = {o.in_charge[w].q[0], o.in_charge[w].q[1]};
                b.out = {o.out_charge[w].q[0], o.out_charge[w].q[1]};
                for (unsigned e = o.entry_begin[w]; e < o.entry_begin[w + 1]; ++e)
                    b.e.push_back({o.row[e], o.col[e], o.row_spin ? o.row_spin[e] : 0u, o.col_spin ? o.col_spin[e] : 0u,
                                   o.coeff[e]});
                op.blocks.push_back(b);
            }
            ops.push_back(op);
        }
        row_b2.assign(rows, {});
        col_b1.assign(cols, {});
        for (unsigned t = 0; t < d.n_terms; ++t) {
            auto key = std::make_pair(d.term_b1[t], d.term_b2[t]);
            if (!at.count(key)) {
                row_b2[key.first].push_back(key.second);
            }
            at[key].push_back({d.term_op[t], d.term_scale[t]});
        }
        for (unsigned b1 = 0; b1 < rows; ++b1)
            for (unsigned b2 : row_b2[b1]) col_b1[b2].push_back(b1);
        for (auto& c : col_b1) std::sort(c.begin(), c.end());
    }
};

// ------------------------------------------------------------------ BoundaryIndex
struct BIndex {
    Idx bra, ket;
    std::map<std::pair<size_t, size_t>, unsigned> lbrb;
    std::vector<std::vector<long>> offsets;
    std::vector<std::vector<double>> cscale;
    std::vector<std::vector<char>> trans;
    std::vector<char> tr_;
    std::vector<size_t> lsz, rsz;
    std::vector<unsigned> nbs;
    unsigned aux = 0;
    BIndex() {}
    BIndex(const Idx& b, const Idx& k, unsigned a) : bra(b), ket(k), aux(a) {}
    BIndex(const mq_bindex_desc& d) : bra(d.bra), ket(d.ket), aux(d.aux_dim) {
        for (unsigned ci = 0; ci < d.n_cohorts; ++ci) {
            lbrb[{d.bra_pos[ci], d.ket_pos[ci]}] = ci;
            offsets.push_back(std::vector<long>(d.offsets + size_t(ci) * aux, d.offsets + size_t(ci + 1) * aux));
            cscale.push_back(std::vector<double>(d.conj_scales + size_t(ci) * aux, d.conj_scales + size_t(ci + 1) * aux));
            trans.push_back(d.transposes ? std::vector<char>(d.transposes + size_t(ci) * aux,
                                                             d.transposes + size_t(ci + 1) * aux)
                                         : std::vector<char>(aux, 0));
            tr_.push_back(d.tr[ci]);
            lsz.push_back(bra.v[d.bra_pos[ci]].second);
            rsz.push_back(ket.v[d.ket_pos[ci]].second);
            nbs.push_back(d.n_blocks[ci]);
        }
    }
    unsigned n_cohorts() const { return unsigned(offsets.size()); }
    long offset(unsigned ci, unsigned b) const { return offsets[ci][b]; }
    bool has_block(unsigned ci, unsigned b) const { return ci < n_cohorts() && offsets[ci][b] > -1; }
    double conjugate_scale(unsigned ci, unsigned b) const { return cscale[ci][b]; }
    bool tr(unsigned ci) const { return tr_[ci]; }
    size_t left_size(unsigned ci) const { return lsz[ci]; }
    size_t right_size(unsigned ci) const { return rsz[ci]; }
    size_t n_blocks(unsigned ci) const { return nbs[ci]; }
    size_t cohort_size(unsigned ci) const { return nbs[ci] * lsz[ci] * rsz[ci]; }
    unsigned cohort_index(size_t lb, size_t rb) const {
        auto it = lbrb.find({lb, rb});
        return it == lbrb.end() ? n_cohorts() : it->second;
    }
    unsigned cohort_index(Chg lc, Chg rc) const { return cohort_index(bra.position(lc), ket.position(rc)); }
    // add_cohort (boundary.h:138-159)
    unsigned add_cohort(size_t lb, size_t rb, const std::vector<long>& off) {
        unsigned ci = n_cohorts();
        lbrb[{lb, rb}] = ci;
        offsets.push_back(off);
        cscale.push_back(std::vector<double>(off.size(), 1.0));
        trans.push_back(std::vector<char>(off.size(), 0));
        tr_.push_back(0);
        lsz.push_back(bra.v[lb].second);
        rsz.push_back(ket.v[rb].second);
        nbs.push_back(unsigned(std::count_if(off.begin(), off.end(), [](long o) { return o > -1; })));
        return ci;
    }
    // complement_transpose (boundary.h:161-191), positions interchangeable since bra == ket
    void complement_transpose(const Herm& herm, bool forward, const Sym& sy) {
        for (size_t rb = 0; rb < ket.size(); ++rb)
            for (size_t lb = 0; lb < bra.size(); ++lb) {
                unsigned ciA = cohort_index(lb, rb);
                if (ciA == n_cohorts()) continue;
                unsigned ciB = cohort_index(ket.v[rb].first, bra.v[lb].first);
                if (ciB == n_cohorts()) {
                    ciB = add_cohort(rb, lb, std::vector<long>(herm.H.size(), -1));
                    tr_[ciB] = 1;
                }
                for (unsigned b = 0; b < herm.H.size(); ++b)
                    if (herm.skip(b, ket.v[rb].first, bra.v[lb].first)) {
                        offsets[ciB][b] = offsets[ciA][herm.H[b]];
                        cscale[ciB][b] = conj_corr(sy, ket.v[rb].first, bra.v[lb].first) *
                                         double(herm.P[forward ? b : herm.H[b]]);
                        trans[ciB][b] = 1;
                    }
            }
    }
    static double conj_corr(const Sym& sy, Chg lc, Chg rc) {
        if (!sy.su2()) return 1.0;
        int S = std::min(sy.spin(rc), sy.spin(lc));
        switch (sy.spin(rc) - sy.spin(lc)) {
            case 0: return 1.;
            case 1: return -std::sqrt((S + 1.) / (S + 2.));
            case -1: return std::sqrt((S + 2.) / (S + 1.));
            case 2: return -std::sqrt((S + 1.) / (S + 3.));
            case -2: return -std::sqrt((S + 3.) / (S + 1.));
        }
        throw std::runtime_error("conjugate_correction: rank > 1");
    }
};

// ------------------------------------------------------------------ Cohort / MPSBlock
// solver/tasks.h:50-348 (host parts only)
struct Cohort {
    struct SUnit {
        unsigned offset = 0, ms = 0, b2count = 0;
        std::vector<unsigned> tidx, b2s, b1;
        std::vector<double> alpha;
        void push_back(double s, unsigned ti, unsigned col) {
            alpha.push_back(s);
            ++b2count;
            tidx.push_back(ti);
            tidx.push_back(col);
        }
        unsigned add_line(unsigned b) {
            unsigned r = b2count;
            if (b2count) {
                b2s.push_back(b2count);
                b1.push_back(b);
                b2count = 0;
            }
            return r;
        }
    };
    unsigned lb, rb, ls, rs, ci, ci_eff, sblock, nSrows, stripe = 0;
    std::vector<long> mpo_offsets;
    std::vector<unsigned> sfold;
    std::vector<SUnit> suv;
    Cohort(const std::vector<size_t>& phys, unsigned l, unsigned r, unsigned lsz, unsigned rsz, unsigned c,
           unsigned ce, unsigned mpodim, bool left = false)
        : lb(l), rb(r), ls(lsz), rs(rsz), ci(c), ci_eff(ce), nSrows(mpodim), mpo_offsets(mpodim, 0) {
        unsigned sum = 0;
        for (auto d : phys) sfold.push_back(sum), sum += unsigned(d);
        suv.resize(sum);
        sblock = left ? rs : ls;
    }
    void push_back(unsigned s, unsigned ss2, double scale, unsigned ti, unsigned col) {
        suv[sfold[s] + ss2].push_back(scale, ti, col);
    }
    void add_line(unsigned b) {
        for (auto& x : suv) mpo_offsets[b] += x.add_line(b);
    }
    void add_unit(unsigned s, unsigned ss, unsigned m, unsigned off) {
        for (unsigned i = 0; i < ss; ++i) {
            suv[sfold[s] + i].offset = i * m + off;
            suv[sfold[s] + i].ms = m;
            stripe += m;
        }
    }
    void finalize() {  // compute_mpo_offsets (tasks.cpp:506-526)
        std::vector<long> cpy = mpo_offsets;
        unsigned n = 0;
        for (auto& b : cpy)
            if (b) b = n++;
        nSrows = n;
        for (auto& x : suv)
            for (auto& b : x.b1) b = unsigned(cpy[b]);
        size_t bs = size_t(ls) * rs;
        unsigned cnt = 0;
        for (auto& b : mpo_offsets) b = b ? long(bs * cnt++) : -1;
    }
    size_t n_tasks() const {
        size_t n = 0;
        for (auto& x : suv) n += x.alpha.size();
        return n;
    }
    size_t S_size() const { return size_t(nSrows) * stripe * sblock; }

    // create_s (tasks.cpp:425-454): left layout, S is stripe x (nSrows*rs)
    std::vector<double> create_s(const std::vector<std::vector<double>>& T) const {
        std::vector<double> ret(S_size(), 0.0);
        for (auto& x : suv) {
            if (x.alpha.empty()) continue;
            std::vector<double> buf(size_t(x.ms) * rs);
            unsigned seek = 0;
            for (size_t b = 0; b < x.b2s.size(); ++b) {
                std::fill(buf.begin(), buf.end(), 0.0);
                for (unsigned ia = seek; ia < seek + x.b2s[b]; ++ia)
                    axpy(buf.size(), x.alpha[ia], &T[x.tidx[2 * ia]][size_t(x.tidx[2 * ia + 1]) * rs], buf.data());
                unsigned bb = x.b1[b];
                for (unsigned c = 0; c < rs; ++c)
                    std::copy(&buf[size_t(c) * x.ms], &buf[size_t(c) * x.ms] + x.ms,
                              ret.data() + size_t(stripe) * (size_t(bb) * rs + c) + x.offset);
                seek += x.b2s[b];
            }
        }
        return ret;
    }
    // create_s_r (tasks.cpp:466-494): right layout, S is (nSrows*ls) x stripe
    std::vector<double> create_s_r(const std::vector<std::vector<double>>& T) const {
        std::vector<double> ret(S_size(), 0.0);
        for (auto& x : suv) {
            if (x.alpha.empty()) continue;
            std::vector<double> buf(size_t(ls) * x.ms);
            unsigned seek = 0;
            for (size_t b = 0; b < x.b2s.size(); ++b) {
                std::fill(buf.begin(), buf.end(), 0.0);
                for (unsigned ia = seek; ia < seek + x.b2s[b]; ++ia)
                    axpy(buf.size(), x.alpha[ia], &T[x.tidx[2 * ia]][size_t(x.tidx[2 * ia + 1]) * ls], buf.data());
                unsigned bb = x.b1[b];
                for (unsigned c = 0; c < x.ms; ++c)
                    std::copy(&buf[size_t(c) * ls], &buf[size_t(c) * ls] + ls,
                              ret.data() + size_t(nSrows) * ls * (x.offset + c) + size_t(bb) * ls);
                seek += x.b2s[b];
            }
        }
        return ret;
    }
    // tr_tile_v (tasks.cpp:49-62)
    static void tr_tile_v(unsigned nrows, unsigned ncols, size_t cnt, const double* in, double* out) {
        std::vector<double> buf(size_t(nrows) * ncols);
        for (size_t b = 0; b < cnt; ++b) {
            size_t off = b * nrows * ncols;
            std::copy(in + off, in + off + size_t(nrows) * ncols, buf.data());
            for (unsigned i = 0; i < nrows; ++i)
                for (unsigned j = 0; j < ncols; ++j) out[off + size_t(ncols) * i + j] = buf[size_t(nrows) * j + i];
        }
    }
    // contract (tasks.cpp:286-313)
    void contract(const std::vector<const double*>& left, const std::vector<std::vector<double>>& T, double* out,
                  std::mutex& mu) const {
        std::vector<double> s = create_s_r(T);
        int M = rs, N = stripe, K = nSrows * ls;
        const double* luse = left[ci_eff];
        std::vector<double> lbuf;
        if (ci != ci_eff) {
            lbuf.resize(size_t(M) * K);
            tr_tile_v(ls, rs, nSrows, left[ci_eff], lbuf.data());
            luse = lbuf.data();
        }
        std::vector<double> buf(size_t(M) * N, 0.0);
        gemm('N', 'N', M, N, K, 1.0, luse, M, s.data(), K, 0.0, buf.data(), M);
        std::lock_guard<std::mutex> lk(mu);
        axpy(buf.size(), 1.0, buf.data(), out);
    }
    // prop_l (tasks.cpp:198-209)
    void prop_l(const double* bra, const std::vector<std::vector<double>>& T, double* out) const {
        std::vector<double> s = create_s(T);
        gemm('T', 'N', ls, nSrows * rs, stripe, 1.0, bra, stripe, s.data(), stripe, 0.0, out, ls);
    }
    // prop_r (tasks.cpp:243-258)
    void prop_r(const double* bra, const std::vector<std::vector<double>>& T, double* out) const {
        std::vector<double> s = create_s_r(T);
        int M = nSrows * ls, N = rs;
        std::vector<double> buf(size_t(M) * N, 0.0);
        gemm('N', 'T', M, N, stripe, 1.0, s.data(), M, bra, rs, 0.0, buf.data(), M);
        for (unsigned b = 0; b < nSrows; ++b)
            for (unsigned col = 0; col < rs; ++col)
                std::copy(&buf[size_t(ls) * b + size_t(col) * M], &buf[size_t(ls) * b + size_t(col) * M] + ls,
                          out + (size_t(b) * rs + col) * ls);
    }
    size_t n_flops() const {
        size_t r = 2ull * rs * ls * nSrows * stripe;
        for (auto& x : suv) r += 2ull * ls * x.ms * x.alpha.size();
        return r;
    }
};

struct MPSBlock {
    std::vector<std::tuple<unsigned, unsigned, unsigned, unsigned, size_t>> t_schedule;
    std::vector<Cohort> cohorts;
    unsigned get_ti(unsigned mps_offset, unsigned ci) const {  // tasks.cpp:766-775
        for (unsigned ti = 0; ti < t_schedule.size(); ++ti)
            if (std::get<0>(t_schedule[ti]) == mps_offset && std::get<1>(t_schedule[ti]) == ci) return ti;
        return NPOS;
    }
    // create_T (tasks.cpp:650-705)
    std::vector<std::vector<double>> create_T(const BIndex& right, const std::vector<const double*>& rdata,
                                              const std::vector<const double*>& mps,
                                              const std::vector<size_t>& lr_ket) const {
        std::vector<std::vector<double>> ret(t_schedule.size());
        for (unsigned ti = 0; ti < t_schedule.size(); ++ti) {
            unsigned off = std::get<0>(t_schedule[ti]), ci = std::get<1>(t_schedule[ti]),
                     cie = std::get<2>(t_schedule[ti]), lbk = std::get<3>(t_schedule[ti]);
            unsigned bls = unsigned(right.left_size(ci)), brs = unsigned(right.right_size(ci));
            int M = int(lr_ket[lbk]), K = bls;
            const double* mpsdata = mps[lbk] + size_t(M) * off;
            ret[ti].assign(size_t(M) * right.n_blocks(cie) * brs, 0.0);
            for (unsigned b = 0; b < right.n_blocks(cie); ++b) {
                int N = brs;
                size_t roff = size_t(b) * K * N, ooff = size_t(b) * M * N;
                if (ci != cie)
                    gemm('N', 'T', M, N, K, 1.0, mpsdata, M, rdata[cie] + roff, N, 0.0, ret[ti].data() + ooff, M);
                else
                    gemm('N', 'N', M, N, K, 1.0, mpsdata, M, rdata[cie] + roff, K, 0.0, ret[ti].data() + ooff, M);
            }
        }
        return ret;
    }
    // create_T_left (tasks.cpp:545-601)
    std::vector<std::vector<double>> create_T_left(const BIndex& left, const std::vector<const double*>& ldata,
                                                   const std::vector<const double*>& mps, size_t rs_ket) const {
        std::vector<std::vector<double>> ret(t_schedule.size());
        for (unsigned ti = 0; ti < t_schedule.size(); ++ti) {
            unsigned off = std::get<0>(t_schedule[ti]), ci = std::get<1>(t_schedule[ti]),
                     cie = std::get<2>(t_schedule[ti]), lbk = std::get<3>(t_schedule[ti]);
            int M = int(left.left_size(ci)), N = int(rs_ket), K = int(left.right_size(ci));
            const double* mpsdata = mps[lbk] + size_t(K) * off;
            size_t nb = left.n_blocks(cie);
            ret[ti].assign(size_t(M) * N * nb, 0.0);
            for (unsigned b = 0; b < nb; ++b) {
                size_t loff = size_t(b) * M * K, ooff = size_t(b) * M * N;
                if (ci != cie)
                    gemm('T', 'N', M, N, K, 1.0, ldata[cie] + loff, K, mpsdata, K, 0.0, ret[ti].data() + ooff, M);
                else
                    gemm('N', 'N', M, N, K, 1.0, ldata[cie] + loff, M, mpsdata, K, 0.0, ret[ti].data() + ooff, M);
            }
        }
        return ret;
    }
};

// ------------------------------------------------------------------ schedule builders
static void op_iterate(const Oper& W, const OpB& blk, const double cp[4], Cohort& cg, unsigned s, unsigned m2,
                       const MPSBlock& mpsb, unsigned mps_offset, unsigned ci, unsigned bcol) {
    for (auto& e : blk.e) {
        unsigned casenr = 0;
        if (e.rs == 2 && e.cs == 2) casenr = 3;
        else if (e.rs == 2) casenr = 1;
        else if (e.cs == 2) casenr = 2;
        unsigned ti = mpsb.get_ti(mps_offset + e.row * m2, ci);
        if (ti != NPOS) cg.push_back(s, e.col, e.coeff * cp[casenr], ti, bcol);
    }
}

struct Site {
    Sym sy;
    Idx left_i, phys_i, right_i;
    Site(const mq_site_desc& d) : left_i(d.left_i), phys_i(d.phys_i), right_i(d.right_i) { sy.kind = d.sym; }
};

static void rshtm_t_tasks(const BIndex& right, const Idx& left_i, const Idx& right_i, const Idx& phys_i,
                          const PB& rpb, unsigned lb_ket, MPSBlock& mpsb) {
    Chg lc_ket = left_i.v[lb_ket].first;
    for (size_t s = 0; s < phys_i.size(); ++s) {
        Chg phys_in = phys_i.v[s].first;
        Chg rc_ket = fuse(lc_ket, phys_in);
        size_t rb_ket = right_i.position(rc_ket);
        if (rb_ket == right_i.size()) continue;
        unsigned rs_ket = unsigned(right_i.v[rb_ket].second);
        unsigned ket_offset = unsigned(rpb(phys_in, rc_ket));
        for (size_t rb_bra = 0; rb_bra < right_i.size(); ++rb_bra) {
            unsigned ci = right.cohort_index(rc_ket, right_i.v[rb_bra].first);
            if (ci == right.n_cohorts()) continue;
            unsigned cie = right.tr(ci) ? right.cohort_index(right_i.v[rb_bra].first, rc_ket) : ci;
            size_t sza = round_up128(left_i.v[lb_ket].second * right_i.v[rb_bra].second * right.n_blocks(cie));
            for (unsigned ss = 0; ss < phys_i.v[s].second; ++ss)
                mpsb.t_schedule.push_back(std::make_tuple(ket_offset + ss * rs_ket, ci, cie, lb_ket, sza));
        }
    }
}

// shtm_tasks (site_hamil_tasks.hpp:41-137)
static void shtm_tasks(const MPO& mpo, const BIndex& left, const BIndex& right, const Site& st, const PB& rpb,
                       unsigned lb_in, MPSBlock& mpsb) {
    const Idx &left_i = st.left_i, &right_i = st.right_i, &phys_i = st.phys_i;
    auto phys_s = phys_i.sizes();
    Chg lc_in = left_i.v[lb_in].first;
    unsigned ls_in = unsigned(left_i.v[lb_in].second);
    for (unsigned lb_out = 0; lb_out < left_i.size(); ++lb_out) {
        Chg lc_out = left_i.v[lb_out].first;
        unsigned ls_out = unsigned(left_i.v[lb_out].second);
        unsigned ci = left.cohort_index(lc_out, lc_in);
        if (ci == left.n_cohorts()) continue;
        unsigned cie = left.tr(ci) ? left.cohort_index(lc_in, lc_out) : ci;
        Cohort cohort(phys_s, lb_in, lb_out, ls_in, ls_out, ci, cie, unsigned(left.n_blocks(cie)));
        for (unsigned s = 0; s < phys_i.size(); ++s) {
            Chg phys_out = phys_i.v[s].first;
            Chg rc_out = fuse(lc_out, phys_out);
            size_t rb_out = right_i.position(rc_out);
            if (rb_out == right_i.size()) continue;
            unsigned rs_out = unsigned(right_i.v[rb_out].second);
            unsigned out_offset = unsigned(rpb(phys_out, rc_out));
            cohort.add_unit(s, unsigned(phys_i.v[s].second), rs_out, out_offset);
            W9j w9j(st.sy, lc_out, lc_in, rc_out);
            for (unsigned b1 = 0; b1 < mpo.rows; ++b1) {
                if (!left.has_block(ci, b1)) continue;
                unsigned left_idx = unsigned(left.offset(ci, b1) / (long(ls_in) * ls_out));
                int A = mpo.lspin[b1];
                if (st.sy.su2() && !triangle(st.sy.spin(lc_in), A, st.sy.spin(lc_out))) continue;
                for (unsigned b2 : mpo.row_b2[b1]) {
                    for (auto& term : mpo.at.at({b1, b2})) {
                        const Oper& W = mpo.ops[term.first];
                        int K = W.spin, Ap = mpo.rspin[b2];
                        for (auto& wb : W.blocks) {
                            if (phys_out != wb.out) continue;
                            Chg phys_in = wb.in;
                            Chg rc_in = fuse(lc_in, phys_in);
                            unsigned ci_right = right.cohort_index(rc_in, rc_out);
                            if (!right.has_block(ci_right, b2)) continue;
                            size_t rb_in = right_i.position(rc_in);
                            if (rb_in == right_i.size()) continue;
                            unsigned rs_in = unsigned(right_i.v[rb_in].second);
                            unsigned in_offset = unsigned(rpb(phys_in, rc_in));
                            long right_offset = right.offset(ci_right, b2);
                            double cp[4];
                            double scale =
                                right.conjugate_scale(ci_right, b2) * term.second * left.conjugate_scale(ci, b1);
                            w9j.set_scale(A, K, Ap, st.sy.spin(rc_in), scale, cp);
                            op_iterate(W, wb, cp, cohort, s, rs_in, mpsb, in_offset, ci_right,
                                       unsigned(right_offset / rs_in));
                        }
                    }
                }
                cohort.add_line(left_idx);
            }
        }
        if (cohort.n_tasks()) mpsb.cohorts.push_back(std::move(cohort));
    }
}

// rshtm_tasks (right_tasks.hpp:80-167)
static void rshtm_tasks(const MPO& mpo, const BIndex& right, const Site& st, const PB& rpb, unsigned lb_ket,
                        MPSBlock& mpsb, bool skip) {
    const Idx &left_i = st.left_i, &right_i = st.right_i, &phys_i = st.phys_i;
    auto phys_s = phys_i.sizes();
    Chg lc_ket = left_i.v[lb_ket].first;
    unsigned ls_ket = unsigned(left_i.v[lb_ket].second);
    for (unsigned lb_bra = 0; lb_bra < left_i.size(); ++lb_bra) {
        Chg lc_bra = left_i.v[lb_bra].first;
        if (std::abs(st.sy.particleNumber(lc_ket) - st.sy.particleNumber(lc_bra)) > 2) continue;
        unsigned ls_bra = unsigned(left_i.v[lb_bra].second);
        Cohort cohort(phys_s, lb_ket, lb_bra, ls_ket, ls_bra, 0, 0, mpo.rows);
        for (unsigned s = 0; s < phys_i.size(); ++s) {
            Chg phys_out = phys_i.v[s].first;
            Chg rc_bra = fuse(lc_bra, phys_out);
            size_t rb_bra = right_i.position(rc_bra);
            if (rb_bra == right_i.size()) continue;
            unsigned rs_bra = unsigned(right_i.v[rb_bra].second);
            unsigned bra_offset = unsigned(rpb(phys_out, rc_bra));
            cohort.add_unit(s, unsigned(phys_i.v[s].second), rs_bra, bra_offset);
            W9j w9j(st.sy, lc_bra, lc_ket, rc_bra);
            for (unsigned b1 = 0; b1 < mpo.rows; ++b1) {
                if (mpo.lh.skip(b1, lc_ket, lc_bra) && skip) continue;
                int A = mpo.lspin[b1];
                if (st.sy.su2() && !triangle(st.sy.spin(lc_ket), A, st.sy.spin(lc_bra))) continue;
                for (unsigned b2 : mpo.row_b2[b1]) {
                    for (auto& term : mpo.at.at({b1, b2})) {
                        const Oper& W = mpo.ops[term.first];
                        int K = W.spin, Ap = mpo.rspin[b2];
                        for (auto& wb : W.blocks) {
                            if (phys_out != wb.out) continue;
                            Chg phys_in = wb.in;
                            Chg rc_ket = fuse(lc_ket, phys_in);
                            unsigned ci = right.cohort_index(rc_ket, rc_bra);
                            if (!right.has_block(ci, b2)) continue;
                            size_t rb_ket = right_i.position(rc_ket);
                            if (rb_ket == right_i.size()) continue;
                            unsigned rs_ket = unsigned(right_i.v[rb_ket].second);
                            unsigned ket_offset = unsigned(rpb(phys_in, rc_ket));
                            long right_offset = right.offset(ci, b2);
                            double cp[4];
                            w9j.set_scale(A, K, Ap, st.sy.spin(rc_ket), right.conjugate_scale(ci, b2) * term.second, cp);
                            op_iterate(W, wb, cp, cohort, s, rs_ket, mpsb, ket_offset, ci, unsigned(right_offset / rs_ket));
                        }
                    }
                }
                cohort.add_line(b1);
            }
        }
        cohort.finalize();
        if (cohort.n_tasks()) mpsb.cohorts.push_back(std::move(cohort));
    }
}

static void lshtm_t_tasks(const BIndex& left, const Idx& left_i, const Idx& right_i, const Idx& phys_i, const PB& rpb,
                          unsigned rb_ket, MPSBlock& mpsb) {
    Chg rc_ket = right_i.v[rb_ket].first;
    unsigned rs_ket = unsigned(right_i.v[rb_ket].second);
    for (size_t s = 0; s < phys_i.size(); ++s) {
        Chg phys_in = phys_i.v[s].first;
        Chg lc_ket = fuse(rc_ket, neg(phys_in));
        size_t lb_ket = left_i.position(lc_ket);
        if (lb_ket == left_i.size()) continue;
        unsigned ket_offset = unsigned(rpb(phys_in, rc_ket));
        for (size_t lb_bra = 0; lb_bra < left_i.size(); ++lb_bra) {
            unsigned ci = left.cohort_index(left_i.v[lb_bra].first, lc_ket);
            if (ci == left.n_cohorts()) continue;
            unsigned cie = left.tr(ci) ? left.cohort_index(lc_ket, left_i.v[lb_bra].first) : ci;
            size_t sza = round_up128(left_i.v[lb_bra].second * rs_ket * left.n_blocks(cie));
            for (unsigned ss = 0; ss < phys_i.v[s].second; ++ss)
                mpsb.t_schedule.push_back(std::make_tuple(ket_offset + ss * rs_ket, ci, cie, unsigned(lb_ket), sza));
        }
    }
}

// lshtm_tasks (left_tasks.hpp:77-172)
static void lshtm_tasks(const MPO& mpo, const BIndex& left, const Site& st, const PB& bra_left_pb,
                        const PB& ket_right_pb, unsigned rb_ket, MPSBlock& mpsb, bool skip) {
    const Idx &left_i = st.left_i, &right_i = st.right_i, &phys_i = st.phys_i;
    auto phys_s = phys_i.sizes();
    Chg rc_ket = right_i.v[rb_ket].first;
    unsigned rs_ket = unsigned(right_i.v[rb_ket].second);
    for (unsigned rb_bra = 0; rb_bra < right_i.size(); ++rb_bra) {
        Chg rc_bra = right_i.v[rb_bra].first;
        if (std::abs(st.sy.particleNumber(rc_bra) - st.sy.particleNumber(rc_ket)) > 2) continue;
        unsigned rs_bra = unsigned(right_i.v[rb_bra].second);
        Cohort cohort(phys_s, rb_bra, rb_ket, rs_bra, rs_ket, 0, 0, mpo.cols, true);
        for (unsigned s = 0; s < phys_i.size(); ++s) {
            Chg phys_out = phys_i.v[s].first;
            Chg lc_bra = fuse(rc_bra, neg(phys_out));
            size_t lb_bra = left_i.position(lc_bra);
            if (lb_bra == left_i.size()) continue;
            unsigned ls_bra = unsigned(left_i.v[lb_bra].second);
            unsigned bra_offset = unsigned(bra_left_pb(phys_out, lc_bra));
            cohort.add_unit(s, unsigned(phys_i.v[s].second), ls_bra, bra_offset);
            for (unsigned b2 = 0; b2 < mpo.cols; ++b2) {
                if (mpo.rh.skip(b2, rc_bra, rc_ket) && skip) continue;
                int Ap = mpo.rspin[b2];
                if (st.sy.su2() && !triangle(st.sy.spin(rc_ket), Ap, st.sy.spin(rc_bra))) continue;
                for (unsigned b1 : mpo.col_b1[b2]) {
                    for (auto& term : mpo.at.at({b1, b2})) {
                        const Oper& W = mpo.ops[term.first];
                        int K = W.spin, A = mpo.lspin[b1];
                        for (auto& wb : W.blocks) {
                            if (phys_out != wb.out) continue;
                            Chg phys_in = wb.in;
                            Chg lc_ket = fuse(rc_ket, neg(phys_in));
                            size_t lb_ket = left_i.position(lc_ket);
                            if (lb_ket == left_i.size()) continue;
                            unsigned ci = left.cohort_index(lc_bra, lc_ket);
                            if (!left.has_block(ci, b1)) continue;
                            unsigned ls_ket = unsigned(left_i.v[lb_ket].second);
                            unsigned ket_offset = unsigned(ket_right_pb(phys_in, rc_ket));
                            long left_offset = left.offset(ci, b1);
                            double cp[4];
                            double scale = left.conjugate_scale(ci, b1) * term.second;
                            if (st.sy.su2()) {
                                const Sym& y = st.sy;
                                set_coupling(y.spin(lc_ket), std::abs(y.spin(fuse(rc_ket, neg(lc_ket)))), y.spin(rc_ket),
                                             A, K, Ap, y.spin(lc_bra), std::abs(y.spin(fuse(rc_bra, neg(lc_bra)))),
                                             y.spin(rc_bra), scale, cp);
                            } else {
                                cp[0] = cp[1] = cp[2] = cp[3] = scale;
                            }
                            op_iterate(W, wb, cp, cohort, s, rs_ket, mpsb, ket_offset, ci, unsigned(left_offset / ls_ket));
                        }
                    }
                }
                cohort.add_line(b2);
            }
        }
        cohort.finalize();
        if (cohort.n_tasks()) mpsb.cohorts.push_back(std::move(cohort));
    }
}

// ------------------------------------------------------------------ handles
struct Sched {
    int kind;   // 0 sigma, 1 left, 2 right
    std::vector<MPSBlock> blocks;
    // flattened copy for export
    std::vector<uint64_t> t_begin, coh_begin, coh_unit_begin, unit_line_begin, unit_task_begin;
    std::vector<uint32_t> t_mps_offset, t_ci, t_ci_eff, t_lb_ket, coh_lb, coh_rb, coh_ls, coh_rs, coh_ci, coh_ci_eff,
        coh_nSrows, coh_stripe, unit_offset, unit_ms, line_b1, line_count, task_ti, task_col;
    std::vector<double> task_alpha;
    std::vector<int64_t> coh_mpo_offsets;
    uint32_t aux = 0;
    size_t flops = 0;
    void flatten() {
        t_begin = {0};
        coh_begin = {0};
        coh_unit_begin = {0};
        unit_line_begin = {0};
        unit_task_begin = {0};
        for (auto& b : blocks) {
            for (auto& t : b.t_schedule) {
                t_mps_offset.push_back(std::get<0>(t));
                t_ci.push_back(std::get<1>(t));
                t_ci_eff.push_back(std::get<2>(t));
                t_lb_ket.push_back(std::get<3>(t));
            }
            t_begin.push_back(t_mps_offset.size());
            for (auto& c : b.cohorts) {
                coh_lb.push_back(c.lb);
                coh_rb.push_back(c.rb);
                coh_ls.push_back(c.ls);
                coh_rs.push_back(c.rs);
                coh_ci.push_back(c.ci);
                coh_ci_eff.push_back(c.ci_eff);
                coh_nSrows.push_back(c.nSrows);
                coh_stripe.push_back(c.stripe);
                if (kind != 0) coh_mpo_offsets.insert(coh_mpo_offsets.end(), c.mpo_offsets.begin(), c.mpo_offsets.end());
                for (auto& x : c.suv) {
                    unit_offset.push_back(x.offset);
                    unit_ms.push_back(x.ms);
                    for (size_t l = 0; l < x.b2s.size(); ++l) {
                        line_b1.push_back(x.b1[l]);
                        line_count.push_back(x.b2s[l]);
                    }
                    unit_line_begin.push_back(line_b1.size());
                    for (size_t a = 0; a < x.alpha.size(); ++a) {
                        task_ti.push_back(x.tidx[2 * a]);
                        task_col.push_back(x.tidx[2 * a + 1]);
                        task_alpha.push_back(x.alpha[a]);
                    }
                    unit_task_begin.push_back(task_ti.size());
                }
                coh_unit_begin.push_back(unit_offset.size());
            }
            coh_begin.push_back(coh_lb.size());
        }
    }
};

struct Bnd {
    BIndex idx;
    std::vector<std::vector<double>> data;
    // exported descriptor storage
    std::vector<mq_charge> bc, kc;
    std::vector<uint32_t> bsz, ksz, bpos, kpos, nb;
    std::vector<uint8_t> tr, trans;
    std::vector<int64_t> off;
    std::vector<double> cs;
};

static std::vector<const double*> view(const mq_bindex_desc& d, const double* const* data) {
    return std::vector<const double*>(data, data + d.n_cohorts);
}

}  // namespace oracle

using namespace oracle;

template <class F>
static int guard(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        last_error = e.what();
        return 1;
    }
}

extern "C" {

const char* mqo_last_error() { return last_error.c_str(); }

void mqo_set_dgemm(void* fn) { ext_dgemm = (dgemm64_t)fn; }

double mqo_wigner9j(int a, int b, int c, int d, int e, int f, int g, int h, int i) {
    return ninej(a, b, c, d, e, f, g, h, i);
}

void mqo_set_coupling(int a, int b, int c, int d, int e, int f, int g, int h, int i, double init, double* out4) {
    set_coupling(a, b, c, d, e, f, g, h, i, init, out4);
}

// create_contraction_schedule (site_hamil_schedule.hpp:40-93).  only (may be NULL): per MPS block
// a flag; blocks without it get an empty schedule, so that sigma is the contribution of the
// flagged ket blocks alone (a bounded sample of the full work for timing).
int mqo_schedule_sigma_blocks(const mq_site_desc* sd, const mq_mpo_desc* md, const mq_bindex_desc* ld,
                              const mq_bindex_desc* rd, const uint8_t* only, void** out) {
    return guard([&] {
        Site st(*sd);
        MPO mpo(*md);
        BIndex left(*ld), right(*rd);
        PB rpb(st.phys_i, st.right_i, true);
        auto* s = new Sched();
        s->kind = 0;
        s->blocks.resize(st.left_i.size());
#pragma omp parallel for schedule(dynamic, 1)
        for (long mb = 0; mb < long(st.left_i.size()); ++mb) {
            if (only && !only[mb]) continue;
            rshtm_t_tasks(right, st.left_i, st.right_i, st.phys_i, rpb, unsigned(mb), s->blocks[mb]);
            shtm_tasks(mpo, left, right, st, rpb, unsigned(mb), s->blocks[mb]);
        }
        // flops (MPSBlock::n_flops, tasks.cpp:777-796)
        for (auto& b : s->blocks) {
            for (auto& t : b.t_schedule)
                s->flops += 2ull * st.left_i.v[std::get<3>(t)].second * right.cohort_size(std::get<2>(t));
            for (auto& c : b.cohorts) s->flops += c.n_flops();
        }
        s->flatten();
        *out = s;
    });
}

int mqo_schedule_export(void* h, mq_schedule_desc* d, uint64_t* flops) {
    return guard([&] {
        Sched* s = (Sched*)h;
        d->n_blocks = uint32_t(s->blocks.size());
        d->t_begin = s->t_begin.data();
        d->t_mps_offset = s->t_mps_offset.data();
        d->t_ci = s->t_ci.data();
        d->t_ci_eff = s->t_ci_eff.data();
        d->t_lb_ket = s->t_lb_ket.data();
        d->coh_begin = s->coh_begin.data();
        d->coh_lb = s->coh_lb.data();
        d->coh_rb = s->coh_rb.data();
        d->coh_ls = s->coh_ls.data();
        d->coh_rs = s->coh_rs.data();
        d->coh_ci = s->coh_ci.data();
        d->coh_ci_eff = s->coh_ci_eff.data();
        d->coh_nSrows = s->coh_nSrows.data();
        d->coh_stripe = s->coh_stripe.data();
        d->coh_unit_begin = s->coh_unit_begin.data();
        d->unit_offset = s->unit_offset.data();
        d->unit_ms = s->unit_ms.data();
        d->unit_line_begin = s->unit_line_begin.data();
        d->line_b1 = s->line_b1.data();
        d->line_count = s->line_count.data();
        d->unit_task_begin = s->unit_task_begin.data();
        d->task_ti = s->task_ti.data();
        d->task_col = s->task_col.data();
        d->task_alpha = s->task_alpha.data();
        d->aux_dim = s->aux;
        d->coh_mpo_offsets = s->coh_mpo_offsets.empty() ? nullptr : s->coh_mpo_offsets.data();
        if (flops) *flops = s->flops;
    });
}

int mqo_schedule_sigma(const mq_site_desc* sd, const mq_mpo_desc* md, const mq_bindex_desc* ld,
                       const mq_bindex_desc* rd, void** out) {
    return mqo_schedule_sigma_blocks(sd, md, ld, rd, nullptr, out);
}

void mqo_schedule_free(void* h) { delete (Sched*)h; }

// super_hamil_mv, CPU branch (super_hamil_mv.hpp:121-132): psi/sigma in the padded layout.
int mqo_sigma_apply(void* h, const mq_site_desc* sd, const mq_bindex_desc* ld, const double* const* ldata,
                    const mq_bindex_desc* rd, const double* const* rdata, const double* psi, double* sigma) {
    return guard([&] {
        Sched* s = (Sched*)h;
        Site st(*sd);
        BIndex right(*rd);
        auto lv = view(*ld, ldata), rv = view(*rd, rdata);
        std::vector<size_t> lr_ket = st.left_i.sizes();
        // block pointers (DavidsonVector view)
        PB rpb(st.phys_i, st.right_i, true);
        std::vector<const double*> mps(st.left_i.size());
        std::vector<double*> out(st.left_i.size());
        size_t off = 0, total = 0;
        std::vector<size_t> bsz(st.left_i.size());
        for (size_t lb = 0; lb < st.left_i.size(); ++lb) {
            size_t cols = 0;
            for (size_t k = 0; k < st.phys_i.size(); ++k) {
                size_t rb = st.right_i.position(fuse(st.left_i.v[lb].first, st.phys_i.v[k].first));
                if (rb < st.right_i.size()) cols += st.phys_i.v[k].second * st.right_i.v[rb].second;
            }
            bsz[lb] = cols * st.left_i.v[lb].second;
            mps[lb] = psi + off;
            out[lb] = sigma + off;
            off += round_up128(bsz[lb]);
        }
        total = off;
        std::fill(sigma, sigma + total, 0.0);
        std::vector<std::mutex> mutexes(st.left_i.size());
#pragma omp parallel for schedule(dynamic, 1)
        for (long b = 0; b < long(s->blocks.size()); ++b) {
            auto T = s->blocks[b].create_T(right, rv, mps, lr_ket);
            for (auto& c : s->blocks[b].cohorts) c.contract(lv, T, out[c.rb], mutexes[c.rb]);
        }
    });
}

// Boundary-step results
static void export_bnd(Bnd* b, mq_bindex_desc* d) {
    BIndex& I = b->idx;
    b->bc.clear();
    b->kc.clear();
    b->bsz.clear();
    b->ksz.clear();
    for (auto& p : I.bra.v) b->bc.push_back({{p.first.a, p.first.b}}), b->bsz.push_back(uint32_t(p.second));
    for (auto& p : I.ket.v) b->kc.push_back({{p.first.a, p.first.b}}), b->ksz.push_back(uint32_t(p.second));
    b->bpos.clear();
    b->kpos.clear();
    b->nb.clear();
    b->tr.clear();
    b->trans.clear();
    b->off.clear();
    b->cs.clear();
    std::vector<std::pair<size_t, size_t>> pos(I.n_cohorts());
    for (auto& kv : I.lbrb) pos[kv.second] = kv.first;
    for (unsigned ci = 0; ci < I.n_cohorts(); ++ci) {
        b->bpos.push_back(uint32_t(pos[ci].first));
        b->kpos.push_back(uint32_t(pos[ci].second));
        b->nb.push_back(uint32_t(I.nbs[ci]));
        b->tr.push_back(uint8_t(I.tr_[ci]));
        for (unsigned k = 0; k < I.aux; ++k) {
            b->off.push_back(I.offsets[ci][k]);
            b->cs.push_back(I.cscale[ci][k]);
            b->trans.push_back(uint8_t(I.trans[ci][k]));
        }
    }
    d->bra = {uint32_t(b->bc.size()), b->bc.data(), b->bsz.data()};
    d->ket = {uint32_t(b->kc.size()), b->kc.data(), b->ksz.data()};
    d->n_cohorts = I.n_cohorts();
    d->aux_dim = I.aux;
    d->bra_pos = b->bpos.data();
    d->ket_pos = b->kpos.data();
    d->tr = b->tr.data();
    d->n_blocks = b->nb.data();
    d->offsets = b->off.data();
    d->conj_scales = b->cs.data();
    d->transposes = b->trans.data();
}

// overlap_mpo_left_step, CPU branch (move_boundary.hpp:147-213, 260-276).  bra in the left-paired
// layout, ket right-paired (padded layouts); result index + cohort data through the handle.
int mqo_left_step(const mq_site_desc* sd, const mq_mpo_desc* md, const double* bra_lp, const double* ket_rp,
                  const mq_bindex_desc* ld, const double* const* ldata, int symmetric, void** out) {
    return guard([&] {
        Site st(*sd);
        MPO mpo(*md);
        BIndex left(*ld);
        auto lv = view(*ld, ldata);
        PB ket_right_pb(st.phys_i, st.right_i, true);
        PB bra_left_pb(st.phys_i, st.left_i, false);
        size_t nr = st.right_i.size();
        std::vector<MPSBlock> tasks(nr);
#pragma omp parallel for schedule(dynamic, 1)
        for (long rb = 0; rb < long(nr); ++rb) {
            lshtm_t_tasks(left, st.left_i, st.right_i, st.phys_i, ket_right_pb, unsigned(rb), tasks[rb]);
            lshtm_tasks(mpo, left, st, bra_left_pb, ket_right_pb, unsigned(rb), tasks[rb], symmetric != 0);
        }
        auto* b = new Bnd();
        b->idx = BIndex(st.right_i, st.right_i, mpo.cols);
        for (unsigned rb = 0; rb < nr; ++rb)
            for (auto& c : tasks[rb].cohorts) b->idx.add_cohort(c.lb, rb, c.mpo_offsets);
        if (symmetric) b->idx.complement_transpose(mpo.rh, true, st.sy);
        b->data.resize(b->idx.n_cohorts());
        for (unsigned ci = 0; ci < b->idx.n_cohorts(); ++ci) b->data[ci].assign(b->idx.cohort_size(ci), 0.0);
        // ket block views (right-paired) and bra block views (left-paired)
        std::vector<const double*> ket(st.left_i.size()), bra(nr);
        size_t off = 0;
        for (size_t lb = 0; lb < st.left_i.size(); ++lb) {
            size_t cols = 0;
            for (size_t k = 0; k < st.phys_i.size(); ++k) {
                size_t r = st.right_i.position(fuse(st.left_i.v[lb].first, st.phys_i.v[k].first));
                if (r < nr) cols += st.phys_i.v[k].second * st.right_i.v[r].second;
            }
            ket[lb] = ket_rp + off;
            off += round_up128(cols * st.left_i.v[lb].second);
        }
        off = 0;
        for (size_t rb = 0; rb < nr; ++rb) {
            size_t rows = 0;
            for (size_t k = 0; k < st.phys_i.size(); ++k) {
                size_t l = st.left_i.position(fuse(st.right_i.v[rb].first, neg(st.phys_i.v[k].first)));
                if (l < st.left_i.size()) rows += st.phys_i.v[k].second * st.left_i.v[l].second;
            }
            bra[rb] = bra_lp + off;
            off += round_up128(rows * st.right_i.v[rb].second);
        }
#pragma omp parallel for schedule(dynamic, 1)
        for (long rb = 0; rb < long(nr); ++rb) {
            auto T = tasks[rb].create_T_left(left, lv, ket, st.right_i.v[rb].second);
            for (auto& c : tasks[rb].cohorts) {
                unsigned ci = b->idx.cohort_index(size_t(c.lb), size_t(rb));
                c.prop_l(bra[c.lb], T, b->data[ci].data());
            }
        }
        *out = b;
    });
}

// overlap_mpo_right_step, CPU branch (move_boundary.hpp:281-337, 385-401).  bra and ket right-paired.
int mqo_right_step(const mq_site_desc* sd, const mq_mpo_desc* md, const double* bra_rp, const double* ket_rp,
                   const mq_bindex_desc* rd, const double* const* rdata, int symmetric, void** out) {
    return guard([&] {
        Site st(*sd);
        MPO mpo(*md);
        BIndex right(*rd);
        auto rv = view(*rd, rdata);
        PB bra_right_pb(st.phys_i, st.right_i, true);
        size_t nl = st.left_i.size();
        std::vector<MPSBlock> tasks(nl);
#pragma omp parallel for schedule(dynamic, 1)
        for (long lb = 0; lb < long(nl); ++lb) {
            rshtm_t_tasks(right, st.left_i, st.right_i, st.phys_i, bra_right_pb, unsigned(lb), tasks[lb]);
            rshtm_tasks(mpo, right, st, bra_right_pb, unsigned(lb), tasks[lb], symmetric != 0);
        }
        auto* b = new Bnd();
        b->idx = BIndex(st.left_i, st.left_i, mpo.rows);
        for (unsigned lb = 0; lb < nl; ++lb)
            for (auto& c : tasks[lb].cohorts) b->idx.add_cohort(lb, c.rb, c.mpo_offsets);
        b->idx.complement_transpose(mpo.lh, false, st.sy);
        b->data.resize(b->idx.n_cohorts());
        for (unsigned ci = 0; ci < b->idx.n_cohorts(); ++ci) b->data[ci].assign(b->idx.cohort_size(ci), 0.0);
        std::vector<const double*> ket(nl), bra(nl);
        size_t off = 0;
        for (size_t lb = 0; lb < nl; ++lb) {
            size_t cols = 0;
            for (size_t k = 0; k < st.phys_i.size(); ++k) {
                size_t r = st.right_i.position(fuse(st.left_i.v[lb].first, st.phys_i.v[k].first));
                if (r < st.right_i.size()) cols += st.phys_i.v[k].second * st.right_i.v[r].second;
            }
            ket[lb] = ket_rp + off;
            bra[lb] = bra_rp + off;
            off += round_up128(cols * st.left_i.v[lb].second);
        }
        std::vector<size_t> lr_ket = st.left_i.sizes();
#pragma omp parallel for schedule(dynamic, 1)
        for (long lb = 0; lb < long(nl); ++lb) {
            auto T = tasks[lb].create_T(right, rv, ket, lr_ket);
            for (auto& c : tasks[lb].cohorts) {
                unsigned ci = b->idx.cohort_index(size_t(lb), size_t(c.rb));
                c.prop_r(bra[c.rb], T, b->data[ci].data());
            }
        }
        *out = b;
    });
}

int mqo_boundary_index(void* h, mq_bindex_desc* d) {
    return guard([&] { export_bnd((Bnd*)h, d); });
}

const double* mqo_boundary_data(void* h, uint32_t ci) { return ((Bnd*)h)->data[ci].data(); }

void mqo_boundary_free(void* h) { delete (Bnd*)h; }

// jacobi_davidson::calculate_eigenvalue (ietl/jacobi.h:386-474) with the jcd_gmres_solver correction
// (ietl/jacobi.h:126-164; plain when gmres_iter = 0) and basic_iteration (ietl/iteration.h:40-78).
// The subspace eigenproblem is solved by cyclic Jacobi (dsyevx range 'I', il = iu = 1).
static double jacobi_min(std::vector<double> A, int n, std::vector<double>& vec) {
    std::vector<double> V(n * n, 0.0);
    for (int i = 0; i < n; ++i) V[i * n + i] = 1;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0;
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j) off += A[i * n + j] * A[i * n + j];
        if (off < 1e-30) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                double apq = A[p * n + q];
                if (std::fabs(apq) < 1e-300) continue;
                double th = (A[q * n + q] - A[p * n + p]) / (2 * apq);
                double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1));
                double c = 1 / std::sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < n; ++k) {  // columns p, q
                    double x = A[k * n + p], y = A[k * n + q];
                    A[k * n + p] = c * x - s * y;
                    A[k * n + q] = s * x + c * y;
                }
                for (int k = 0; k < n; ++k) {  // rows p, q
                    double x = A[p * n + k], y = A[q * n + k];
                    A[p * n + k] = c * x - s * y;
                    A[q * n + k] = s * x + c * y;
                }
                for (int k = 0; k < n; ++k) {
                    double x = V[k * n + p], y = V[k * n + q];
                    V[k * n + p] = c * x - s * y;
                    V[k * n + q] = s * x + c * y;
                }
            }
    }
    int im = 0;
    for (int i = 1; i < n; ++i)
        if (A[i * n + i] < A[im * n + im]) im = i;
    vec.resize(n);
    for (int k = 0; k < n; ++k) vec[k] = V[k * n + im];
    return A[im * n + im];
}

int mqo_subspace_min(const double* A, int n, double* vec, double* val) {
    return guard([&] {
        std::vector<double> a(A, A + n * n), v;
        *val = jacobi_min(a, n, v);
        std::copy(v.begin(), v.end(), vec);
    });
}

// ietl_gmres (ietl/gmres.h:78-140) on A x = b from x0, A the projected correction operator
// (1 - uu^T)(H - theta)(1 - uu^T) (jcd_solver_operator, ietl/jacobi.h:266-284); its bookkeeping kept:
// at most max_iter - 1 Arnoldi steps, and on convergence at step i the update uses i basis vectors.
static std::vector<double> gmres_restated(const std::function<void(const std::vector<double>&, std::vector<double>&)>& A,
                                          const std::vector<double>& b, const std::vector<double>& x0, double abs_tol,
                                          int max_iter) {
    const size_t n = b.size();
    auto dot = [&](const std::vector<double>& a, const std::vector<double>& c) {
        double r = 0;
        for (size_t i = 0; i < n; ++i) r += a[i] * c[i];
        return r;
    };
    std::vector<double> sv(max_iter + 1, 0.0), cs(max_iter + 1, 0.0), sn(max_iter + 1, 0.0);
    std::vector<std::vector<double>> v(max_iter + 1);
    std::vector<double> w;
    A(x0, v[0]);
    for (size_t i = 0; i < n; ++i) v[0][i] = b[i] - v[0][i];
    sv[0] = std::sqrt(dot(v[0], v[0]));
    if (std::fabs(sv[0]) < abs_tol) return x0;
    for (auto& x : v[0]) x /= sv[0];
    std::vector<double> H(size_t(max_iter + 1) * (max_iter + 1), 0.0);
    auto h = [&](int i, int j) -> double& { return H[size_t(i) * (max_iter + 1) + j]; };
    auto rot = [](double& dx, double& dy, double c, double s_) {
        double r0 = c * dx + s_ * dy, r1 = -s_ * dx + c * dy;
        dx = r0, dy = r1;
    };
    int i = 0;
    for (; i < max_iter - 1; ++i) {
        A(v[i], w);
        for (int k = 0; k <= i; ++k) {
            h(k, i) = dot(w, v[k]);
            axpy(n, -h(k, i), v[k].data(), w.data());
        }
        h(i + 1, i) = std::sqrt(dot(w, w));
        v[i + 1] = w;
        for (auto& x : v[i + 1]) x /= h(i + 1, i);
        for (int k = 0; k < i; ++k) rot(h(k, i), h(k + 1, i), cs[k], sn[k]);
        double dx = h(i, i), dy = h(i + 1, i);
        if (dy == 0.0) {
            cs[i] = 1, sn[i] = 0;
        } else if (std::fabs(dy) > std::fabs(dx)) {
            double t = dx / dy;
            sn[i] = 1.0 / std::sqrt(1.0 + t * t), cs[i] = t * sn[i];
        } else {
            double t = dy / dx;
            cs[i] = 1.0 / std::sqrt(1.0 + t * t), sn[i] = t * cs[i];
        }
        rot(h(i, i), h(i + 1, i), cs[i], sn[i]);
        rot(sv[i], sv[i + 1], cs[i], sn[i]);
        if (std::fabs(sv[i + 1]) < abs_tol) break;
    }
    std::vector<double> y(sv.begin(), sv.begin() + i);
    for (int a = i - 1; a >= 0; --a) {
        y[a] /= h(a, a);
        for (int c = a - 1; c >= 0; --c) y[c] -= h(c, a) * y[a];
    }
    std::vector<double> x = x0;
    for (int k = 0; k < i; ++k) axpy(n, y[k], v[k].data(), x.data());
    return x;
}

int mqo_jd_ex(void* h, const mq_site_desc* sd, const mq_bindex_desc* ld, const double* const* ldata,
              const mq_bindex_desc* rd, const double* const* rdata, uint64_t n, const double* init, double* out,
              double tol, int maxiter, int gmres_iter, int n_ortho, const double* const* ortho, double* energy,
              int* iters) {
    return guard([&] {
        auto dot = [&](const std::vector<double>& a, const std::vector<double>& b) {
            double s = 0;
            for (size_t i = 0; i < n; ++i) s += a[i] * b[i];
            return s;
        };
        auto mult = [&](const std::vector<double>& x, std::vector<double>& y) {
            y.assign(n, 0.0);
            int rc = mqo_sigma_apply(h, sd, ld, ldata, rd, rdata, x.data(), y.data());
            if (rc) throw std::runtime_error(last_error);
        };
        // ortho vectors: Gram-Schmidt'ed copies (solver/ietl_jacobi_davidson.h:41-47), projected out by
        // DavidsonVS::project (solver/ietl_interface_dv.h:94-99)
        if (uint64_t(n_ortho) >= n) n_ortho = int(n) - 1;
        std::vector<std::vector<double>> O(n_ortho);
        for (int j = 0; j < n_ortho; ++j) {
            O[j].assign(ortho[j], ortho[j] + n);
            for (int j0 = 0; j0 < j; ++j0) axpy(n, -dot(O[j0], O[j]) / dot(O[j0], O[j0]), O[j0].data(), O[j].data());
        }
        auto project = [&](std::vector<double>& t) {
            for (int j = 0; j < n_ortho; ++j) axpy(n, -dot(O[j], t) / dot(O[j], O[j]), O[j].data(), t.data());
        };
        std::vector<std::vector<double>> V(maxiter + 1), VA(maxiter);
        std::vector<double> M(size_t(maxiter) * maxiter, 0.0), s;
        V[0].assign(init, init + n);
        project(V[0]);
        int it = 0;
        double theta = 0;
        std::vector<double> u, uA;
        while (true) {
            std::vector<double>& t = V[it];
            double tau = std::sqrt(dot(t, t));
            for (int i = 1; i <= it; ++i) axpy(n, -dot(V[i - 1], t), V[i - 1].data(), t.data());
            if (std::sqrt(dot(t, t)) < 0.25 * tau)
                for (int i = 1; i <= it; ++i) axpy(n, -dot(V[i - 1], t), V[i - 1].data(), t.data());
            project(t);
            double nt = std::sqrt(dot(t, t));
            for (auto& x : t) x /= nt;
            mult(t, VA[it]);
            for (int i = 0; i <= it; ++i) M[size_t(i) * maxiter + it] = dot(V[i], VA[it]);
            int dim = it + 1;
            std::vector<double> Msub(dim * dim);
            for (int i = 0; i < dim; ++i)
                for (int j = 0; j < dim; ++j)
                    Msub[i * dim + j] = (i <= j) ? M[size_t(i) * maxiter + j] : M[size_t(j) * maxiter + i];
            theta = jacobi_min(Msub, dim, s);
            u.assign(n, 0.0);
            uA.assign(n, 0.0);
            for (int j = 0; j <= it; ++j) {
                axpy(n, s[j], V[j].data(), u.data());
                axpy(n, s[j], VA[j].data(), uA.data());
            }
            project(uA);
            axpy(n, -theta, u.data(), uA.data());
            ++it;
            double rn = std::sqrt(dot(uA, uA));
            if (rn <= tol * std::fabs(theta) || rn < tol || it >= maxiter) break;
            double dru = dot(uA, u), duu = dot(u, u);
            V[it].assign(n, 0.0);
            for (size_t i = 0; i < n; ++i) V[it][i] = -uA[i] + dru / duu * u[i];
            if (gmres_iter > 0) {   // jcd_gmres_solver (ietl/jacobi.h:141-157), tolerance 2^-(it+1)
                const std::vector<double> uu = u;
                const double th = theta;
                auto op = [&](const std::vector<double>& x, std::vector<double>& y) {
                    std::vector<double> t2 = x;
                    axpy(n, -dot(uu, x), uu.data(), t2.data());
                    mult(t2, y);
                    axpy(n, -th, t2.data(), y.data());
                    axpy(n, -dot(uu, y), uu.data(), y.data());
                };
                std::vector<double> b(n);
                for (size_t i = 0; i < n; ++i) b[i] = -uA[i];
                V[it] = gmres_restated(op, b, V[it], 1.0 / std::pow(2.0, double(it + 1)), gmres_iter);
            }
        }
        std::copy(u.begin(), u.end(), out);
        *energy = theta;
        *iters = it;
    });
}

int mqo_jd(void* h, const mq_site_desc* sd, const mq_bindex_desc* ld, const double* const* ldata,
           const mq_bindex_desc* rd, const double* const* rdata, uint64_t n, const double* init, double* out,
           double tol, int maxiter, double* energy, int* iters) {
    return mqo_jd_ex(h, sd, ld, ldata, rd, rdata, n, init, out, tol, maxiter, 0, 0, nullptr, energy, iters);
}

}  // extern "C"
