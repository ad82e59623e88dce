"""Charges, indices and product bases of the host front-end.

Charges are 2-tuples of ints for every group (U1 uses the first entry only), fused by
component-wise addition (block_matrix/symmetry/u1.h, 2u1.h, su2u1.h of the reference).  An Index
keeps its sectors in descending lexicographic order like the reference's sorted Index
(block_matrix/indexing_stable.hpp:61-66, 166-176).
"""

SYM_U1, SYM_TWOU1, SYM_SU2U1 = 1, 2, 3


def fuse(a, b):
    return (a[0] + b[0], a[1] + b[1])


def neg(a):
    return (-a[0], -a[1])


def sub(a, b):
    return (a[0] - b[0], a[1] - b[1])


class Index:
    def __init__(self, charges, sizes, sort=True):
        pairs = [(tuple(int(x) for x in c), int(s)) for c, s in zip(charges, sizes)]
        if sort:
            pairs.sort(key=lambda p: p[0], reverse=True)
        self.charges = [c for c, _ in pairs]
        self.sizes = [s for _, s in pairs]
        self._pos = {c: i for i, c in enumerate(self.charges)}
        if len(self._pos) != len(self.charges):
            raise ValueError("duplicate charge in Index")

    def __len__(self):
        return len(self.charges)

    def position(self, c):
        return self._pos.get(tuple(c), len(self.charges))

    def has(self, c):
        return tuple(c) in self._pos

    def size_of(self, c):
        i = self._pos.get(tuple(c))
        return 0 if i is None else self.sizes[i]

    def total(self):
        return sum(self.sizes)

    def __eq__(self, o):
        return isinstance(o, Index) and self.charges == o.charges and self.sizes == o.sizes

    def __repr__(self):
        return "Index(%s)" % ", ".join("%s:%d" % (c, s) for c, s in zip(self.charges, self.sizes))

    @staticmethod
    def from_dict(d):
        return Index(list(d.keys()), list(d.values()))


class ProductBasis:
    """Offsets of (a-sector, b-sector) pairs inside the fused sector f(a, b), a outer, b inner
    (reference block_matrix/indexing_stable.hpp:310-378).  sign=-1 fuses (-a, b)."""

    def __init__(self, a, b, sign=+1):
        self.off = {}
        self.size = {}
        for i, ca in enumerate(a.charges):
            for j, cb in enumerate(b.charges):
                pc = sub(cb, ca) if sign < 0 else fuse(ca, cb)
                s = self.size.get(pc, 0)
                self.off[(ca, cb)] = s
                self.size[pc] = s + a.sizes[i] * b.sizes[j]

    def __call__(self, ca, cb):
        return self.off[(tuple(ca), tuple(cb))]


def product_index(a, b):
    """Index of the fused sectors of a (x) b and the ProductBasis giving sub-index offsets."""
    pb = ProductBasis(a, b, +1)
    return Index(list(pb.size.keys()), list(pb.size.values())), pb


class SiteLayout:
    """Right- and left-paired block layouts of a site tensor (mirrors the C-ABI conventions)."""

    def __init__(self, left, phys, right):
        self.left, self.phys, self.right = left, phys, right
        self.rpb = ProductBasis(phys, right, -1)
        self.lpb = ProductBasis(phys, left, +1)
        self.rp_cols, self.rp_offset = [], []
        tot = 0
        for lc, ls in zip(left.charges, left.sizes):
            cols = 0
            for pc, d in zip(phys.charges, phys.sizes):
                rs = right.size_of(fuse(lc, pc))
                cols += d * rs
            self.rp_cols.append(cols)
            self.rp_offset.append(tot)
            tot += -(-(ls * cols) // 128) * 128
        self.rp_total = tot
        self.lp_rows, self.lp_offset = [], []
        tot = 0
        for rc, rs in zip(right.charges, right.sizes):
            rows = 0
            for pc, d in zip(phys.charges, phys.sizes):
                rows += d * left.size_of(sub(rc, pc))
            self.lp_rows.append(rows)
            self.lp_offset.append(tot)
            tot += -(-(rows * rs) // 128) * 128
        self.lp_total = tot
