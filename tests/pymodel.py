"""Independent Python-integer model of alt_bn128 used to pin the oracle.

Written from the mathematics only (modular arithmetic with Python ints,
affine short-Weierstrass group law, Fermat inverse, DFT by definition); it
shares no code with the C++ oracle, so a dropped term, wrong sign or
transposed operand in either shows up as a mismatch.
"""
import numpy as np

P = 0x30644E72E131A029B85045B68181585D97816A916871CA8D3C208C16D87CFD47
R = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001
MONT = 1 << 256


def limbs_to_int(a):
    a = [int(x) for x in np.asarray(a, dtype=np.uint64).reshape(-1)]
    return sum(v << (64 * i) for i, v in enumerate(a))


def int_to_limbs(x, nlimbs=4):
    return np.array([(x >> (64 * i)) & ((1 << 64) - 1) for i in range(nlimbs)], dtype=np.uint64)


def elems(buf, width=4):
    """Split a uint64 buffer into Python ints of `width` limbs each."""
    b = np.asarray(buf, dtype=np.uint64).reshape(-1, width)
    return [limbs_to_int(row) for row in b]


def to_mont(x, m):
    return (x * MONT) % m


def from_mont(x, m):
    return (x * pow(MONT, -1, m)) % m


# --- Fq2 = Fq[u]/(u^2+1): tuples (c0, c1) of ints (canonical)
def f2_add(a, b):
    return ((a[0] + b[0]) % P, (a[1] + b[1]) % P)


def f2_sub(a, b):
    return ((a[0] - b[0]) % P, (a[1] - b[1]) % P)


def f2_mul(a, b):
    return ((a[0] * b[0] - a[1] * b[1]) % P, (a[0] * b[1] + a[1] * b[0]) % P)


def f2_inv(a):
    n = pow((a[0] * a[0] + a[1] * a[1]) % P, P - 2, P)
    return (a[0] * n % P, (-a[1] * n) % P)


class Curve:
    """Affine group law over a field given by add/sub/mul/inv callables."""

    def __init__(self, add, sub, mul, inv, zero, b):
        self.add, self.sub, self.mul, self.inv, self.zero, self.b = add, sub, mul, inv, zero, b

    def on_curve(self, pt):
        if pt is None:
            return True
        x, y = pt
        return self.mul(y, y) == self.add(self.mul(self.mul(x, x), x), self.b)

    def neg(self, pt):
        if pt is None:
            return None
        return (pt[0], self.sub(self.zero, pt[1]))

    def padd(self, p1, p2):
        if p1 is None:
            return p2
        if p2 is None:
            return p1
        x1, y1 = p1
        x2, y2 = p2
        if x1 == x2:
            if y1 != y2 or y1 == self.zero:
                return None
            three = self.add(self.add(self.mul(x1, x1), self.mul(x1, x1)), self.mul(x1, x1))
            m = self.mul(three, self.inv(self.add(y1, y1)))
        else:
            m = self.mul(self.sub(y2, y1), self.inv(self.sub(x2, x1)))
        x3 = self.sub(self.sub(self.mul(m, m), x1), x2)
        y3 = self.sub(self.mul(m, self.sub(x1, x3)), y1)
        return (x3, y3)

    def smul(self, k, pt):
        acc = None
        addend = pt
        while k:
            if k & 1:
                acc = self.padd(acc, addend)
            addend = self.padd(addend, addend)
            k >>= 1
        return acc


G1C = Curve(lambda a, b: (a + b) % P, lambda a, b: (a - b) % P, lambda a, b: a * b % P,
            lambda a: pow(a, P - 2, P), 0, 3)
_TWIST_B = f2_mul((3, 0), f2_inv((9, 1)))
G2C = Curve(f2_add, f2_sub, f2_mul, f2_inv, (0, 0), _TWIST_B)

G1_GEN = (1, 2)
G2_GEN = ((0x1800DEEF121F1E76426A00665E5C4479674322D4F75EDADD46DEBD5CD992F6ED,
           0x198E9393920D483A7260BFB731FB5D25F1AA493335A9E71297E485B7AEF312C2),
          (0x12C85EA5DB8C6DEB4AAB71808DCB408FE3D1E7690C43D37B4CE6CC0166FA7DAA,
           0x090689D0585FF075EC9E99AD690C3395BC4B313370B38EF355ACDADCD122975B))


# --- conversions between buffers and model points
def g1_aff_from_buf(buf):
    """64-byte affine (Montgomery) -> (x, y) canonical ints or None."""
    x, y = elems(buf)
    if x == 0 and y == 0:
        return None
    return (from_mont(x, P), from_mont(y, P))


def g1_jac_from_buf(buf):
    """96-byte normalised Jacobian -> point or None (libff infinity has Z = 0)."""
    x, y, z = elems(buf)
    if z == 0:
        return None
    assert from_mont(z, P) == 1, "not normalised"
    return (from_mont(x, P), from_mont(y, P))


def g2_aff_from_buf(buf):
    v = elems(buf)
    if all(t == 0 for t in v):
        return None
    v = [from_mont(t, P) for t in v]
    return ((v[0], v[1]), (v[2], v[3]))


def g2_jac_from_buf(buf):
    v = elems(buf)
    if v[4] == 0 and v[5] == 0:
        return None
    v = [from_mont(t, P) for t in v]
    assert (v[4], v[5]) == (1, 0), "not normalised"
    return ((v[0], v[1]), (v[2], v[3]))


def g1_to_aff_buf(pt):
    if pt is None:
        return np.zeros(8, dtype=np.uint64)
    return np.concatenate([int_to_limbs(to_mont(pt[0], P)), int_to_limbs(to_mont(pt[1], P))])


def g2_to_aff_buf(pt):
    if pt is None:
        return np.zeros(16, dtype=np.uint64)
    (a, b), (c, d) = pt
    return np.concatenate([int_to_limbs(to_mont(t, P)) for t in (a, b, c, d)])


def dft(xs, w, m):
    """X_k = sum_j x_j w^(jk) mod m, by definition."""
    n = len(xs)
    return [sum(xs[j] * pow(w, j * k, m) for j in range(n)) % m for k in range(n)]
