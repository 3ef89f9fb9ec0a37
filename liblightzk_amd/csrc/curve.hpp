// Device elliptic-curve arithmetic for alt_bn128 G1 (over Fq) and G2 (over
// Fq2, one point per lane), short Weierstrass with a = 0.
//
// Bucket accumulators use extended Jacobian "XYZZ" coordinates
// (x = X/ZZ, y = Y/ZZZ; Bernstein-Lange EFD, shortw-xyzz, 2008 formulas):
//   madd-2008-s   XYZZ += affine        8M + 2S
//   mdbl-2008-s-1 affine doubled         2M + 2S
//   add-2008-s    XYZZ += XYZZ         12M + 2S
//   dbl-2008-s-1  XYZZ doubled           6M + 3S
// The point at infinity is ZZ = ZZZ = 0; an affine input (0, 0) is infinity
// (not on y^2 = x^3 + b). The reference accumulates in Jacobian coordinates
// (include/device-arith/curve/jacobian_t.cuh:156-216 aadd); XYZZ is chosen
// here because its mixed addition is 2 multiplications cheaper and its
// doubling/infinity cases are cheap to detect (P == 0 from U2 - X1).
#pragma once
#include "field.hpp"

namespace hzk {

template <class F>
struct affine_t {
    F x, y;
    HZK_DEV bool is_inf() const { return x.is_zero() && y.is_zero(); }
    static HZK_DEV affine_t inf() { return {F::zero(), F::zero()}; }
    HZK_DEV affine_t neg() const { return {x, y.neg()}; }
};

template <class F>
struct xyzz_t {
    F X, Y, ZZ, ZZZ;

    static HZK_DEV xyzz_t inf() { return {F::zero(), F::zero(), F::zero(), F::zero()}; }
    HZK_DEV bool is_inf() const { return ZZ.is_zero(); }
    static HZK_DEV xyzz_t from_affine(const affine_t<F>& a) {
        if (a.is_inf()) return inf();
        return {a.x, a.y, F::one(), F::one()};
    }
    HZK_DEV xyzz_t neg() const { return {X, Y.neg(), ZZ, ZZZ}; }

    // 2 * (x, y) for a finite affine point (mdbl-2008-s-1, a = 0)
    static HZK_DEV xyzz_t dbl_affine(const affine_t<F>& a) {
        F U = a.y.dbl();
        F V = U.sqr();
        F W = U * V;
        F S = a.x * V;
        F xx = a.x.sqr();
        F M = xx.dbl() + xx;
        F X3 = M.sqr() - S.dbl();
        F Y3 = M * (S - X3) - W * a.y;
        return {X3, Y3, V, W};
    }

    // doubling (dbl-2008-s-1, a = 0); infinity stays infinity (ZZ = 0 propagates)
    HZK_DEV xyzz_t dbl() const {
        F U = Y.dbl();
        F V = U.sqr();
        F W = U * V;
        F S = X * V;
        F xx = X.sqr();
        F M = xx.dbl() + xx;
        F X3 = M.sqr() - S.dbl();
        F Y3 = M * (S - X3) - W * Y;
        return {X3, Y3, V * ZZ, W * ZZZ};
    }

    // this += affine (madd-2008-s) with every special case handled:
    // this = inf, a = inf, a = this (doubling), a = -this (infinity).
    HZK_DEV void add_affine(const affine_t<F>& a) {
        if (a.is_inf()) return;
        if (is_inf()) { *this = from_affine(a); return; }
        F U2 = a.x * ZZ;
        F S2 = a.y * ZZZ;
        F P = U2 - X;
        F R = S2 - Y;
        if (P.is_zero()) {
            if (R.is_zero()) *this = dbl_affine(a);
            else *this = inf();
            return;
        }
        F PP = P.sqr();
        F PPP = P * PP;
        F Q = X * PP;
        F X3 = R.sqr() - PPP - Q.dbl();
        F Y3 = R * (Q - X3) - Y * PPP;
        ZZ = ZZ * PP;
        ZZZ = ZZZ * PPP;
        X = X3;
        Y = Y3;
    }

    // this += o (add-2008-s) with every special case handled
    HZK_DEV void add(const xyzz_t& o) {
        if (o.is_inf()) return;
        if (is_inf()) { *this = o; return; }
        F U1 = X * o.ZZ;
        F U2 = o.X * ZZ;
        F S1 = Y * o.ZZZ;
        F S2 = o.Y * ZZZ;
        F P = U2 - U1;
        F R = S2 - S1;
        if (P.is_zero()) {
            if (R.is_zero()) *this = dbl();
            else *this = inf();
            return;
        }
        F PP = P.sqr();
        F PPP = P * PP;
        F Q = U1 * PP;
        F X3 = R.sqr() - PPP - Q.dbl();
        F Y3 = R * (Q - X3) - S1 * PPP;
        ZZ = ZZ * o.ZZ * PP;
        ZZZ = ZZZ * o.ZZZ * PPP;
        X = X3;
        Y = Y3;
    }
};

typedef affine_t<fq_t> g1_affine;
typedef xyzz_t<fq_t> g1_xyzz;
typedef affine_t<fq2_t> g2_affine;
typedef xyzz_t<fq2_t> g2_xyzz;

}  // namespace hzk
