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

    // Products are issued in groups of independent ones (F::mul2 / mul3 /
    // sqr2), which the field layer runs as interleaved chains.

    // 2 * (x, y) for a finite affine point (mdbl-2008-s-1, a = 0)
    static HZK_DEV xyzz_t dbl_affine(const affine_t<F>& a) {
        F U = a.y.dbl(), V, xx;
        F::sqr2(V, U, xx, a.x);
        F M = xx.dbl() + xx;
        F W, S, MM;
        F::mul3(W, U, V, S, a.x, V, MM, M, M);
        F X3 = MM - S.dbl();
        F t, Wy;
        F::mul2(t, M, S - X3, Wy, W, a.y);
        return {X3, t - Wy, V, W};
    }

    // doubling (dbl-2008-s-1, a = 0); infinity stays infinity (ZZ = 0 propagates)
    HZK_DEV xyzz_t dbl() const {
        F U = Y.dbl(), V, xx;
        F::sqr2(V, U, xx, X);
        F W, S, ZZ3;
        F::mul3(W, U, V, S, X, V, ZZ3, V, ZZ);
        F M = xx.dbl() + xx;
        F MM, ZZZ3, Wy;
        F::mul3(MM, M, M, ZZZ3, W, ZZZ, Wy, W, Y);
        F X3 = MM - S.dbl();
        F Y3 = M * (S - X3) - Wy;
        return {X3, Y3, ZZ3, ZZZ3};
    }

    // this += affine (madd-2008-s) with every special case handled:
    // this = inf, a = inf, a = this (doubling), a = -this (infinity).
    HZK_DEV void add_affine(const affine_t<F>& a) {
        if (a.is_inf()) return;
        if (is_inf()) { *this = from_affine(a); return; }
        F U2, S2;
        F::mul2(U2, a.x, ZZ, S2, a.y, ZZZ);
        F P = U2 - X;
        F R = S2 - Y;
        if (P.is_zero()) {
            if (R.is_zero()) *this = dbl_affine(a);
            else *this = inf();
            return;
        }
        F PP, RR;
        F::sqr2(PP, P, RR, R);
        F PPP, Q, ZZ3;
        F::mul3(PPP, P, PP, Q, X, PP, ZZ3, ZZ, PP);
        F X3 = RR - PPP - Q.dbl();
        F t, YP, ZZZ3;
        F::mul3(t, R, Q - X3, YP, Y, PPP, ZZZ3, ZZZ, PPP);
        X = X3;
        Y = t - YP;
        ZZ = ZZ3;
        ZZZ = ZZZ3;
    }

    // this += o (add-2008-s) with every special case handled
    HZK_DEV void add(const xyzz_t& o) {
        if (o.is_inf()) return;
        if (is_inf()) { *this = o; return; }
        F U1, U2, S1;
        F::mul3(U1, X, o.ZZ, U2, o.X, ZZ, S1, Y, o.ZZZ);
        F S2, ZZ12, ZZZ12;
        F::mul3(S2, o.Y, ZZZ, ZZ12, ZZ, o.ZZ, ZZZ12, ZZZ, o.ZZZ);
        F P = U2 - U1;
        F R = S2 - S1;
        if (P.is_zero()) {
            if (R.is_zero()) *this = dbl();
            else *this = inf();
            return;
        }
        F PP, RR;
        F::sqr2(PP, P, RR, R);
        F PPP, Q, ZZ3;
        F::mul3(PPP, P, PP, Q, U1, PP, ZZ3, ZZ12, PP);
        F X3 = RR - PPP - Q.dbl();
        F t, SP, ZZZ3;
        F::mul3(t, R, Q - X3, SP, S1, PPP, ZZZ3, ZZZ12, PPP);
        X = X3;
        Y = t - SP;
        ZZ = ZZ3;
        ZZZ = ZZZ3;
    }
};

typedef affine_t<fq_t> g1_affine;
typedef xyzz_t<fq_t> g1_xyzz;
typedef affine_t<fq2_t> g2_affine;
typedef xyzz_t<fq2_t> g2_xyzz;

}  // namespace hzk
