// ORACLE — test infrastructure only (see orc_field.hpp).
//
// alt_bn128 G1 (over Fq, y^2 = x^3 + 3) and G2 (over Fq2, y^2 = x^3 + 3/(9+u))
// in Jacobian coordinates, restating libff's group law:
//   operator+  add-2007-bl       src/host-arith/curve/alt_bn128/alt_bn128_g1.cu:135-191
//   mixed_add  madd-2007-bl      alt_bn128_g1.cu:248-312
//   dbl        dbl-2009-l        alt_bn128_g1.cu:314-347
//   to_affine  (inf -> (0,1,0))  alt_bn128_g1.cu:56-73
//   scalar_mul MSB-first double-and-add  include/host-arith/curve/curve_utils.cuh:29-50
// and the same for G2 (alt_bn128_g2.cu:60-352).
#pragma once
#include "orc_field.hpp"

namespace orc {

template <class F>
struct Jac {
    F X, Y, Z;

    static Jac zero() { return {F::zero(), F::one(), F::zero()}; }
    bool is_zero() const { return Z.is_zero(); }

    Jac dbl() const {
        if (is_zero()) return *this;
        F A = X.squared();
        F B = Y.squared();
        F C = B.squared();
        F D = ((X + B).squared() - A - C).dbl();
        F E = A.dbl() + A;
        F Fv = E.squared();
        F X3 = Fv - D.dbl();
        F C8 = C.dbl().dbl().dbl();
        F Y3 = E * (D - X3) - C8;
        F Z3 = (Y * Z).dbl();
        return {X3, Y3, Z3};
    }

    Jac operator+(const Jac& o) const {
        if (is_zero()) return o;
        if (o.is_zero()) return *this;
        F Z1Z1 = Z.squared(), Z2Z2 = o.Z.squared();
        F U1 = X * Z2Z2, U2 = o.X * Z1Z1;
        F S1 = Y * (o.Z * Z2Z2), S2 = o.Y * (Z * Z1Z1);
        if (U1 == U2 && S1 == S2) return dbl();
        F H = U2 - U1;
        F S2mS1 = S2 - S1;
        F I = H.dbl().squared();
        F J = H * I;
        F r = S2mS1.dbl();
        F V = U1 * I;
        F X3 = r.squared() - J - V.dbl();
        F S1J = S1 * J;
        F Y3 = r * (V - X3) - S1J.dbl();
        F Z3 = ((Z + o.Z).squared() - Z1Z1 - Z2Z2) * H;
        return {X3, Y3, Z3};
    }

    // o must have Z = 1 (affine, finite)
    Jac mixed_add(const Jac& o) const {
        if (is_zero()) return o;
        if (o.is_zero()) return *this;
        F Z1Z1 = Z.squared();
        F U2 = o.X * Z1Z1;
        F S2 = o.Y * (Z * Z1Z1);
        if (X == U2 && Y == S2) return dbl();
        F H = U2 - X;
        F HH = H.squared();
        F I = HH.dbl().dbl();
        F J = H * I;
        F r = (S2 - Y).dbl();
        F V = X * I;
        F X3 = r.squared() - J - V.dbl();
        F Y1J = Y * J;
        F Y3 = r * (V - X3) - Y1J.dbl();
        F Z3 = (Z + H).squared() - Z1Z1 - HH;
        return {X3, Y3, Z3};
    }

    Jac neg() const { return {X, -Y, Z}; }

    // Normalise to Z = 1; infinity becomes (0, 1, 0).
    Jac to_affine() const {
        if (is_zero()) return zero();
        F zi = Z.inverse();
        F zi2 = zi.squared();
        return {X * zi2, Y * (zi2 * zi), F::one()};
    }

    bool operator==(const Jac& o) const {
        if (is_zero()) return o.is_zero();
        if (o.is_zero()) return false;
        F Z1s = Z.squared(), Z2s = o.Z.squared();
        if (X * Z2s != o.X * Z1s) return false;
        return Y * (o.Z * Z2s) == o.Y * (Z * Z1s);
    }

    // MSB-first double-and-add over a canonical little-endian scalar (4 limbs).
    Jac scalar_mul(const uint64_t k[4]) const {
        Jac acc = zero();
        bool found = false;
        for (int i = 255; i >= 0; --i) {
            if (found) acc = acc.dbl();
            if ((k[i / 64] >> (i % 64)) & 1) { found = true; acc = acc + *this; }
        }
        return acc;
    }
};

typedef Jac<Fq> G1;
typedef Jac<Fq2> G2;

inline Fq g1_b() { return Fq::from_u64(3); }
// b' = 3 / (9 + u)  (alt_bn128_init.cu:116-118)
inline Fq2 g2_b() {
    Fq2 xi = {Fq::from_u64(9), Fq::from_u64(1)};
    return Fq2{Fq::from_u64(3), Fq::zero()} * xi.inverse();
}

// Generators (alt_bn128_init.cu:130-132, :191-195).
inline G1 g1_generator() { return {Fq::from_u64(1), Fq::from_u64(2), Fq::one()}; }
inline G2 g2_generator() {
    return {{Fq::from_hex("1800deef121f1e76426a00665e5c4479674322d4f75edadd46debd5cd992f6ed"),
             Fq::from_hex("198e9393920d483a7260bfb731fb5d25f1aa493335a9e71297e485b7aef312c2")},
            {Fq::from_hex("12c85ea5db8c6deb4aab71808dcb408fe3d1e7690c43d37b4ce6cc0166fa7daa"),
             Fq::from_hex("090689d0585ff075ec9e99ad690c3395bc4b313370b38ef355acdadcd122975b")},
            Fq2::one()};
}

template <class F>
inline bool on_curve(const Jac<F>& P, const F& b) {
    if (P.is_zero()) return true;
    F X2 = P.X.squared(), Y2 = P.Y.squared(), Z2 = P.Z.squared();
    F Z6 = (P.Z * Z2).squared();
    return Y2 == P.X * X2 + b * Z6;
}
inline bool on_curve_g1(const G1& P) { return on_curve(P, g1_b()); }
inline bool on_curve_g2(const G2& P) {
    if (P.is_zero()) return true;
    Fq2 X2 = P.X.squared(), Y2 = P.Y.squared(), Z2 = P.Z.squared();
    Fq2 Z6 = (P.Z * Z2).squared();
    return Y2 == P.X * X2 + g2_b() * Z6;
}

}  // namespace orc
