// ORACLE — test infrastructure only (see orc_field.hpp).
//
// BN254 optimal-ate pairing on the host, used to verify Groth16 proofs. The
// reference carries only the constants (ate loop count and final exponent at
// src/host-arith/curve/alt_bn128/alt_bn128_init.cu:250-254; Fq6/Fq12 tower
// commented out at :83-112) and no pairing code, so this follows the textbook
// construction:
//   * Fq12 = Fq[w] / (w^12 - 18 w^6 + 82); Fq2 embeds by u -> w^6 - 9, so
//     w^6 = xi = 9 + u, the twist's non-residue (alt_bn128_init.cu:117).
//   * A G2 point (x, y) on y^2 = x^3 + 3/xi maps to (x w^2, y w^3) on
//     y^2 = x^3 + 3 over Fq12 (D-type sextic twist).
//   * Miller loop over 6x+2 = 0x19d797039be763ba8 with affine line functions,
//     two Frobenius correction steps (Q1 = pi(Q), -Q2 = -pi^2(Q)), then
//     f^((p^12 - 1)/r).
//   * The Frobenius on twist coordinates uses xi^((p-1)/3) and xi^((p-1)/2),
//     the values the reference stores as alt_bn128_twist_mul_by_q_X/Y
//     (alt_bn128_init.cu:121-124).
#include <vector>
#include "orc_curve.hpp"

using namespace orc;

namespace {

struct Fq12 {
    Fq c[12];
    static Fq12 zero() { Fq12 r; for (auto& x : r.c) x = Fq::zero(); return r; }
    static Fq12 one() { Fq12 r = zero(); r.c[0] = Fq::one(); return r; }
    Fq12 operator*(const Fq12& o) const {
        Fq t[23];
        for (auto& x : t) x = Fq::zero();
        for (int i = 0; i < 12; ++i) {
            if (c[i].is_zero()) continue;
            for (int j = 0; j < 12; ++j) t[i + j] += c[i] * o.c[j];
        }
        static const Fq k18 = Fq::from_u64(18), k82 = Fq::from_u64(82);
        for (int k = 22; k >= 12; --k) {
            if (t[k].is_zero()) continue;
            t[k - 6] += t[k] * k18;
            t[k - 12] -= t[k] * k82;
        }
        Fq12 r; for (int i = 0; i < 12; ++i) r.c[i] = t[i];
        return r;
    }
    Fq12 operator+(const Fq12& o) const { Fq12 r; for (int i = 0; i < 12; ++i) r.c[i] = c[i] + o.c[i]; return r; }
    Fq12 operator-(const Fq12& o) const { Fq12 r; for (int i = 0; i < 12; ++i) r.c[i] = c[i] - o.c[i]; return r; }
    bool operator==(const Fq12& o) const { for (int i = 0; i < 12; ++i) if (c[i] != o.c[i]) return false; return true; }
    Fq12 pow(const uint64_t* e, int nlimbs) const {
        Fq12 acc = one(); bool found = false;
        for (int i = nlimbs * 64 - 1; i >= 0; --i) {
            if (found) acc = acc * acc;
            if ((e[i / 64] >> (i % 64)) & 1) { found = true; acc = acc * *this; }
        }
        return acc;
    }
};

Fq12 embed_fq(const Fq& a) { Fq12 r = Fq12::zero(); r.c[0] = a; return r; }
Fq12 embed_fq2(const Fq2& a) {
    static const Fq nine = Fq::from_u64(9);
    Fq12 r = Fq12::zero();
    r.c[0] = a.c0 - a.c1 * nine;
    r.c[6] = a.c1;
    return r;
}
Fq12 w_pow(int k) { Fq12 r = Fq12::zero(); r.c[k] = Fq::one(); return r; }

// (p^12 - 1) / r, little-endian 64-bit words (45 words, 2790 bits); equals
// alt_bn128_final_exponent (alt_bn128_init.cu:252).
std::vector<uint64_t> final_exponent() {
    static const char* hex =
        "2f4b6dc97020fddadf107d20bc842d43bf6369b1ff6a1c71015f3f7be2e1e30a73bb94fec0daf15466b2383a5d3ec3d15ad524d8f70c54efee1bd8c3b21377e563a09a1b705887e72eceaddea3790364a61f676baaf977870e88d5c6c8fef0781361e443ae77f5b63a2a2264487f2940a8b1ddb3d15062cd0fb2015dfc6668449aed3cc48a82d0d602d268c7daab6a41294c0cc4ebe5664568dfc50e1648a45a4a1e3a5195846a3ed011a337a02088ec80e0ebae8755cfe107acf3aafb40494e406f804216bb10cf430b0f37856b42db8dc5514724ee93dfb10826f0dd4a0364b9580291d2cd65664814fde37ca80bb4ea44eacc5e641bbadf423f9a2cbf813b8d145da90029baee7ddadda71c7f3811c4105262945bba1668c3be69a3c230974d83561841d766f9c9d570bb7fbe04c7e8a6c3c760c0de81def35692da361102b6b9b2b918837fa97896e84abb40a4efb7e54523a486964b64ca86f120";
    size_t len = strlen(hex);
    std::vector<uint64_t> w((len + 15) / 16, 0);
    for (size_t i = 0; i < len; ++i) {
        char ch = hex[len - 1 - i];
        uint64_t d = (ch >= '0' && ch <= '9') ? ch - '0' : ch - 'a' + 10;
        w[i / 16] |= d << (4 * (i % 16));
    }
    return w;
}

struct AffG2 { Fq2 x, y; bool inf; };
struct AffG1 { Fq x, y; bool inf; };

Fq2 fq2_pow(const Fq2& a, const uint64_t e[4]) {
    Fq2 acc = Fq2::one(); bool found = false;
    for (int i = 255; i >= 0; --i) {
        if (found) acc = acc.squared();
        if ((e[i / 64] >> (i % 64)) & 1) { found = true; acc = acc * a; }
    }
    return acc;
}
// xi^((p-1)/3) and xi^((p-1)/2)
void frob_consts(Fq2& gx, Fq2& gy) {
    uint64_t pm1[4]; const uint64_t one_raw[4] = {1, 0, 0, 0};
    sub4(pm1, FQ_PARAMS.p, one_raw);
    // (p-1)/2
    uint64_t h[4]; for (int i = 0; i < 4; ++i) h[i] = (pm1[i] >> 1) | (i < 3 ? pm1[i + 1] << 63 : 0);
    // (p-1)/3 by long division
    uint64_t t[4]; u128 rem = 0;
    for (int i = 3; i >= 0; --i) { u128 cur = (rem << 64) | pm1[i]; t[i] = (uint64_t)(cur / 3); rem = cur % 3; }
    Fq2 xi = {Fq::from_u64(9), Fq::from_u64(1)};
    gx = fq2_pow(xi, t);
    gy = fq2_pow(xi, h);
}
AffG2 frobenius(const AffG2& Q) {
    static Fq2 gx, gy; static bool init = false;
    if (!init) { frob_consts(gx, gy); init = true; }
    return {Q.x.conj() * gx, Q.y.conj() * gy, Q.inf};
}

// Line through R and S (twist coordinates) evaluated at P, as an Fq12 element.
// With x = x' w^2, y = y' w^3 the slope is m' w, so
//   l = m' w (xP - x1' w^2) - (yP - y1' w^3).
Fq12 line(const AffG2& R, const AffG2& S, const AffG1& P) {
    Fq12 xP = embed_fq(P.x), yP = embed_fq(P.y);
    Fq12 x1 = embed_fq2(R.x) * w_pow(2), y1 = embed_fq2(R.y) * w_pow(3);
    Fq2 m;
    if (R.x != S.x) {
        m = (S.y - R.y) * (S.x - R.x).inverse();
    } else if (R.y == S.y) {
        Fq2 xx = R.x.squared();
        m = (xx.dbl() + xx) * R.y.dbl().inverse();
    } else {
        return xP - x1;   // vertical line
    }
    Fq12 mw = embed_fq2(m) * w_pow(1);
    return mw * (xP - x1) - (yP - y1);
}
AffG2 add_aff(const AffG2& R, const AffG2& S) {
    if (R.inf) return S;
    if (S.inf) return R;
    Fq2 m;
    if (R.x == S.x) {
        if (R.y != S.y || R.y.is_zero()) return {Fq2::zero(), Fq2::zero(), true};
        Fq2 xx = R.x.squared();
        m = (xx.dbl() + xx) * R.y.dbl().inverse();
    } else {
        m = (S.y - R.y) * (S.x - R.x).inverse();
    }
    Fq2 x3 = m.squared() - R.x - S.x;
    Fq2 y3 = m * (R.x - x3) - R.y;
    return {x3, y3, false};
}

Fq12 miller_loop(const AffG2& Q, const AffG1& P) {
    if (Q.inf || P.inf) return Fq12::one();
    const uint64_t loop = 0x9d797039be763ba8ULL;   // 6x+2 = 0x1_9d797039be763ba8 (65 bits)
    AffG2 R = Q;
    Fq12 f = Fq12::one();
    for (int i = 63; i >= 0; --i) {
        f = f * f * line(R, R, P);
        R = add_aff(R, R);
        if ((loop >> i) & 1) {
            f = f * line(R, Q, P);
            R = add_aff(R, Q);
        }
    }
    AffG2 Q1 = frobenius(Q);
    AffG2 nQ2 = frobenius(Q1); nQ2.y = -nQ2.y;
    f = f * line(R, Q1, P);
    R = add_aff(R, Q1);
    f = f * line(R, nQ2, P);
    return f;
}

Fq12 final_exp(const Fq12& f) {
    static std::vector<uint64_t> e = final_exponent();
    return f.pow(e.data(), (int)e.size());
}

AffG1 g1_aff(const uint64_t* a) {
    AffG1 r{Fq::from_raw(a), Fq::from_raw(a + 4), false};
    r.inf = r.x.is_zero() && r.y.is_zero();
    return r;
}
AffG2 g2_aff(const uint64_t* a) {
    AffG2 r{{Fq::from_raw(a), Fq::from_raw(a + 4)}, {Fq::from_raw(a + 8), Fq::from_raw(a + 12)}, false};
    r.inf = r.x.is_zero() && r.y.is_zero();
    return r;
}

}  // namespace

extern "C" {

// e(P, Q) for affine P (64 B) and Q (128 B); output 12 Fq coefficients (Montgomery) in the w-basis.
int orc_pairing(const uint64_t* p_aff, const uint64_t* q_aff, uint64_t* out12) {
    Fq12 f = final_exp(miller_loop(g2_aff(q_aff), g1_aff(p_aff)));
    for (int i = 0; i < 12; ++i) std::memcpy(out12 + 4 * i, f.c[i].v, 32);
    return 0;
}

// prod_i e(P_i, Q_i) == 1 ?  (one shared final exponentiation)
int orc_pairing_product_is_one(const uint64_t* ps_aff, const uint64_t* qs_aff, size_t n) {
    Fq12 f = Fq12::one();
    for (size_t i = 0; i < n; ++i) f = f * miller_loop(g2_aff(qs_aff + 16 * i), g1_aff(ps_aff + 8 * i));
    return final_exp(f) == Fq12::one() ? 1 : 0;
}

// The Frobenius twist constants xi^((p-1)/3), xi^((p-1)/2) (2 x Fq2 = 128 B).
void orc_twist_frobenius_consts(uint64_t* out) {
    Fq2 gx, gy; frob_consts(gx, gy);
    std::memcpy(out, gx.c0.v, 32); std::memcpy(out + 4, gx.c1.v, 32);
    std::memcpy(out + 8, gy.c0.v, 32); std::memcpy(out + 12, gy.c1.v, 32);
}

}  // extern "C"
