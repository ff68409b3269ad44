// Point decoding and hash-to-curve on the device.
//
//   decompress   32-byte compressed point -> affine; restates helios_frombytes
//                (helios/src/x64/helios_frombytes.cpp:59-109) and selene_frombytes
//                (selene/src/x64/selene_frombytes.cpp:52-111): bit 255 = y parity, x must be
//                canonical, y = sqrt(x^3 - 3x + b) or reject, negate y to match the parity.  The
//                bytes ABI maps 0^32 to the identity like the reference test loader
//                (src/tests.cpp:7479-7496).
//   sswu         simplified SWU map (RFC 9380 6.6.2) with Z = 7 (Helios) and Z = -4 (Selene),
//                restating helios/src/x64/helios_map_to_curve.cpp:98-198 and
//                selene/src/x64/selene_map_to_curve.cpp:59-73: u is read with bit 255 masked and
//                the sign of y follows sgn0(u).
//   sqrt         p = 5 mod 8: beta = z^((p+3)/8), times sqrt(-1) when beta^2 = -z
//                (fp/src/x64/fp_sqrt.cpp:67+); q = 3 mod 4: z^((q+1)/4) (fq/src/x64/fq_sqrt.cpp:45).
#pragma once
#include "curve.h"

namespace hs
{

struct HeliosK
{
    // b (helios_constants.h:41-44)
    static HS_HD void b(fe &r) { set(r, 0xedd73ad4u, 0x22b49ee1u, 0x4113bf35u, 0x043f6f38u, 0x3ebb7b3bu, 0xbe94a76bu, 0xb0ea70b8u, 0x22e8c739u); }
    static HS_HD void sswu_z(fe &r) { set(r, 7, 0, 0, 0, 0, 0, 0, 0); }
    static HS_HD void neg_b_over_a(fe &r) { set(r, 0xa49d139cu, 0x60e6dfa0u, 0x15b13fbcu, 0x016a7a68u, 0x6a3e7e69u, 0x94dc37ceu, 0x3af8d03du, 0x0ba2ed13u); }
    static HS_HD void b_over_za(fe &r) { set(r, 0xc3e98f69u, 0x16ba7256u, 0x2178f6e5u, 0x48f0c9f1u, 0x39f71283u, 0x33e08a50u, 0xf7934ff7u, 0x7e56706au); }
    // fe from eight little-endian 32-bit words
    static HS_HD void set(fe &r, uint32_t a0, uint32_t a1, uint32_t a2, uint32_t a3, uint32_t a4, uint32_t a5, uint32_t a6, uint32_t a7)
    {
        const uint32_t w[8] = {a0, a1, a2, a3, a4, a5, a6, a7};
        fe_from_words(r, w);
    }
};

struct SeleneK
{
    // b (selene_constants.h:40-43)
    static HS_HD void b(fe &r) { HeliosK::set(r, 0xcb5a4558u, 0x68de0983u, 0x06ae900eu, 0xf3944bdfu, 0x95ffe279u, 0x7f51bba5u, 0x695876c1u, 0x70127713u); }
    static HS_HD void sswu_z(fe &r) { HeliosK::set(r, 0x7927c79bu, 0x6eb6d272u, 0xb7656b58u, 0xbf7f782cu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0x7fffffffu); }
    static HS_HD void neg_b_over_a(fe &r) { HeliosK::set(r, 0x43c8c1c8u, 0xcd9f5881u, 0x023a3004u, 0xa686c3f5u, 0x31fff628u, 0x7fc5e937u, 0x7872d240u, 0x255b7d06u); }
    static HS_HD void b_over_za(fe &r) { HeliosK::set(r, 0x50f23072u, 0x3367d620u, 0x408e8c01u, 0x29a1b0fdu, 0xcc7ffd8au, 0x1ff17a4du, 0x9e1cb490u, 0x0956df41u); }
};

// sqrt in Fp; returns false (and 0) for a non-residue
HS_HD bool fe_sqrt_p(fe &r, const fe &z)
{
    const uint32_t e[8] = {0xfffffffeu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0x0fffffffu};
    fe beta, bsq, nz;
    fe_pow<Fp>(beta, z, e);
    fe_sq<Fp>(bsq, beta);
    if (fe_eq<Fp>(bsq, z))
    {
        r = beta;
        return true;
    }
    fe_neg<Fp>(nz, z);
    if (fe_eq<Fp>(bsq, nz))
    {
        fe i;
        HeliosK::set(i, 0x4a0ea0b0u, 0xc4ee1b27u, 0xad2fe478u, 0x2f431806u, 0x3dfbd7a7u, 0x2b4d0099u, 0x4fc1df0bu, 0x2b832480u);
        fe_mul<Fp>(r, beta, i);
        return true;
    }
    fe_zero(r);
    return false;
}

HS_HD bool fe_sqrt_q(fe &r, const fe &z)
{
    const uint32_t e[8] = {0x9e49f1e8u, 0x1badb49cu, 0x2dd95ad6u, 0xefdfde0bu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0x1fffffffu};
    fe y, ysq;
    fe_pow<Fq>(y, z, e);
    fe_sq<Fq>(ysq, y);
    r = y;
    return fe_eq<Fq>(ysq, z);
}

template <class F> HS_HD bool fe_sqrt(fe &r, const fe &z);
template <> HS_HD bool fe_sqrt<Fp>(fe &r, const fe &z) { return fe_sqrt_p(r, z); }
template <> HS_HD bool fe_sqrt<Fq>(fe &r, const fe &z) { return fe_sqrt_q(r, z); }

// y^2 = x^3 - 3x + b
template <class F, class K> HS_HD void curve_rhs(fe &r, const fe &x)
{
    fe x2, x3, t, b;
    fe_sq<F>(x2, x);
    fe_mul<F>(x3, x2, x);
    fe_add<F>(t, x, x);
    fe_add<F>(t, t, x);
    fe_sub<F>(r, x3, t);
    K::b(b);
    fe_add<F>(r, r, b);
}

// compressed bytes (8 LE words) -> affine.  returns 0 ok, 1 identity (0^32), -1 invalid.
template <class F, class K> HS_HD int decompress(aff &out, const uint32_t w[8])
{
    uint32_t orr = 0;
#pragma unroll
    for (int i = 0; i < 8; i++)
        orr |= w[i];
    if (orr == 0)
    {
        fe_zero(out.x);
        fe_zero(out.y);
        return 1;
    }
    const uint32_t parity = w[7] >> 31;
    uint32_t xw[8], xcw[8];
#pragma unroll
    for (int i = 0; i < 8; i++)
        xw[i] = w[i];
    xw[7] &= 0x7fffffffu;
    fe x;
    fe_from_words(x, xw);
    fe_to_words_canon<F>(xcw, x);
#pragma unroll
    for (int i = 0; i < 8; i++)
        if (xcw[i] != xw[i])
            return -1; // x >= modulus
    fe rhs, y;
    curve_rhs<F, K>(rhs, x);
    if (!fe_sqrt<F>(y, rhs))
        return -1;
    if ((uint32_t)fe_is_odd<F>(y) != parity)
        fe_neg<F>(y, y);
    fe_canon<F>(out.y, y);
    out.x = x;
    return 0;
}

// simplified SWU: u (8 LE words, bit 255 ignored) -> affine point
template <class F, class K> HS_HD void sswu(aff &out, const uint32_t w[8])
{
    fe u, t;
    uint32_t uw[8];
#pragma unroll
    for (int i = 0; i < 8; i++)
        uw[i] = w[i];
    uw[7] &= 0x7fffffffu;
    fe_from_words(t, uw);
    fe_canon<F>(u, t);
    fe Z, A, one, u2, Zu2, den, x1, gx1, x2, gx2, y;
    K::sswu_z(Z);
    fe_one(one);
    fe three;
    fe_small(three, 3);
    fe_neg<F>(A, three);
    fe_sq<F>(u2, u);
    fe_mul<F>(Zu2, Z, u2);
    fe_sq<F>(t, Zu2);
    fe_add<F>(den, t, Zu2); // Z^2 u^4 + Z u^2
    if (fe_is_zero<F>(den))
        K::b_over_za(x1);
    else
    {
        fe di, nba;
        fe_inv<F>(di, den);
        fe_add<F>(di, di, one);
        K::neg_b_over_a(nba);
        fe_mul<F>(x1, nba, di);
    }
    // g(x) = x^3 + A x + B
    fe b;
    K::b(b);
    fe_sq<F>(t, x1);
    fe_mul<F>(gx1, t, x1);
    fe_mul<F>(t, A, x1);
    fe_add<F>(gx1, gx1, t);
    fe_add<F>(gx1, gx1, b);
    fe_mul<F>(x2, Zu2, x1);
    fe x;
    if (fe_sqrt<F>(y, gx1))
        x = x1;
    else
    {
        fe_sq<F>(t, x2);
        fe_mul<F>(gx2, t, x2);
        fe_mul<F>(t, A, x2);
        fe_add<F>(gx2, gx2, t);
        fe_add<F>(gx2, gx2, b);
        fe_sqrt<F>(y, gx2);
        x = x2;
    }
    if (fe_is_odd<F>(y) != ((u.v[0] & 1u) != 0))
        fe_neg<F>(y, y);
    fe_canon<F>(out.y, y);
    fe_canon<F>(out.x, x);
}

} // namespace hs
