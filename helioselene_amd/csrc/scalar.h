// Scalar preparation and signed-window recoding for the Pippenger path.
//
// Reference behaviour restated (helios/src/x64/helios_msm_vartime.cpp):
//   * The reference recodes every scalar into ceil(256/w) signed w-bit digits with
//     encode_signed_wbit (:127-162) and never keeps the carry out of the last digit.  Only for
//     w = 8 (chosen for 864 <= n < 2592, :248-263) can that carry be non-zero, and then the
//     reference effectively uses s - 2^256 for that scalar.  In reference-compatible mode (the
//     default) the GPU applies exactly that substitution, so results match the reference bytes;
//     in exact mode it computes sum s_i P_i.
//   * Everything else is exact group arithmetic, so any correct recoding gives the same point.
//     Scalars are reduced modulo the (prime) group order first, which leaves every group element
//     unchanged, and are then recoded with this build's own window width c.
#pragma once
#include "field.h"

namespace hs
{

// 1 if the reference's width-8 recode of these 32 bytes drops its final carry.
HS_HD uint32_t w8_drops_carry(const uint32_t s[8])
{
    uint32_t carry = 0;
#pragma unroll
    for (int i = 0; i < 32; i++)
    {
        uint32_t v = ((s[i >> 2] >> (8 * (i & 3))) & 0xffu) + carry;
        carry = v >= 128u ? 1u : 0u;
    }
    return carry;
}

// s (any 256-bit value) -> canonical s mod order, with the reference's window-8 substitution when
// `w8_compat` is set.  O is the field whose modulus is the group order.
template <class O> HS_HD void scalar_prepare(fe &out, const uint32_t s[8], bool w8_compat)
{
    fe v;
    fe_from_words(v, s);
    fe_canon<O>(out, v);
    if (w8_compat && w8_drops_carry(s))
    {
        // s - 2^256 = s - D (mod order)
        fe d;
        fe_zero(d);
        d.v[0] = O::D0;
        d.v[1] = O::D1;
        d.v[2] = O::D2;
        d.v[3] = O::D3;
        fe t;
        fe_sub<O>(t, out, d);
        fe_canon<O>(out, t);
    }
}

// Digit i of the signed width-c recoding of a canonical scalar (< 2^255, so c*W >= 256 digits
// absorb the final carry).  Digits lie in [-(2^(c-1)) + 1, 2^(c-1)].  Encoded as 0 for a zero
// digit, else |d| | (d < 0) << 31.  `carry` is threaded through consecutive calls (i = 0, 1, ...).
HS_HD uint32_t recode_digit(const fe &s, int c, int i, uint32_t &carry)
{
    const int bit = c * i, wi = bit >> 5, off = bit & 31;
    uint64_t lo = wi < 8 ? s.v[wi] : 0u;
    uint64_t hi = wi + 1 < 8 ? s.v[wi + 1] : 0u;
    uint32_t raw = (uint32_t)(((lo | (hi << 32)) >> off) & ((1ull << c) - 1));
    uint32_t v = raw + carry;
    const uint32_t half = 1u << (c - 1);
    if (v > half)
    {
        carry = 1;
        return ((1u << c) - v) | 0x80000000u;
    }
    carry = 0;
    return v;
}

} // namespace hs
