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

// w -= m when w >= m (8 words)
template <class O> HS_HD void words_sub_mod_once(uint32_t w[8])
{
    uint32_t d[8];
    int64_t br = 0;
#pragma unroll
    for (int i = 0; i < 8; i++)
    {
        int64_t x = (int64_t)w[i] - O::mw(i) + br;
        d[i] = (uint32_t)x;
        br = x >> 32;
    }
    const uint32_t keep = br == 0 ? 0xffffffffu : 0u;
#pragma unroll
    for (int i = 0; i < 8; i++)
        w[i] = (d[i] & keep) | (w[i] & ~keep);
}

// s (any 256-bit value) -> canonical s mod order (8 words), with the reference's window-8
// substitution s - 2^256 when `w8_compat` is set.  O is the field whose modulus is the order.
template <class O> HS_HD void scalar_prepare(uint32_t out[8], const uint32_t s[8], bool w8_compat)
{
#pragma unroll
    for (int i = 0; i < 8; i++)
        out[i] = s[i];
    // s < 2^256 = 2 * order + 2 * (2^255 mod order): two conditional subtractions
    words_sub_mod_once<O>(out);
    words_sub_mod_once<O>(out);
    if (w8_compat && w8_drops_carry(s))
    {
        // out - (2^256 mod order), where 2^256 mod order = 2 * (2^255 mod order)
        int64_t br = 0;
        uint32_t prev = 0;
#pragma unroll
        for (int i = 0; i < 8; i++)
        {
            const uint32_t ci = O::cw(i);
            const uint32_t d = (ci << 1) | (prev >> 31);
            prev = ci;
            int64_t x = (int64_t)out[i] - d + br;
            out[i] = (uint32_t)x;
            br = x >> 32;
        }
        if (br)
        {
            uint64_t t = 0;
#pragma unroll
            for (int i = 0; i < 8; i++)
            {
                t = (uint64_t)out[i] + O::mw(i) + (t >> 32);
                out[i] = (uint32_t)t;
            }
        }
    }
}

// Digit i of the signed width-c recoding of a canonical scalar (< 2^255, so c*W >= 256 digits
// absorb the final carry).  Digits lie in [-(2^(c-1)) + 1, 2^(c-1)].  Encoded as 0 for a zero
// digit, else |d| | (d < 0) << 31.  `carry` is threaded through consecutive calls (i = 0, 1, ...).
HS_HD uint32_t recode_digit(const uint32_t s[8], int c, int i, uint32_t &carry)
{
    const int bit = c * i, wi = bit >> 5, off = bit & 31;
    uint64_t lo = wi < 8 ? s[wi] : 0u;
    uint64_t hi = wi + 1 < 8 ? s[wi + 1] : 0u;
    uint32_t raw = (uint32_t)(((lo | (hi << 32)) >> off) & ((1ull << c) - 1));
    uint32_t v = raw + carry;
    const uint32_t half = 1u << (c - 1);
    if (v > half)
    {
        // v - 2^c: a negative digit, or zero when v == 2^c (all-ones window plus a carry)
        carry = 1;
        const uint32_t mag = (1u << c) - v;
        return mag ? (mag | 0x80000000u) : 0u;
    }
    carry = 0;
    return v;
}

} // namespace hs
