// 255-bit prime-field arithmetic for the two Helios/Selene coordinate fields, written for
// CDNA4 VALU: eight 32-bit limbs, 32x32->64 partial products on v_mad_u64_u32 chains, and a
// dedicated reduction per modulus (no MFMA: this is exact integer work).
//
//   Fp: p = 2^255 - 19                (reference fp/include/x64/fp51_inline.h:46-243, fp_ops.h:42-91)
//   Fq: q = 2^255 - gamma, gamma < 2^127 (reference fq/include/x64/fq51.h:37-86,
//                                          fq51_inline.h:1086-1225, fq_ops.h:50-101)
//
// Representation: a value is any 256-bit integer v < 2^256 standing for v mod m ("redundant"
// form).  Since 2^256 = D (mod m) with D = 38 for p and D = 2*gamma (128 bits) for q, every
// operation folds its overflow back with D.  Only canonical() maps to [0, m); equality and
// zero tests canonicalise first.  The reference keeps lazy radix-2^51 limbs instead; both denote
// the same residues, so encoded outputs are identical.
#pragma once
#include <stdint.h>

#ifndef HS_HD
#define HS_HD __host__ __device__ __forceinline__
#endif

namespace hs
{

struct fe
{
    uint32_t v[8];
};

struct Fp
{
    // p = 2^255 - 19
    static constexpr uint32_t M0 = 0xffffffedu, M1 = 0xffffffffu, M2 = 0xffffffffu, M3 = 0xffffffffu;
    static constexpr uint32_t M4 = 0xffffffffu, M5 = 0xffffffffu, M6 = 0xffffffffu, M7 = 0x7fffffffu;
    // D = 2^256 mod p (one limb), C = 2^255 mod p
    static constexpr int DLIMBS = 1;
    static constexpr uint32_t D0 = 38, D1 = 0, D2 = 0, D3 = 0;
    static constexpr uint32_t C0 = 19, C1 = 0, C2 = 0, C3 = 0;
};

struct Fq
{
    // q = 2^255 - gamma, gamma = 0x408087d3489a94a791492d8d86d83861 (reference fq51.h:39-62)
    static constexpr uint32_t M0 = 0x7927c79fu, M1 = 0x6eb6d272u, M2 = 0xb7656b58u, M3 = 0xbf7f782cu;
    static constexpr uint32_t M4 = 0xffffffffu, M5 = 0xffffffffu, M6 = 0xffffffffu, M7 = 0x7fffffffu;
    static constexpr int DLIMBS = 4;
    // D = 2*gamma = 0x81010fa69135294f22925b1b0db070c2
    static constexpr uint32_t D0 = 0x0db070c2u, D1 = 0x22925b1bu, D2 = 0x9135294fu, D3 = 0x81010fa6u;
    // C = gamma
    static constexpr uint32_t C0 = 0x86d83861u, C1 = 0x91492d8du, C2 = 0x489a94a7u, C3 = 0x408087d3u;
};

template <class F> HS_HD uint32_t mod_limb(int i)
{
    switch (i)
    {
    case 0: return F::M0;
    case 1: return F::M1;
    case 2: return F::M2;
    case 3: return F::M3;
    case 4: return F::M4;
    case 5: return F::M5;
    case 6: return F::M6;
    default: return F::M7;
    }
}

template <class F> HS_HD uint32_t d_limb(int i)
{
    switch (i)
    {
    case 0: return F::D0;
    case 1: return F::D1;
    case 2: return F::D2;
    case 3: return F::D3;
    default: return 0;
    }
}

template <class F> HS_HD uint32_t c_limb(int i)
{
    switch (i)
    {
    case 0: return F::C0;
    case 1: return F::C1;
    case 2: return F::C2;
    case 3: return F::C3;
    default: return 0;
    }
}

HS_HD uint64_t mad64(uint32_t a, uint32_t b, uint64_t c) { return (uint64_t)a * b + c; }

HS_HD void fe_zero(fe &r)
{
#pragma unroll
    for (int i = 0; i < 8; i++)
        r.v[i] = 0;
}

HS_HD void fe_one(fe &r)
{
    r.v[0] = 1;
#pragma unroll
    for (int i = 1; i < 8; i++)
        r.v[i] = 0;
}

// r += k*D (mod 2^256) for a small multiplier k (k*D*... < 2^32 per limb product); returns the
// carry out of bit 256.
template <class F> HS_HD uint32_t add_kd(fe &r, uint32_t k)
{
    uint64_t c = 0;
#pragma unroll
    for (int i = 0; i < 8; i++)
    {
        if (i < F::DLIMBS)
            c = mad64(k, d_limb<F>(i), (uint64_t)r.v[i] + (c >> 32));
        else
            c = (uint64_t)r.v[i] + (c >> 32);
        r.v[i] = (uint32_t)c;
    }
    return (uint32_t)(c >> 32);
}

// Final fold shared by every operation: value = r + carry*2^256 with carry small; result < 2^256.
template <class F> HS_HD void fold_carry(fe &r, uint32_t carry)
{
    uint32_t c2 = add_kd<F>(r, carry);
    // A second carry leaves r < k*D, so adding D once more cannot overflow.
    add_kd<F>(r, c2);
}

template <class F> HS_HD void fe_add(fe &r, const fe &a, const fe &b)
{
    uint64_t c = 0;
#pragma unroll
    for (int i = 0; i < 8; i++)
    {
        c = (uint64_t)a.v[i] + b.v[i] + (c >> 32);
        r.v[i] = (uint32_t)c;
    }
    fold_carry<F>(r, (uint32_t)(c >> 32));
}

// r -= k*D (mod 2^256); returns the borrow out.
template <class F> HS_HD uint32_t sub_kd(fe &r, uint32_t k)
{
    int64_t b = 0;
#pragma unroll
    for (int i = 0; i < 8; i++)
    {
        uint64_t sub = (i < F::DLIMBS) ? (uint64_t)k * d_limb<F>(i) : 0;
        int64_t t = (int64_t)r.v[i] - (int64_t)(sub & 0xffffffffu) + b;
        r.v[i] = (uint32_t)t;
        b = (t >> 32) - (int64_t)(sub >> 32);
    }
    return (uint32_t)(-b);
}

template <class F> HS_HD void fe_sub(fe &r, const fe &a, const fe &b)
{
    int64_t br = 0;
#pragma unroll
    for (int i = 0; i < 8; i++)
    {
        int64_t t = (int64_t)a.v[i] - b.v[i] + br;
        r.v[i] = (uint32_t)t;
        br = t >> 32;
    }
    // a wrapped borrow added 2^256 = D (mod m): take D back off, at most twice
    uint32_t b1 = (uint32_t)(-br);
    uint32_t b2 = sub_kd<F>(r, b1);
    sub_kd<F>(r, b2);
}

template <class F> HS_HD void fe_neg(fe &r, const fe &a)
{
    fe z;
    fe_zero(z);
    fe_sub<F>(r, z, a);
}

// r = a * 2 (a + a)
template <class F> HS_HD void fe_dbl(fe &r, const fe &a) { fe_add<F>(r, a, a); }

// Reduce a 512-bit product t[16] to < 2^256 using 2^256 = D (mod m).
template <class F> HS_HD void reduce512(fe &r, const uint32_t t[16])
{
    if (F::DLIMBS == 1)
    {
        // r = t_lo + D * t_hi  (D = 38)
        uint64_t c = 0;
#pragma unroll
        for (int j = 0; j < 8; j++)
        {
            c = mad64(t[8 + j], F::D0, (uint64_t)t[j] + (c >> 32));
            r.v[j] = (uint32_t)c;
        }
        fold_carry<F>(r, (uint32_t)(c >> 32));
    }
    else
    {
        // first fold: w = t_lo + t_hi * D, 12 limbs (t_hi: 8 limbs, D: 4 limbs)
        uint32_t w[12];
#pragma unroll
        for (int j = 0; j < 12; j++)
            w[j] = j < 8 ? t[j] : 0;
#pragma unroll
        for (int i = 0; i < 4; i++)
        {
            uint64_t c = 0;
#pragma unroll
            for (int j = 0; j < 8; j++)
            {
                c = mad64(t[8 + j], d_limb<F>(i), (uint64_t)w[i + j] + (c >> 32));
                w[i + j] = (uint32_t)c;
            }
            // propagate the row carry into the remaining limbs
            uint64_t cc = c >> 32;
#pragma unroll
            for (int k = i + 8; k < 12; k++)
            {
                cc += w[k];
                w[k] = (uint32_t)cc;
                cc >>= 32;
            }
        }
        // second fold: r = w_lo + w_hi * D, w_hi = w[8..11] (4 limbs)
        uint32_t x[9];
#pragma unroll
        for (int j = 0; j < 8; j++)
            x[j] = w[j];
        x[8] = 0;
#pragma unroll
        for (int i = 0; i < 4; i++)
        {
            uint64_t c = 0;
#pragma unroll
            for (int j = 0; j < 4; j++)
            {
                c = mad64(w[8 + j], d_limb<F>(i), (uint64_t)x[i + j] + (c >> 32));
                x[i + j] = (uint32_t)c;
            }
            uint64_t cc = c >> 32;
#pragma unroll
            for (int k = i + 4; k < 9; k++)
            {
                cc += x[k];
                x[k] = (uint32_t)cc;
                cc >>= 32;
            }
        }
#pragma unroll
        for (int j = 0; j < 8; j++)
            r.v[j] = x[j];
        fold_carry<F>(r, x[8]);
    }
}

template <class F> HS_HD void fe_mul(fe &r, const fe &a, const fe &b)
{
    uint32_t t[16];
    {
        uint64_t c = 0;
#pragma unroll
        for (int j = 0; j < 8; j++)
        {
            c = mad64(a.v[0], b.v[j], c >> 32);
            t[j] = (uint32_t)c;
        }
        t[8] = (uint32_t)(c >> 32);
    }
#pragma unroll
    for (int i = 1; i < 8; i++)
    {
        uint64_t c = 0;
#pragma unroll
        for (int j = 0; j < 8; j++)
        {
            c = mad64(a.v[i], b.v[j], (uint64_t)t[i + j] + (c >> 32));
            t[i + j] = (uint32_t)c;
        }
        t[i + 8] = (uint32_t)(c >> 32);
    }
    reduce512<F>(r, t);
}

template <class F> HS_HD void fe_sq(fe &r, const fe &a)
{
    uint32_t t[16];
#pragma unroll
    for (int i = 0; i < 16; i++)
        t[i] = 0;
    // off-diagonal products a_i a_j, i < j
#pragma unroll
    for (int i = 0; i < 7; i++)
    {
        uint64_t c = 0;
#pragma unroll
        for (int j = i + 1; j < 8; j++)
        {
            c = mad64(a.v[i], a.v[j], (uint64_t)t[i + j] + (c >> 32));
            t[i + j] = (uint32_t)c;
        }
        t[i + 8] = (uint32_t)(c >> 32);
    }
    // double
    t[15] = t[14] >> 31;
#pragma unroll
    for (int i = 14; i > 0; i--)
        t[i] = (t[i] << 1) | (t[i - 1] >> 31);
    t[0] <<= 1;
    // add the squares a_i^2 at limb 2i
    uint64_t c = 0;
#pragma unroll
    for (int i = 0; i < 8; i++)
    {
        uint64_t s = (uint64_t)a.v[i] * a.v[i];
        c = (uint64_t)t[2 * i] + (uint32_t)s + (c >> 32);
        t[2 * i] = (uint32_t)c;
        c = (uint64_t)t[2 * i + 1] + (uint32_t)(s >> 32) + (c >> 32);
        t[2 * i + 1] = (uint32_t)c;
    }
    reduce512<F>(r, t);
}

// small-constant multiply r = a * k (k < 2^32)
template <class F> HS_HD void fe_mul_small(fe &r, const fe &a, uint32_t k)
{
    uint64_t c = 0;
#pragma unroll
    for (int i = 0; i < 8; i++)
    {
        c = mad64(a.v[i], k, c >> 32);
        r.v[i] = (uint32_t)c;
    }
    fold_carry<F>(r, (uint32_t)(c >> 32));
}

// Canonical residue in [0, m): v < 2^256 < 2m + D, so at most two subtractions of m.
template <class F> HS_HD void fe_canon(fe &r, const fe &a)
{
    r = a;
#pragma unroll
    for (int pass = 0; pass < 2; pass++)
    {
        // t = r - m; keep it when there is no borrow
        uint32_t t[8];
        int64_t br = 0;
#pragma unroll
        for (int i = 0; i < 8; i++)
        {
            int64_t x = (int64_t)r.v[i] - mod_limb<F>(i) + br;
            t[i] = (uint32_t)x;
            br = x >> 32;
        }
        uint32_t keep = br == 0 ? 0xffffffffu : 0u;
#pragma unroll
        for (int i = 0; i < 8; i++)
            r.v[i] = (t[i] & keep) | (r.v[i] & ~keep);
    }
}

template <class F> HS_HD bool fe_is_zero(const fe &a)
{
    fe c;
    fe_canon<F>(c, a);
    uint32_t o = 0;
#pragma unroll
    for (int i = 0; i < 8; i++)
        o |= c.v[i];
    return o == 0;
}

template <class F> HS_HD bool fe_eq(const fe &a, const fe &b)
{
    fe d;
    fe_sub<F>(d, a, b);
    return fe_is_zero<F>(d);
}

// Exponentiation by a fixed 256-bit exponent given as 8 little-endian 32-bit words, 4-bit
// fixed windows (255 squarings + 64 multiplications).
template <class F> HS_HD void fe_pow(fe &r, const fe &a, const uint32_t e[8])
{
    fe tab[16];
    fe_one(tab[0]);
    tab[1] = a;
#pragma unroll
    for (int i = 2; i < 16; i++)
        fe_mul<F>(tab[i], tab[i - 1], a);
    fe acc;
    fe_one(acc);
#pragma unroll 1
    for (int w = 63; w >= 0; w--)
    {
        if (w != 63)
        {
            fe_sq<F>(acc, acc);
            fe_sq<F>(acc, acc);
            fe_sq<F>(acc, acc);
            fe_sq<F>(acc, acc);
        }
        uint32_t nib = (e[w >> 3] >> ((w & 7) * 4)) & 15u;
        fe t;
        // select without a data-dependent index into registers
        t = tab[0];
#pragma unroll
        for (int k = 1; k < 16; k++)
            if (nib == (uint32_t)k)
                t = tab[k];
        fe_mul<F>(acc, acc, t);
    }
    r = acc;
}

// m - 2 as an exponent
template <class F> HS_HD void exp_m_minus_2(uint32_t e[8])
{
#pragma unroll
    for (int i = 0; i < 8; i++)
        e[i] = mod_limb<F>(i);
    e[0] -= 2; // both moduli have a low limb >= 2
}

// a^-1 (Fermat, a^(m-2)); maps 0 to 0 like the reference's fp_invert / fq_invert.
template <class F> HS_HD void fe_inv(fe &r, const fe &a)
{
    uint32_t e[8];
    exp_m_minus_2<F>(e);
    fe_pow<F>(r, a, e);
}

template <class F> HS_HD bool fe_is_odd(const fe &a)
{
    fe c;
    fe_canon<F>(c, a);
    return c.v[0] & 1;
}

// --------------------------------------------------------------- conversions
// 32-byte little-endian -> fe (value < 2^256, not reduced)
HS_HD void fe_from_words(fe &r, const uint32_t w[8])
{
#pragma unroll
    for (int i = 0; i < 8; i++)
        r.v[i] = w[i];
}

// radix-2^51 limbs (the reference's fp_fe / fq_fe, uint64_t[5], limbs may exceed 51 bits) ->
// reduced fe.  value = sum l_i 2^(51 i) < 2^(204 + 64).
template <class F> HS_HD void fe_from_r51(fe &r, const uint64_t l[5])
{
    uint32_t w[10];
#pragma unroll
    for (int i = 0; i < 10; i++)
        w[i] = 0;
#pragma unroll
    for (int i = 0; i < 5; i++)
    {
        const int bit = 51 * i, lw = bit >> 5, sh = bit & 31;
        // add l_i << sh into words lw, lw+1, lw+2 (and carry on)
        uint64_t lo = l[i] << sh;                        // low 64 bits of the shifted limb
        uint64_t hi = sh ? (l[i] >> (64 - sh)) : 0;      // bits above 64
        uint64_t c = (uint64_t)w[lw] + (uint32_t)lo;
        w[lw] = (uint32_t)c;
        c = (uint64_t)w[lw + 1] + (uint32_t)(lo >> 32) + (c >> 32);
        w[lw + 1] = (uint32_t)c;
        c >>= 32;
        c += hi;
#pragma unroll
        for (int k = lw + 2; k < 10; k++)
        {
            c += w[k];
            w[k] = (uint32_t)c;
            c >>= 32;
        }
    }
    // value = w[0..7] + 2^256 * (w[8] + 2^32 w[9]); 2^256 = D, and w[8..9] < 2^13
    fe lo;
#pragma unroll
    for (int i = 0; i < 8; i++)
        lo.v[i] = w[i];
    // w[9] is zero: the input is below 2^268
    fold_carry<F>(lo, w[8]);
    r = lo;
}

// fe -> canonical radix-2^51 limbs
template <class F> HS_HD void fe_to_r51(uint64_t l[5], const fe &a)
{
    fe c;
    fe_canon<F>(c, a);
    uint64_t w[4];
#pragma unroll
    for (int i = 0; i < 4; i++)
        w[i] = (uint64_t)c.v[2 * i] | ((uint64_t)c.v[2 * i + 1] << 32);
    const uint64_t M = (1ull << 51) - 1;
    l[0] = w[0] & M;
    l[1] = ((w[0] >> 51) | (w[1] << 13)) & M;
    l[2] = ((w[1] >> 38) | (w[2] << 26)) & M;
    l[3] = ((w[2] >> 25) | (w[3] << 39)) & M;
    l[4] = (w[3] >> 12) & M;
}

} // namespace hs
