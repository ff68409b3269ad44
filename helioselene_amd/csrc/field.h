// 255-bit prime-field arithmetic for the two Helios/Selene coordinate fields, written for
// CDNA4 VALU (no MFMA: this is exact integer work).
//
//   Fp: p = 2^255 - 19                (reference fp/include/x64/fp51_inline.h:46-243, fp_ops.h:42-91)
//   Fq: q = 2^255 - gamma, gamma < 2^127 (reference fq/include/x64/fq51.h:37-86,
//                                          fq51_inline.h:1086-1225, fq_ops.h:50-101)
//
// Representation: nine 29-bit limbs, value = sum v[i] 2^(29 i) < 2^261 + small, standing for
// value mod m.  "Normalised" means every limb <= 2^29.  A product is formed column by column:
// each of the 81 partial products a_i * b_j is one v_mad_u64_u32 accumulating straight into a
// 64-bit column register (9 * 2^58 < 2^62, so a column never overflows and needs no carry
// handling until the end).  The reduction then uses 2^261 = R (mod m): R = 1216 for p
// (2^261 = 2^6 * 19) and R = 64 * gamma (133 bits, five limbs) for q -- the Crandall fold.
// Only fe_to_words_canon() produces the canonical residue in [0, m) (8 x 32-bit words); equality
// and zero tests go through it.  The reference keeps lazy radix-2^51 limbs instead; both denote
// the same residues, so encoded outputs are identical.
#pragma once
#include <stdint.h>

#ifndef HS_HD
#define HS_HD __host__ __device__ __forceinline__
#endif

namespace hs
{

constexpr int NL = 9;
constexpr uint32_t M29 = 0x1fffffffu;

struct fe
{
    uint32_t v[NL];
};

HS_HD uint64_t mad64(uint32_t a, uint32_t b, uint64_t c) { return (uint64_t)a * b + c; }


struct Fp
{
    static constexpr int NR = 1; // limbs of R = 2^261 mod p
    static HS_HD uint32_t mod(int i) { return i == 0 ? 0x1fffffedu : i == 8 ? 0x7fffffu : 0x1fffffffu; }
    static HS_HD uint32_t r(int i) { return i == 0 ? 1216u : 0u; }
    static HS_HD uint32_t g(int i) { return i == 0 ? 19u : 0u; }
    // 129 * p with every limb in [2^30, 2^31): subtraction bias
    static HS_HD uint32_t bias(int i)
    {
        return i == 0 ? 0x5ffff66du : i == 8 ? 0x407ffffdu : 0x5ffffffdu;
    }
    // 257 p with every limb in [2^31, 2^31 + 2^29): bias of a subtraction of three normalised values
    static HS_HD uint32_t bias3(int i) { return i == 0 ? 0x9fffecedu : i == 8 ? 0x807ffffbu : 0x9ffffffbu; }
    // 65 p with every limb in [2^29, 2^30): bias of the small-limb lazy forms
    static HS_HD uint32_t bias2(int i) { return i == 0 ? 0x3ffffb2du : i == 8 ? 0x207ffffeu : 0x3ffffffeu; }
    // modulus and 2^255 mod m as 32-bit words (canonicalisation)
    static HS_HD uint32_t mw(int i) { return i == 0 ? 0xffffffedu : i == 7 ? 0x7fffffffu : 0xffffffffu; }
    static HS_HD uint32_t cw(int i) { return i == 0 ? 19u : 0u; }
    static constexpr int NCW = 1;
};

struct Fq
{
    static constexpr int NR = 5;
    static HS_HD uint32_t mod(int i)
    {
        switch (i)
        {
        case 0: return 0x1927c79fu;
        case 1: return 0x15b69393u;
        case 2: return 0x195ad61bu;
        case 3: return 0x1ef0596eu;
        case 4: return 0x1ffffbf7u;
        case 8: return 0x7fffffu;
        default: return 0x1fffffffu;
        }
    }
    // R = 2^261 mod q = 64 * gamma
    static HS_HD uint32_t r(int i)
    {
        switch (i)
        {
        case 0: return 0x160e1840u;
        case 1: return 0x125b1b0du;
        case 2: return 0x094a7914u;
        case 3: return 0x03e9a44du;
        case 4: return 0x00010202u;
        default: return 0u;
        }
    }
    // gamma = 2^255 mod q in 29-bit limbs (127 bits)
    static HS_HD uint32_t g(int i)
    {
        switch (i)
        {
        case 0: return 0x06d83861u;
        case 1: return 0x0a496c6cu;
        case 2: return 0x06a529e4u;
        case 3: return 0x010fa691u;
        case 4: return 0x00000408u;
        default: return 0u;
        }
    }
    static HS_HD uint32_t bias(int i)
    {
        switch (i)
        {
        case 0: return 0x4d0b971fu;
        case 1: return 0x51005d76u;
        case 2: return 0x46c5e3f0u;
        case 3: return 0x571d10d2u;
        case 4: return 0x5ffdf7f1u;
        case 8: return 0x407ffffdu;
        default: return 0x5ffffffdu;
        }
    }
    static HS_HD uint32_t bias3(int i)
    {
        switch (i)
        {
        case 0: return 0x80ef669fu;
        case 1: return 0x8c4a2759u;
        case 2: return 0x9430f1c5u;
        case 3: return 0x8f49c835u;
        case 4: return 0x9ffbf3ebu;
        case 8: return 0x807ffffbu;
        default: return 0x9ffffffbu;
        }
    }
    // 65 q with every limb in [2^29, 2^30)
    static HS_HD uint32_t bias2(int i)
    {
        switch (i)
        {
        case 0: return 0x2319af5fu;
        case 1: return 0x235b7885u;
        case 2: return 0x30105d06u;
        case 3: return 0x3b06b520u;
        case 4: return 0x3ffef9f4u;
        case 8: return 0x207ffffeu;
        default: return 0x3ffffffeu;
        }
    }
    // q = 2^255 - gamma, gamma = 0x408087d3489a94a791492d8d86d83861 (reference fq51.h:39-62)
    static HS_HD uint32_t mw(int i)
    {
        switch (i)
        {
        case 0: return 0x7927c79fu;
        case 1: return 0x6eb6d272u;
        case 2: return 0xb7656b58u;
        case 3: return 0xbf7f782cu;
        case 7: return 0x7fffffffu;
        default: return 0xffffffffu;
        }
    }
    static HS_HD uint32_t cw(int i)
    {
        switch (i)
        {
        case 0: return 0x86d83861u;
        case 1: return 0x91492d8du;
        case 2: return 0x489a94a7u;
        case 3: return 0x408087d3u;
        default: return 0u;
        }
    }
    static constexpr int NCW = 4;
};

// Hides what the compiler knows about the bits of a value's limbs.  Given visible limb bounds it
// can split a following product into 32-bit carry chains with many extra register moves.
HS_HD void fe_opaque(fe &a)
{
#ifdef __HIP_DEVICE_COMPILE__
#pragma unroll
    for (int i = 0; i < NL; i++)
        asm volatile("" : "+v"(a.v[i]));
#else
    (void)a;
#endif
}

HS_HD void fe_zero(fe &r)
{
#pragma unroll
    for (int i = 0; i < NL; i++)
        r.v[i] = 0;
}

HS_HD void fe_one(fe &r)
{
    r.v[0] = 1;
#pragma unroll
    for (int i = 1; i < NL; i++)
        r.v[i] = 0;
}

HS_HD void fe_small(fe &r, uint32_t k)
{
    fe_zero(r);
    r.v[0] = k;
}

// r (limbs < 2^29, plus the value c * 2^261 carried out of the top) -> normalised.
// Adds c * R, c small (< 2^14).
template <class F> HS_HD void fold_top(fe &r, uint32_t c)
{
    if (F::NR == 1)
    {
        uint32_t t = r.v[0] + c * F::r(0);
        r.v[0] = t & M29;
        r.v[1] += t >> 29; // r.v[1] <= 2^29 afterwards
    }
    else
    {
        // fold everything from bit 255 up: value = lo + 2^255 top = lo + gamma top with
        // top = bits 255.. of limb 8 plus c * 2^6 (< 2^20), so the sum stays below
        // 2^255 + 2^147 and one carry pass leaves every limb < 2^29
        const uint32_t top = (r.v[NL - 1] >> 23) + (c << 6);
        r.v[NL - 1] &= (1u << 23) - 1;
        uint64_t u = 0;
#pragma unroll
        for (int i = 0; i < 5; i++)
        {
            u = mad64(top, F::g(i), (uint64_t)r.v[i] + (u >> 29));
            r.v[i] = (uint32_t)u & M29;
        }
        uint32_t cc = (uint32_t)(u >> 29);
#pragma unroll
        for (int i = 5; i < NL; i++)
        {
            const uint32_t t = r.v[i] + cc;
            r.v[i] = t & M29;
            cc = t >> 29;
        }
    }
}

// limbs < 2^31.7 -> normalised
template <class F> HS_HD void fe_norm(fe &r)
{
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < NL; i++)
    {
        uint32_t t = r.v[i] + c;
        r.v[i] = t & M29;
        c = t >> 29;
    }
    fold_top<F>(r, c);
}

template <class F> HS_HD void fe_add(fe &r, const fe &a, const fe &b)
{
#pragma unroll
    for (int i = 0; i < NL; i++)
        r.v[i] = a.v[i] + b.v[i];
    fe_norm<F>(r);
}

template <class F> HS_HD void fe_sub(fe &r, const fe &a, const fe &b)
{
#pragma unroll
    for (int i = 0; i < NL; i++)
        r.v[i] = a.v[i] + F::bias(i) - b.v[i];
    fe_norm<F>(r);
}

// r = a - b - c - d for normalised inputs, one normalisation (bias 257 m, limbs < 2^31.6)
template <class F> HS_HD void fe_sub3(fe &r, const fe &a, const fe &b, const fe &c, const fe &d)
{
#pragma unroll
    for (int i = 0; i < NL; i++)
        r.v[i] = a.v[i] + F::bias3(i) - b.v[i] - c.v[i] - d.v[i];
    fe_norm<F>(r);
}

// r = a - b - c for normalised inputs, one normalisation
template <class F> HS_HD void fe_sub2(fe &r, const fe &a, const fe &b, const fe &c)
{
#pragma unroll
    for (int i = 0; i < NL; i++)
        r.v[i] = a.v[i] + F::bias3(i) - b.v[i] - c.v[i];
    fe_norm<F>(r);
}

// r = a - b without normalisation: limbs < 2^29 + 2^31, value < 2^262.3.  Only valid as one
// operand of fe_mul whose other operand is normalised (column sums stay below 2^63.6, inside the
// bounds reduce_cols is written for); never as an input of fe_sq, fe_add or fe_sub.
template <class F> HS_HD void fe_sub_lazy(fe &r, const fe &a, const fe &b)
{
#pragma unroll
    for (int i = 0; i < NL; i++)
        r.v[i] = a.v[i] + F::bias(i) - b.v[i];
}

// r = a + b without normalisation (limbs <= 2^30): as one operand of fe_mul only, like fe_sub_lazy
template <class F> HS_HD void fe_add_lazy(fe &r, const fe &a, const fe &b)
{
#pragma unroll
    for (int i = 0; i < NL; i++)
        r.v[i] = a.v[i] + b.v[i];
}

// Small-limb lazy forms for fe_mul_sum2: r = a - b (limbs < 2^29 + 2^30) and r = -a (limbs
// < 2^30) for normalised a, b, via the bias 65 m whose limbs lie in [2^29, 2^30).
template <class F> HS_HD void fe_sub_lazy2(fe &r, const fe &a, const fe &b)
{
#pragma unroll
    for (int i = 0; i < NL; i++)
        r.v[i] = a.v[i] + F::bias2(i) - b.v[i];
}

template <class F> HS_HD void fe_neg_lazy2(fe &r, const fe &a)
{
#pragma unroll
    for (int i = 0; i < NL; i++)
        r.v[i] = F::bias2(i) - a.v[i];
}

// r = 257 m - b - c - d for normalised b, c, d: limbs in (2^29, 2^31 + 2^29), an addend for
// fe_mul_add / fe_sq_add
template <class F> HS_HD void fe_neg3_lazy(fe &r, const fe &b, const fe &c, const fe &d)
{
#pragma unroll
    for (int i = 0; i < NL; i++)
        r.v[i] = F::bias3(i) - b.v[i] - c.v[i] - d.v[i];
}

template <class F> HS_HD void fe_neg(fe &r, const fe &a)
{
    fe z;
    fe_zero(z);
    fe_sub<F>(r, z, a);
}

template <class F> HS_HD void fe_dbl(fe &r, const fe &a) { fe_add<F>(r, a, a); }

// 17 product columns -> normalised residue.  Columns are < 9 * 2^58 < 2^61.2 for normalised
// operands and < 9 * 2^29 * (2^29 + 2^31) < 2^63.6 when one operand comes from fe_sub_lazy; every
// intermediate below is sized for the latter.
template <class F> HS_HD void reduce_cols(fe &r, const uint64_t c[2 * NL - 1])
{
    if (F::NR == 1)
    {
        // Fold the high columns while they are still 64-bit sums: with h = lo + 2^32 hi,
        //   h 2^(29 (9 + k)) = R lo 2^(29 k) + 8 R hi 2^(29 (k + 1))   (2^261 = R),
        // adding < 2^43.5 to a column, so one carry pass over nine columns finishes the job.
        uint64_t w[NL];
#pragma unroll
        for (int k = 0; k < NL; k++)
            w[k] = c[k];
#pragma unroll
        for (int k = 0; k < NL - 1; k++)
        {
            const uint64_t h = c[NL + k];
            w[k] = mad64((uint32_t)h, F::r(0), w[k]);
            w[k + 1] = mad64((uint32_t)(h >> 32), 8u * F::r(0), w[k + 1]);
        }
        uint64_t t = 0;
#pragma unroll
        for (int k = 0; k < NL; k++)
        {
            t = w[k] + (t >> 29);
            r.v[k] = (uint32_t)t & M29;
        }
        // t >> 29 < 2^33.2 times 2^261 = R
        uint64_t u = (t >> 29) * (uint64_t)F::r(0) + r.v[0];
        r.v[0] = (uint32_t)u & M29;
        u = (u >> 29) + r.v[1];
        r.v[1] = (uint32_t)u & M29;
        r.v[2] += (uint32_t)(u >> 29); // 0 or 1: r.v[2] <= 2^29
    }
    else
    {
        // High columns 9..16 as one number H = sum c[9+k] 2^(29 k), carried into digits h[0..9];
        // the low columns stay 64-bit sums.  value = L + R H (2^261 = R, five limbs).
        uint32_t h[NL + 1];
        uint64_t t = 0;
#pragma unroll
        for (int k = 0; k < NL - 1; k++)
        {
            t = c[NL + k] + (t >> 29);
            h[k] = (uint32_t)t & M29;
        }
        t >>= 29;
        h[NL - 1] = (uint32_t)t & M29;
        h[NL] = (uint32_t)(t >> 29);
        // first fold into columns 0..13: each gains at most five products h_i R_j (< 2^58 each)
        // on top of its < 45 * 2^58
        uint64_t w[NL + F::NR];
#pragma unroll
        for (int k = 0; k < NL + F::NR; k++)
            w[k] = k < NL ? c[k] : 0;
#pragma unroll
        for (int i = 0; i <= NL; i++)
#pragma unroll
            for (int j = 0; j < F::NR; j++)
                w[i + j] = mad64(h[i], F::r(j), w[i + j]);
        // columns 9..13 again as one number M (< 2^139 since H < 2^267 and L + R H < 2^400),
        // carried into digits m[0..4]; the second fold m_i R_j (i + j <= 8) lands in columns 0..8,
        // five more products at most
        uint32_t m[F::NR];
        t = 0;
#pragma unroll
        for (int k = 0; k < F::NR; k++)
        {
            t = w[NL + k] + (t >> 29);
            m[k] = (uint32_t)t & M29;
        }
#pragma unroll
        for (int i = 0; i < F::NR; i++)
#pragma unroll
            for (int j = 0; j < F::NR; j++)
                if (i + j < NL)
                    w[i + j] = mad64(m[i], F::r(j), w[i + j]);
        // Column 8's bits from 2^261 up (top < 2^35) fold back as top * R = tlo R + thi R 2^29
        // into columns 0..5 before the carry pass (every column stays below 56 * 2^58 < 2^64).
        // The carries into limb 8 are then < 2^35, so what leaves limb 8 is < 2^7.
        const uint64_t top = w[NL - 1] >> 29;
        w[NL - 1] &= M29;
        const uint32_t tlo = (uint32_t)top & M29, thi = (uint32_t)(top >> 29);
#pragma unroll
        for (int j = 0; j < F::NR; j++)
        {
            w[j] = mad64(tlo, F::r(j), w[j]);
            w[j + 1] = mad64(thi, F::r(j), w[j + 1]);
        }
        t = 0;
#pragma unroll
        for (int k = 0; k < NL; k++)
        {
            t = w[k] + (t >> 29);
            r.v[k] = (uint32_t)t & M29;
        }
        fold_top<F>(r, (uint32_t)(t >> 29));
    }
}

template <class F> HS_HD void fe_mul(fe &r, const fe &a, const fe &b)
{
    uint64_t c[2 * NL - 1];
#pragma unroll
    for (int k = 0; k < 2 * NL - 1; k++)
        c[k] = 0;
#pragma unroll
    for (int i = 0; i < NL; i++)
#pragma unroll
        for (int j = 0; j < NL; j++)
            c[i + j] = mad64(a.v[i], b.v[j], c[i + j]);
    reduce_cols<F>(r, c);
}

// r = a b + c d with one reduction.  Operands: a and d normalised, b from fe_sub_lazy2 and c from
// fe_neg_lazy2 (or normalised): columns stay below 9 * 2^58 * (3 + 2) < 2^63.5, inside the
// bounds of reduce_cols.
template <class F> HS_HD void fe_mul_sum2(fe &r, const fe &a, const fe &b, const fe &c, const fe &d)
{
    uint64_t w[2 * NL - 1];
#pragma unroll
    for (int k = 0; k < 2 * NL - 1; k++)
        w[k] = 0;
#pragma unroll
    for (int i = 0; i < NL; i++)
#pragma unroll
        for (int j = 0; j < NL; j++)
        {
            w[i + j] = mad64(a.v[i], b.v[j], w[i + j]);
            w[i + j] = mad64(c.v[i], d.v[j], w[i + j]);
        }
    reduce_cols<F>(r, w);
}

// r = a b + c and r = a^2 + c with one reduction: c (limbs < 2^32, such as a lazy negation)
// starts the nine low columns, so a product followed by a subtraction needs no second
// normalisation.  Operands of the product as for fe_mul.
template <class F> HS_HD void fe_mul_add(fe &r, const fe &a, const fe &b, const fe &c)
{
    uint64_t w[2 * NL - 1];
#pragma unroll
    for (int k = 0; k < 2 * NL - 1; k++)
        w[k] = k < NL ? c.v[k] : 0;
#pragma unroll
    for (int i = 0; i < NL; i++)
#pragma unroll
        for (int j = 0; j < NL; j++)
            w[i + j] = mad64(a.v[i], b.v[j], w[i + j]);
    reduce_cols<F>(r, w);
}

template <class F> HS_HD void fe_sq_add(fe &r, const fe &a, const fe &c)
{
    uint64_t w[2 * NL - 1];
    uint32_t d[NL];
#pragma unroll
    for (int k = 0; k < 2 * NL - 1; k++)
        w[k] = k < NL ? c.v[k] : 0;
#pragma unroll
    for (int i = 0; i < NL; i++)
        d[i] = a.v[i] << 1;
#pragma unroll
    for (int i = 0; i < NL; i++)
    {
        w[2 * i] = mad64(a.v[i], a.v[i], w[2 * i]);
#pragma unroll
        for (int j = i + 1; j < NL; j++)
            w[i + j] = mad64(a.v[i], d[j], w[i + j]);
    }
    reduce_cols<F>(r, w);
}

template <class F> HS_HD void fe_sq(fe &r, const fe &a)
{
    uint64_t c[2 * NL - 1];
    uint32_t d[NL];
#pragma unroll
    for (int k = 0; k < 2 * NL - 1; k++)
        c[k] = 0;
#pragma unroll
    for (int i = 0; i < NL; i++)
        d[i] = a.v[i] << 1;
#pragma unroll
    for (int i = 0; i < NL; i++)
    {
        c[2 * i] = mad64(a.v[i], a.v[i], c[2 * i]);
#pragma unroll
        for (int j = i + 1; j < NL; j++)
            c[i + j] = mad64(a.v[i], d[j], c[i + j]);
    }
    reduce_cols<F>(r, c);
}

// r = a * k for a small constant k (< 2^29)
template <class F> HS_HD void fe_mul_small(fe &r, const fe &a, uint32_t k)
{
    uint64_t t = 0;
#pragma unroll
    for (int i = 0; i < NL; i++)
    {
        t = mad64(a.v[i], k, t >> 29);
        r.v[i] = (uint32_t)t & M29;
    }
    fold_top<F>(r, (uint32_t)(t >> 29));
}

// r = a - k b for normalised a, b and small k (<= 16) with one carry pass: limb i is
// a_i + k (bias2_i - b_i) < 2^29 + k 2^30, so the carries stay below 2^6
template <class F> HS_HD void fe_sub_small_mul(fe &r, const fe &a, const fe &b, uint32_t k)
{
    uint64_t t = 0;
#pragma unroll
    for (int i = 0; i < NL; i++)
    {
        t = mad64(F::bias2(i) - b.v[i], k, (uint64_t)a.v[i] + (t >> 29));
        r.v[i] = (uint32_t)t & M29;
    }
    fold_top<F>(r, (uint32_t)(t >> 29));
}

// ---------------------------------------------------------------- conversions
// 8 little-endian words (any value < 2^256) -> limbs (normalised, not reduced)
HS_HD void fe_from_words(fe &r, const uint32_t w[8])
{
#pragma unroll
    for (int i = 0; i < NL; i++)
    {
        const int bit = 29 * i, wi = bit >> 5, sh = bit & 31;
        uint64_t lo = wi < 8 ? w[wi] : 0u;
        uint64_t hi = wi + 1 < 8 ? w[wi + 1] : 0u;
        r.v[i] = (uint32_t)(((lo | (hi << 32)) >> sh) & M29);
    }
}

// canonical residue in [0, m) as 8 little-endian words
template <class F> HS_HD void fe_to_words_canon(uint32_t out[8], const fe &a)
{
    // value = sum v[i] 2^(29 i) with v[i] <= 2^29 + 1: rebuild 9 words exactly
    uint32_t w[9];
#pragma unroll
    for (int i = 0; i < 9; i++)
        w[i] = 0;
#pragma unroll
    for (int i = 0; i < NL; i++)
    {
        const int bit = 29 * i, wi = bit >> 5, sh = bit & 31;
        uint64_t x = (uint64_t)a.v[i] << sh; // < 2^62
        uint64_t t = (uint64_t)w[wi] + (uint32_t)x;
        w[wi] = (uint32_t)t;
        t = (t >> 32) + (x >> 32);
#pragma unroll
        for (int k = wi + 1; k < 9; k++)
        {
            t += w[k];
            w[k] = (uint32_t)t;
            t >>= 32;
        }
    }
    // fold bits >= 255: value = L + H * 2^255 = L + H * (2^255 mod m), H < 2^8
    uint32_t H = (w[7] >> 31) | (w[8] << 1);
    w[7] &= 0x7fffffffu;
    uint64_t t = 0;
#pragma unroll
    for (int i = 0; i < 8; i++)
    {
        t = (i < F::NCW ? mad64(H, F::cw(i), w[i]) : (uint64_t)w[i]) + (t >> 32);
        w[i] = (uint32_t)t;
    }
    // now value < 2^255 + 2^8 * 2^127 < 2m: subtract m once if needed (twice for safety)
#pragma unroll
    for (int pass = 0; pass < 2; pass++)
    {
        uint32_t d[8];
        int64_t br = 0;
#pragma unroll
        for (int i = 0; i < 8; i++)
        {
            int64_t x = (int64_t)w[i] - F::mw(i) + br;
            d[i] = (uint32_t)x;
            br = x >> 32;
        }
        const uint32_t keep = br == 0 ? 0xffffffffu : 0u;
#pragma unroll
        for (int i = 0; i < 8; i++)
            w[i] = (d[i] & keep) | (w[i] & ~keep);
    }
#pragma unroll
    for (int i = 0; i < 8; i++)
        out[i] = w[i];
}

// canonical (fully reduced) limbs
template <class F> HS_HD void fe_canon(fe &r, const fe &a)
{
    uint32_t w[8];
    fe_to_words_canon<F>(w, a);
    fe_from_words(r, w);
}

template <class F> HS_HD bool fe_is_zero_canon(const fe &a)
{
    uint32_t w[8];
    fe_to_words_canon<F>(w, a);
    uint32_t o = 0;
#pragma unroll
    for (int i = 0; i < 8; i++)
        o |= w[i];
    return o == 0;
}

// Zero test with a cheap exclusion first.  After one carry pass the value is D (strict 29-bit
// digits d[0..8]) plus c * 2^261.  With c = 0, D < 2^261 and D = lo + 2^255 hi (hi < 2^6), so
// D = 0 (mod m) needs lo + hi * (2^255 mod m) to be 0 or m (it is < 2m).  Both m = 2^255 - 19
// and m = 2^255 - gamma (gamma < 2^127) have all bits 145..254 set, as does m - k (2^255 mod m)
// for k <= 64, so digits 5..7 (bits 145..231) are then all zero or all ones.  Otherwise the
// value is certainly nonzero; the full canonical test runs only in the rare remaining cases.
template <class F> HS_HD bool fe_is_zero(const fe &a)
{
    uint32_t d[NL];
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < NL; i++)
    {
        const uint32_t t = a.v[i] + c;
        d[i] = t & M29;
        c = t >> 29;
    }
    const uint32_t o = d[5] | d[6] | d[7], n = d[5] & d[6] & d[7];
    if (c == 0 && o != 0 && n != M29)
        return false;
    return fe_is_zero_canon<F>(a);
}

// The same test for the output of fe_norm / fe_add / fe_sub / reduce_cols, without the carry
// pass: there only limbs 0..2 can equal 2^29, limbs 3..8 are strict digits, and the value is
// below 2^261 + 2^88, so a multiple of m is k m with k <= 64.  A carry from limbs 0..2 changes
// digits 5..7 only by running through digit 4 as 2^29 - 1 and leaving it 0, but every such k m
// (k >= 1) has digit 4 nonzero (bits 133..144 of k m are all ones for both moduli); so for a
// multiple of m the limbs 5..7 are its strict digits, and mixed limbs 5..7 prove the value
// nonzero.
template <class F> HS_HD bool fe_is_zero_norm(const fe &a)
{
    const uint32_t o = a.v[5] | a.v[6] | a.v[7], n = a.v[5] & a.v[6] & a.v[7];
    if (o != 0 && n != M29)
        return false;
    return fe_is_zero_canon<F>(a);
}

// The cheap half of fe_is_zero_norm: false proves the value nonzero; true leaves the canonical test
// (fe_is_zero_canon) to decide.
HS_HD bool fe_maybe_zero_norm(const fe &a)
{
    const uint32_t o = a.v[5] | a.v[6] | a.v[7], n = a.v[5] & a.v[6] & a.v[7];
    return o == 0 || n == M29;
}

template <class F> HS_HD bool fe_eq(const fe &a, const fe &b)
{
    fe d;
    fe_sub<F>(d, a, b);
    return fe_is_zero<F>(d);
}

template <class F> HS_HD bool fe_is_odd(const fe &a)
{
    uint32_t w[8];
    fe_to_words_canon<F>(w, a);
    return w[0] & 1u;
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
        fe t = tab[0];
#pragma unroll
        for (int k = 1; k < 16; k++)
            if (nib == (uint32_t)k)
                t = tab[k];
        fe_mul<F>(acc, acc, t);
    }
    r = acc;
}

// a^-1 = a^(m-2) (Fermat); maps 0 to 0 like the reference's fp_invert / fq_invert.  The kernels use
// fe_inv (divsteps, inverse.h); this form checks it in the host tests.
template <class F> HS_HD void fe_inv_fermat(fe &r, const fe &a)
{
    uint32_t e[8];
#pragma unroll
    for (int i = 0; i < 8; i++)
        e[i] = F::mw(i);
    e[0] -= 2; // both moduli have a low word >= 2
    fe_pow<F>(r, a, e);
}

// radix-2^51 limbs (the reference's fp_fe / fq_fe, uint64_t[5], limbs may exceed 51 bits) ->
// normalised fe.  value = sum l_i 2^(51 i) < 2^(204 + 64).
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
        uint64_t lo = l[i] << sh;
        uint64_t hi = sh ? (l[i] >> (64 - sh)) : 0;
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
    // value = w[0..7] + 2^256 * w[8] (w[8] < 2^12, w[9] = 0); 2^256 = 2^-5 * 2^261 is not a limb
    // boundary, so fold through 2^261: 2^256 * w8 = 2^261 * (w8 >> 5) + 2^256 * (w8 & 31)
    fe_from_words(r, w);
    // w8 & 31 goes to bits 256..260 of limb 8 (bit 24..28 of limb 8)
    r.v[8] |= (w[8] & 31u) << 24;
    fold_top<F>(r, w[8] >> 5);
}

// fe -> canonical radix-2^51 limbs
template <class F> HS_HD void fe_to_r51(uint64_t l[5], const fe &a)
{
    uint32_t c[8];
    fe_to_words_canon<F>(c, a);
    uint64_t w[4];
#pragma unroll
    for (int i = 0; i < 4; i++)
        w[i] = (uint64_t)c[2 * i] | ((uint64_t)c[2 * i + 1] << 32);
    const uint64_t M = (1ull << 51) - 1;
    l[0] = w[0] & M;
    l[1] = ((w[0] >> 51) | (w[1] << 13)) & M;
    l[2] = ((w[1] >> 38) | (w[2] << 26)) & M;
    l[3] = ((w[2] >> 25) | (w[3] << 39)) & M;
    l[4] = (w[3] >> 12) & M;
}

} // namespace hs

#include "inverse.h"
