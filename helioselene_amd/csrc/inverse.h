// Modular inversion in F_p and F_q by divsteps (Bernstein and Yang, "Fast constant-time gcd
// computation and modular inversion", TCHES 2019, with the batched transition matrices and the
// (d, e) bookkeeping of its section 11), written for one GPU lane: signed 30-bit limbs, 30 divsteps
// per transition matrix (every entry fits an int32), v_mad_i64_i32 column sums for the
// matrix-vector products.
//
// It replaces Fermat's a^(m-2), about 320 field products on one lane, by about 20 batches of 30
// divsteps, each a few hundred 32-bit instructions: the inversions of a plain MSM's final encoding
// and of the ingest of points with Z != 1 sit on latency-bound chains.  0 maps to 0, as the
// reference's fp_invert / fq_invert do (fp/src/x64/fp_invert.cpp:32-67, fq/src/x64/fq_invert.cpp:41-72).
//
// State (all in signed radix 2^30, limbs 0..7 in [0, 2^30), limb 8 signed):
//   f, g with f = m, g = x initially; each batch maps (f, g) to T (f, g) / 2^30 exactly;
//   d, e with f = d x and g = e x (mod m); T applies to them too, with a multiple of m added so the
//   division by 2^30 is exact; they stay in (-2m, m).
// When g reaches 0, f = +-1 and x^-1 = +-d (mod m).
#pragma once
#include "field.h"

namespace hs
{

struct s30
{
    int32_t v[9];
};

constexpr int32_t M30 = 0x3fffffff;

// the modulus in radix 2^30 and its inverse modulo 2^30
template <class F> struct Inv30;
template <> struct Inv30<Fp>
{
    static HS_HD int32_t m(int i) { return i == 0 ? 0x3fffffed : i == 8 ? 0x7fff : 0x3fffffff; }
    static constexpr uint32_t minv = 0x179435e5u;
};
template <> struct Inv30<Fq>
{
    static HS_HD int32_t m(int i)
    {
        switch (i)
        {
        case 0: return 0x3927c79f;
        case 1: return 0x3adb49c9;
        case 2: return 0x3656b586;
        case 3: return 0x1fde0b2d;
        case 4: return 0x3fffffbf;
        case 8: return 0x7fff;
        default: return 0x3fffffff;
        }
    }
    static constexpr uint32_t minv = 0x290b945fu;
};

// 30 divsteps on the low bits of f and g (f odd): returns the transition matrix scaled by 2^30,
//   [f', g'] = [[u, v], [q, r]] [f, g] / 2^30,
// and advances delta.  One step, with delta > 0 and g odd: (f, g) <- (g, (g - f) / 2), delta <- 1 - delta;
// with g odd otherwise: g <- (g + f) / 2, delta <- 1 + delta; with g even: g <- g / 2, delta <- 1 + delta.
// Branch-free: the first case is a swap with negation followed by the second case.  After i steps
// the low 32 - i bits of fl and gl are still exact, enough for the parity each step reads.
HS_HD void divsteps30(int32_t &delta, uint32_t fl, uint32_t gl, int32_t &u_, int32_t &v_, int32_t &q_, int32_t &r_)
{
    uint32_t u = 1, v = 0, q = 0, r = 1; // two's-complement arithmetic on uint32_t
    int32_t dl = delta;
#pragma unroll
    for (int i = 0; i < 30; i++)
    {
        const uint32_t odd = gl & 1u;
        const uint32_t ms = 0u - (odd & (uint32_t)(dl > 0)); // swap-and-negate case
        uint32_t t = (fl ^ gl) & ms;
        fl ^= t;
        gl ^= t;
        gl = (gl ^ ms) - ms;
        t = (u ^ q) & ms;
        u ^= t;
        q ^= t;
        q = (q ^ ms) - ms;
        t = (v ^ r) & ms;
        v ^= t;
        r ^= t;
        r = (r ^ ms) - ms;
        dl = (int32_t)(((uint32_t)dl ^ ms) - ms);
        const uint32_t mo = 0u - odd;
        gl += fl & mo;
        q += u & mo;
        r += v & mo;
        dl += 1;
        gl >>= 1;
        u <<= 1;
        v <<= 1;
    }
    delta = dl;
    u_ = (int32_t)u;
    v_ = (int32_t)v;
    q_ = (int32_t)q;
    r_ = (int32_t)r;
}

HS_HD int64_t smad(int32_t a, int32_t b, int64_t c) { return (int64_t)a * b + c; }

// (f, g) <- T (f, g) / 2^30 (exact)
HS_HD void s30_update_fg(s30 &f, s30 &g, int32_t u, int32_t v, int32_t q, int32_t r)
{
    int64_t cf = smad(v, g.v[0], smad(u, f.v[0], 0));
    int64_t cg = smad(r, g.v[0], smad(q, f.v[0], 0));
    cf >>= 30;
    cg >>= 30;
#pragma unroll
    for (int i = 1; i < 9; i++)
    {
        const int32_t fi = f.v[i], gi = g.v[i];
        cf = smad(v, gi, smad(u, fi, cf));
        cg = smad(r, gi, smad(q, fi, cg));
        f.v[i - 1] = (int32_t)cf & M30;
        g.v[i - 1] = (int32_t)cg & M30;
        cf >>= 30;
        cg >>= 30;
    }
    f.v[8] = (int32_t)cf;
    g.v[8] = (int32_t)cg;
}

// (d, e) <- (T (d, e) + m (md, me)) / 2^30 with md, me chosen so the low 30 bits vanish; d and e
// stay in (-2m, m) when they start there (a negative input first adds its matrix column's
// entries to md / me, which keeps the result above -2m).
template <class F> HS_HD void s30_update_de(s30 &d, s30 &e, int32_t u, int32_t v, int32_t q, int32_t r)
{
    using I = Inv30<F>;
    const int32_t sd = d.v[8] >> 31, se = e.v[8] >> 31; // all ones when negative
    int32_t md = (u & sd) + (v & se);
    int32_t me = (q & sd) + (r & se);
    int64_t cd = smad(v, e.v[0], smad(u, d.v[0], 0));
    int64_t ce = smad(r, e.v[0], smad(q, d.v[0], 0));
    md -= (int32_t)((I::minv * (uint32_t)cd + (uint32_t)md) & (uint32_t)M30);
    me -= (int32_t)((I::minv * (uint32_t)ce + (uint32_t)me) & (uint32_t)M30);
    cd = smad(I::m(0), md, cd);
    ce = smad(I::m(0), me, ce);
    cd >>= 30;
    ce >>= 30;
#pragma unroll
    for (int i = 1; i < 9; i++)
    {
        const int32_t di = d.v[i], ei = e.v[i], mi = I::m(i);
        cd = smad(mi, md, smad(v, ei, smad(u, di, cd)));
        ce = smad(mi, me, smad(r, ei, smad(q, di, ce)));
        d.v[i - 1] = (int32_t)cd & M30;
        e.v[i - 1] = (int32_t)ce & M30;
        cd >>= 30;
        ce >>= 30;
    }
    d.v[8] = (int32_t)cd;
    e.v[8] = (int32_t)ce;
}

// x + k m for k = +-1 (limb-wise, then one carry pass back to the radix-2^30 format)
template <class F> HS_HD void s30_add_m(s30 &x, int32_t k)
{
    int32_t c = 0;
#pragma unroll
    for (int i = 0; i < 8; i++)
    {
        const int32_t t = x.v[i] + k * Inv30<F>::m(i) + c;
        x.v[i] = t & M30;
        c = t >> 30;
    }
    x.v[8] += k * Inv30<F>::m(8) + c;
}

HS_HD void s30_neg(s30 &x)
{
    int32_t c = 0;
#pragma unroll
    for (int i = 0; i < 8; i++)
    {
        const int32_t t = c - x.v[i];
        x.v[i] = t & M30;
        c = t >> 30;
    }
    x.v[8] = c - x.v[8];
}

HS_HD bool s30_is_zero(const s30 &x)
{
    int32_t o = 0;
#pragma unroll
    for (int i = 0; i < 9; i++)
        o |= x.v[i];
    return o == 0;
}

// a^-1 (mod m), 0 -> 0.  Batches of 30 divsteps until g = 0; 25 batches (750 divsteps) bound the
// 741 that 256-bit inputs can need (the paper's theorem 11.2), so the loop always ends.
template <class F> HS_HD void fe_inv(fe &r, const fe &a)
{
    // the input is moved into vector registers first: a wave-uniform input (one lane finishing an
    // MSM) would otherwise put the whole chain on the scalar unit, whose multiplies are far slower
    fe x = a;
    fe_opaque(x);
    uint32_t w[8];
    fe_to_words_canon<F>(w, x);
    s30 f, g, d, e;
#pragma unroll
    for (int i = 0; i < 9; i++)
    {
        // bits 30 i .. 30 i + 29 of the canonical value
        const int bit = 30 * i, wi = bit >> 5, sh = bit & 31;
        const uint64_t lo = wi < 8 ? w[wi] : 0u, hi = wi + 1 < 8 ? w[wi + 1] : 0u;
        g.v[i] = (int32_t)(((lo | (hi << 32)) >> sh) & (uint64_t)M30);
        f.v[i] = Inv30<F>::m(i);
        d.v[i] = 0;
        e.v[i] = i == 0 ? 1 : 0;
    }
    int32_t delta = 1;
    for (int b = 0; b < 25 && !s30_is_zero(g); b++)
    {
        int32_t u, v, q, rr;
        divsteps30(delta, (uint32_t)f.v[0] | ((uint32_t)f.v[1] << 30), (uint32_t)g.v[0] | ((uint32_t)g.v[1] << 30), u, v,
                   q, rr);
        s30_update_fg(f, g, u, v, q, rr);
        s30_update_de<F>(d, e, u, v, q, rr);
    }
    // f = +-1 (or f = m, d = 0 for a = 0): x^-1 = sign(f) d, brought from (-2m, 2m) into [0, m)
    if (f.v[8] < 0)
        s30_neg(d);
    if (d.v[8] < 0)
        s30_add_m<F>(d, 1);
    if (d.v[8] < 0)
        s30_add_m<F>(d, 1);
    s30 t = d;
    s30_add_m<F>(t, -1);
    if (t.v[8] >= 0)
        d = t;
    // radix 2^30 (canonical, < m) -> 8 words -> limbs
    uint32_t o[8];
#pragma unroll
    for (int k = 0; k < 8; k++)
        o[k] = 0;
#pragma unroll
    for (int i = 0; i < 9; i++)
    {
        const int bit = 30 * i, wi = bit >> 5, sh = bit & 31;
        const uint64_t x = (uint64_t)(uint32_t)d.v[i] << sh;
        if (wi < 8)
            o[wi] |= (uint32_t)x;
        if (wi + 1 < 8)
            o[wi + 1] |= (uint32_t)(x >> 32);
    }
    fe_from_words(r, o);
}

} // namespace hs
