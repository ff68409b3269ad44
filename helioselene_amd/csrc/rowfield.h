// Field arithmetic with one element spread over a DPP row (16 lanes), for the latency-bound serial
// chains (Horner combines): lane l of a row holds limb l of the nine 29-bit limbs of field.h's
// format, lanes 9..15 hold 0.  A wavefront has four rows, so one instruction stream computes four
// independent products, one per row, and a product takes about 50 (F_p) / 105 (F_q) wave
// instructions instead of field.h's 136 / 248 on a single lane.
//
// Product a b: lane k accumulates column k = sum_i a_i b_(k-i) as nine v_mad_u64_u32, with a_i
// broadcast across the row (row_newbcast:i) and b shifted up by i lanes (row_shr:i, zero fill);
// column 16 = a_8 b_8 is formed by every lane.  The reduction then works on all lanes at once:
// carry rounds (every lane keeps its low 29 bits and passes the rest one lane up, row_shr:1) and
// folds of the positions from 9 up by 2^261 = R (mod m), which bring each high position down by
// nine lanes (row_shl) and multiply it by R's limbs.
//
// "Row-normalised" (RN) below: limbs < 2^29 + 2^18 in lanes 0..8, zeros in lanes 9..15.  Every
// result is RN unless its comment says otherwise.  Product operands: the first is read only in
// lanes 0..8 (broadcast), the second must have zeros in lanes 9..15 (it is shifted); their limbs
// must keep 9 A B < 2^63.8 (A, B the limb bounds): RN x RN, RN x a lazy sum of two RN values,
// RN x a lazy difference (limbs < 2^31.5) all qualify.
//
// The lane primitives below (DPP row moves, 64-bit pieces, selects) are the only cross-lane
// operations; a host build can supply a lane-by-lane model of them (HS_ROW_EMU) so the same
// arithmetic runs on the CPU with every intermediate checked for wrap-around.
#pragma once
#include "field.h"

namespace hs
{
namespace row
{

#ifndef HS_ROW_EMU
using u32v = uint32_t;
using u64v = uint64_t;
using bv = bool;
#define HS_RW __device__ __forceinline__

HS_RW u32v lane_id() { return __lane_id(); }
// row_shr:K / row_shl:K with zero fill at the row edge, row_newbcast:K (lane K of each row to all
// sixteen lanes of the row)
template <int K> HS_RW u32v shr(u32v x) { return (u32v)__builtin_amdgcn_update_dpp(0, (int)x, 0x110 + K, 0xf, 0xf, true); }
template <int K> HS_RW u32v shl(u32v x) { return (u32v)__builtin_amdgcn_update_dpp(0, (int)x, 0x100 + K, 0xf, 0xf, true); }
template <int K> HS_RW u32v bcast(u32v x) { return (u32v)__builtin_amdgcn_update_dpp(0, (int)x, 0x150 + K, 0xf, 0xf, true); }
HS_RW u32v lo(u64v x) { return (u32v)x; }
HS_RW u32v hi(u64v x) { return (u32v)(x >> 32); }
HS_RW u64v wide(u32v x) { return (u64v)x; }
HS_RW u64v mad64(u32v a, u32v b, u64v c) { return (u64v)a * b + c; }
// a b + c for a, b < 2^24 (v_mad_u32_u24)
HS_RW u32v mad24(u32v a, u32v b, u32v c) { return (a & 0xffffffu) * (b & 0xffffffu) + c; }
HS_RW u64v add64(u64v a, u64v b) { return a + b; }
HS_RW u64v sh29w(u64v x) { return x >> 29; }
// (x >> 29) for x < 2^61
HS_RW u32v sh29(u64v x) { return (u32v)(x >> 29); }
HS_RW bv eq(u32v a, uint32_t b) { return a == b; }
HS_RW bv lt(u32v a, uint32_t b) { return a < b; }
HS_RW bv nz(u32v a) { return a != 0; }
template <class T> HS_RW T sel(bv c, T a, T b) { return c ? a : b; }
// (a, b) <- rows (a0, b0, a2, b2), (a1, b1, a3, b3)
HS_RW void swap16(u32v &a, u32v &b)
{
    const auto r = __builtin_amdgcn_permlane16_swap(a, b, false, false);
    a = r[0];
    b = r[1];
}
// (a, b) <- rows (a0, a1, b0, b1), (a2, a3, b2, b3)
HS_RW void swap32(u32v &a, u32v &b)
{
    const auto r = __builtin_amdgcn_permlane32_swap(a, b, false, false);
    a = r[0];
    b = r[1];
}
HS_RW uint32_t readlane(u32v x, int l) { return (uint32_t)__builtin_amdgcn_readlane((int)x, l); }
// lanes with c load p[idx], the others get 0
HS_RW u32v load_if(bv c, const uint32_t *p, u32v idx)
{
    u32v x = 0;
    if (c)
        x = p[idx];
    return x;
}
#endif

template <int K> HS_RW u64v shr64(u64v x) { return wide(shr<K>(lo(x))) | (wide(shr<K>(hi(x))) << 32); }
template <int K> HS_RW u64v bcast64(u64v x) { return wide(bcast<K>(lo(x))) | (wide(bcast<K>(hi(x))) << 32); }

struct rfe
{
    u32v v;
};

// Per-lane constants and lane predicates of one field, set up once per kernel.
template <class F> struct RowK
{
    u32v l;                       // lane within the row
    u32v bias, bias2, bias3;      // 129 m, 65 m, 257 m limbs (field.h), zero in lanes 9..15
    u32v gam;                     // gamma = 2^255 mod m, limbs in lanes 0..4 (F_q)
    u32v r0, r1;                  // F_p: R = 1216 in lane 0, 8 R in lane 1
    bv lt9, is0, is7, is8, is9;   // lane predicates
    bv rodd, rhi, row0;           // row 1 or 3, row 2 or 3, row 0
};

template <class G> HS_RW u32v per_limb(u32v l, G g)
{
    u32v r = 0;
#pragma unroll
    for (int i = 0; i < NL; i++)
        r = sel(eq(l, (uint32_t)i), (u32v)g(i), r);
    return r;
}

template <class F> HS_RW void rowk_init(RowK<F> &k)
{
    const u32v lane = lane_id();
    k.l = lane & 15u;
    k.lt9 = lt(k.l, 9);
    k.is0 = eq(k.l, 0);
    k.is7 = eq(k.l, 7);
    k.is8 = eq(k.l, 8);
    k.is9 = eq(k.l, 9);
    k.rodd = nz(lane & 16u);
    k.rhi = nz(lane & 32u);
    k.row0 = lt(lane, 16);
    k.bias = per_limb(k.l, [](int i) { return F::bias(i); });
    k.bias2 = per_limb(k.l, [](int i) { return F::bias2(i); });
    k.bias3 = per_limb(k.l, [](int i) { return F::bias3(i); });
    k.gam = per_limb(k.l, [](int i) { return F::NR == 1 ? 0u : F::g(i); });
    k.r0 = sel(k.is0, (u32v)(F::NR == 1 ? F::r(0) : 0u), (u32v)0u);
    k.r1 = sel(eq(k.l, 1), (u32v)(F::NR == 1 ? 8u * F::r(0) : 0u), (u32v)0u);
}

// ---------------------------------------------------------------- reductions
// F_q: fold the bits from 2^255 up with gamma.  d: lanes 0..8 < 2^32, lane 9 < 2^11 (weight
// 2^261), lanes 10..15 ignored.  The result is RN with limb 8 < 2^23 + 2^17.
template <class F> HS_RW rfe fold255(const RowK<F> &k, u32v d)
{
    const u32v top = (bcast<8>(d) >> 23) + (bcast<9>(d) << 6); // < 2^17.1
    const u32v dm = sel(k.is8, d & 0x7fffffu, sel(k.lt9, d, (u32v)0u));
    const u64v w = mad64(top, k.gam, wide(dm)); // lanes 0..4 < 2^45
    return {(lo(w) & M29) + shr<1>(sh29(w))};
}

// limbs < 2^32 (zeros in lanes 9..15) -> RN
template <class F> HS_RW rfe rnorm(const RowK<F> &k, u32v x)
{
    const u32v h = x >> 29;
    const u32v d = (x & M29) + shr<1>(h); // lane 9 takes the carry out of limb 8 (< 2^3)
    if (F::NR == 1)
    {
        // that carry has weight 2^261 = R: back into lane 0
        return {sel(k.lt9, mad24(bcast<8>(h), k.r0, d), (u32v)0u)};
    }
    return fold255<F>(k, d);
}

// limbs < 2^34 (zeros in lanes 9..15) -> RN
template <class F> HS_RW rfe rnorm64(const RowK<F> &k, u64v x)
{
    const u32v h = sh29(x);
    const u32v d = (lo(x) & M29) + shr<1>(h);
    if (F::NR == 1)
        return {sel(k.lt9, mad24(bcast<8>(h), k.r0, d), (u32v)0u)};
    return fold255<F>(k, d);
}

// F_q: positions 9.. of digit row d (lanes 0..15, limbs < 2^29 + 2^7) plus the digits e16, e17, e18
// of positions 16..18 (zero when absent), folded by R = 2^261 mod q (five limbs) into lanes
// 0..13; lanes 0..8 keep their own digits.  Every lane gains at most five products below 2^58.1.
template <class F> HS_RW u64v fold261_q(const RowK<F> &k, u32v d, u32v e16, u32v e17, u32v e18)
{
    u32v E = shl<9>(d); // E_j = digit at position 9 + j
    E = sel(k.is7, e16, E);
    E = sel(k.is8, e17, E);
    E = sel(k.is9, e18, E);
    u64v w = wide(sel(k.lt9, d, (u32v)0u));
    w = mad64(E, F::r(0), w);
    w = mad64(shr<1>(E), F::r(1), w);
    w = mad64(shr<2>(E), F::r(2), w);
    w = mad64(shr<3>(E), F::r(3), w);
    w = mad64(shr<4>(E), F::r(4), w);
    return w;
}

// 64-bit lanes (< 2^62) -> digits: two carry rounds, every lane < 2^29 + 2^4 afterwards (the
// carries out of lane 15 are dropped: callers know those positions are empty)
HS_RW u32v carry2(u64v w)
{
    const u64v v = add64(shr64<1>(sh29w(w)), wide(lo(w) & M29)); // < 2^33.1
    return (lo(v) & M29) + shr<1>(sh29(v));
}

// columns (lanes 0..15: columns 0..15, c16 = column 16 in every lane) -> RN
template <class F> HS_RW rfe rreduce(const RowK<F> &k, u64v acc, u64v c16)
{
    if (F::NR == 1)
    {
        // Column c = lo + 2^32 hi at position 9 + j adds R lo at position j and 8 R hi at j + 1
        // (2^32 = 8 * 2^29); lanes 9..15 keep stale columns, cleared at the end (only ever moved up).
        const u32v L = sel(k.is7, lo(c16), shl<9>(lo(acc)));
        const u32v H = sel(k.is8, hi(c16), sel(k.is0, (u32v)0u, shl<8>(hi(acc))));
        u64v w = mad64(L, F::r(0), acc);
        w = mad64(H, 8u * F::r(0), w); // < 2^64 for columns < 2^63.9
        // carry round on 64-bit lanes; the carry out of limb 8 (weight 2^261 = R) goes back to
        // lanes 0 (its low word times R) and 1 (its high word times 8 R)
        const u64v h = sh29w(w);      // < 2^35
        const u64v t8 = bcast64<8>(h);
        u64v v = add64(shr64<1>(h), wide(lo(w) & M29));
        v = mad64(lo(t8), k.r0, v);
        v = mad64(hi(t8), k.r1, v);   // lane 0 < 2^42.3, the others < 2^35.1
        // second round: 32-bit digits
        const u32v h1 = sh29(v);      // lane 0 < 2^13.3, the others < 2^6.1
        const u32v d = (lo(v) & M29) + shr<1>(h1);
        return {sel(k.lt9, mad24(bcast<8>(h1), k.r0, d), (u32v)0u)};
    }
    else
    {
        // all 17 positions to digits; the carries out of position 15 join column 16
        const u64v h = sh29w(acc); // < 2^35
        const u64v v = add64(shr64<1>(h), wide(lo(acc) & M29));
        u64v p = add64(c16, bcast64<15>(h));
        const u32v h1 = sh29(v); // < 2^6.1
        const u32v d = (lo(v) & M29) + shr<1>(h1);
        p = add64(p, wide(bcast<15>(h1))); // positions 16.., < 2^61
        // first fold (positions 9..18 -> 0..13; the value is then < 2^395), two carry rounds,
        // second fold (positions 9..13 -> 0..8; value < 2^266.3), two carry rounds (position 9 takes
        // what is left above 2^261, < 2^6), then the top fold at 2^255
        const u64v w1 = fold261_q<F>(k, d, lo(p) & M29, sh29(p) & M29, hi(p) >> 26);
        const u32v d1 = carry2(w1);
        const u64v w2 = fold261_q<F>(k, d1, 0u, 0u, 0u);
        const u32v d2 = carry2(w2);
        return fold255<F>(k, d2);
    }
}

template <class F> HS_RW rfe rmul(const RowK<F> &k, const rfe &a, const rfe &b)
{
    const u32v x = a.v, y = b.v;
    const u32v a8 = bcast<8>(x);
    u64v acc = mad64(bcast<0>(x), y, 0);
    acc = mad64(bcast<1>(x), shr<1>(y), acc);
    acc = mad64(bcast<2>(x), shr<2>(y), acc);
    acc = mad64(bcast<3>(x), shr<3>(y), acc);
    acc = mad64(bcast<4>(x), shr<4>(y), acc);
    acc = mad64(bcast<5>(x), shr<5>(y), acc);
    acc = mad64(bcast<6>(x), shr<6>(y), acc);
    acc = mad64(bcast<7>(x), shr<7>(y), acc);
    acc = mad64(a8, shr<8>(y), acc);
    const u64v c16 = mad64(a8, bcast<8>(y), 0);
    return rreduce<F>(k, acc, c16);
}

template <class F> HS_RW rfe rsq(const RowK<F> &k, const rfe &a) { return rmul<F>(k, a, a); }

// ---------------------------------------------------------------- linear operations
// a + b without normalisation (limbs < 2^30 + 2^18): a product operand
HS_RW rfe radd_lazy(const rfe &a, const rfe &b) { return {a.v + b.v}; }
template <class F> HS_RW rfe radd(const RowK<F> &k, const rfe &a, const rfe &b) { return rnorm<F>(k, a.v + b.v); }
// a - b + 129 m without normalisation (limbs < 2^31.5): a product operand beside an RN one
template <class F> HS_RW rfe rsub_lazy(const RowK<F> &k, const rfe &a, const rfe &b) { return {a.v + k.bias - b.v}; }
template <class F> HS_RW rfe rsub(const RowK<F> &k, const rfe &a, const rfe &b) { return rnorm<F>(k, a.v + k.bias - b.v); }
// a - b - c and a - b - c - d (bias 257 m, limbs in [2^31, 2^31 + 2^29))
template <class F> HS_RW rfe rsub2(const RowK<F> &k, const rfe &a, const rfe &b, const rfe &c)
{
    return rnorm<F>(k, a.v + k.bias3 - b.v - c.v);
}
template <class F> HS_RW rfe rsub3(const RowK<F> &k, const rfe &a, const rfe &b, const rfe &c, const rfe &d)
{
    return rnorm<F>(k, a.v + k.bias3 - b.v - c.v - d.v);
}
// a s for s <= 7
template <class F> HS_RW rfe rmul_small(const RowK<F> &k, const rfe &a, uint32_t s) { return rnorm<F>(k, a.v * s); }
// a - s b for s <= 16 (bias 65 m, limbs in [2^29 + 2^25, 2^30) above every RN limb)
template <class F> HS_RW rfe rsub_small_mul(const RowK<F> &k, const rfe &a, const rfe &b, uint32_t s)
{
    return rnorm64<F>(k, mad64(k.bias2 - b.v, s, wide(a.v)));
}

// ---------------------------------------------------------------- rows <-> values
// row r takes a_r (selects on the row-index bits)
template <class F> HS_RW rfe rsel4(const RowK<F> &k, const rfe &a0, const rfe &a1, const rfe &a2, const rfe &a3)
{
    return {sel(k.rhi, sel(k.rodd, a3.v, a2.v), sel(k.rodd, a1.v, a0.v))};
}
// rows 0, 1, 2, 3 take a0, a1, a2, a2
template <class F> HS_RW rfe rsel3(const RowK<F> &k, const rfe &a0, const rfe &a1, const rfe &a2)
{
    return {sel(k.rhi, a2.v, sel(k.rodd, a1.v, a0.v))};
}
// rows 0, 1, 2, 3 take a0, a1, a1, a1
template <class F> HS_RW rfe rsel2(const RowK<F> &k, const rfe &a0, const rfe &a1)
{
    return {sel(k.row0, a0.v, a1.v)};
}
// row r's value of p to every row: p0..p3 (permlane32 then permlane16 swaps)
HS_RW void rdist4(rfe &p0, rfe &p1, rfe &p2, rfe &p3, const rfe &p)
{
    u32v a = p.v, b = p.v;
    swap32(a, b); // a: rows (0, 1, 0, 1), b: rows (2, 3, 2, 3)
    u32v c = a, d = b;
    swap16(a, c);
    swap16(b, d);
    p0.v = a;
    p1.v = c;
    p2.v = b;
    p3.v = d;
}
HS_RW void rdist3(rfe &p0, rfe &p1, rfe &p2, const rfe &p)
{
    u32v a = p.v, b = p.v;
    swap32(a, b);
    u32v c = a, d = b;
    swap16(a, c);
    swap16(b, d);
    p0.v = a;
    p1.v = c;
    p2.v = b;
}
HS_RW void rdist2(rfe &p0, rfe &p1, const rfe &p)
{
    u32v a = p.v, b = p.v;
    swap32(a, b);
    u32v c = a;
    swap16(a, c);
    p0.v = a;
    p1.v = c;
}

// row 0's limbs as a (wave-uniform) field.h element
HS_RW void rfe_get(fe &r, const rfe &a)
{
#pragma unroll
    for (int i = 0; i < NL; i++)
        r.v[i] = readlane(a.v, i);
}

template <class F> HS_RW bool ris_zero(const rfe &a)
{
    fe t;
    rfe_get(t, a);
    return fe_is_zero<F>(t);
}

// a field.h element in memory (limbs < 2^29 + 2^17) -> the same element in every row
template <class F> HS_RW rfe rload(const RowK<F> &k, const fe *p) { return {load_if(k.lt9, p->v, k.l)}; }

} // namespace row
} // namespace hs
