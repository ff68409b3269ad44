// Fixed-base multi-scalar multiplication over resident tables of multiples.
//
// Restates the algorithm of helios_msm_fixed (helios/include/helios_msm_fixed.h:60-155) and its
// tables (helios_scalarmult_fixed_precompute, helios/include/helios_scalarmult_fixed.h:54-65):
// every base P_i carries the 16 affine multiples [1P_i .. 16P_i]; scalars are split into 52
// signed 5-bit digits, and the result is
//     sum_w 32^w * S_w,   S_w = sum_i sign(d_iw) * T_i[|d_iw|]
// so all MSMs over one base set share the 5 doublings per window (255 in total).
//
// GPU mapping: one wavefront per (MSM k, window w) accumulates S_kw over all bases (lane l takes
// bases l, l+64, ...) with mixed additions, then a 64-way LDS tree sum.  One quad of lanes per MSM
// runs the Horner chain over its 52 window sums.  The digit of window w is read locally from the
// canonical scalar: d = bits[5w, 5w+5) + bit(5w-1) - 32 * bit(5w+4), in [-16, 16]; this signed
// recoding differs from helios_scalar_recode_signed5 digit by digit but represents the same
// integer, and group arithmetic here is exact, so every result is the same point.
//
// Resident sets built from points use a 16-tooth signed-binary comb (FixedW<FIXED_COMB16>: 16
// columns, one shared doubling per column) or, when its tables do not fit, the same windowed
// scheme with wider windows (up to 14 bits), chosen by the table memory they need; the drop-in
// msm_fixed keeps the reference's width-5 tables.
//
// Table entries are canonical affine (x, y) as 16 little-endian words; an all-zero entry is the
// identity (the table of an identity base).
#pragma once
#include "msm_kernels.h"

namespace hs
{

constexpr int FIXED_WINDOWS = 52;
constexpr int FIXED_ENTRIES = 16;

// Table geometry for window width WB: 2^(WB-1) multiples per base, ceil(256 / WB) signed windows
// of WB doublings each.  WB = 5 is the reference's layout (the drop-in tables); resident sets built
// from points use WB = 8..14 (32..19 additions per base and MSM instead of 52).
template <int WB> struct FixedW
{
    static constexpr int entries = 1 << (WB - 1);
    static constexpr int windows = (256 + WB - 1) / WB;
    static constexpr int step = WB;
};

// The signed-binary comb with 16 teeth (width code 0).  A scalar s is made odd (s' = s, or
// s + order) and written as s' = sum_{i<256} b_i 2^i with every b_i = +-1.  Digit i = 16 j + r is
// tooth j of column r, so s' P = sum_r 2^r sum_j b_{16j+r} Q_j with Q_j = 2^(16 j) P.  Entry e of
// a base is Q_15 + sum_{j<15} (bit j of e ? Q_j : -Q_j): 2^15 entries (2 MiB per base), one
// column = one table lookup per base, and a pattern with b_{16*15+r} = -1 uses the entry of its
// negation, negated.  16 additions per base and MSM, like 16-bit windows, but only 15 shared
// doublings per MSM instead of 240.
constexpr int FIXED_COMB16 = 0;
template <> struct FixedW<FIXED_COMB16>
{
    static constexpr int entries = 1 << 15;
    static constexpr int windows = 16;
    static constexpr int step = 1;
};

// Bases from reference Jacobian structs to canonical affine entries (all-zero for the identity),
// once per base, ahead of k_fixed_tables.
template <class C>
__global__ void __launch_bounds__(256) k_fixed_bases(const uint64_t *__restrict__ pts, uint32_t n,
                                                     uint32_t *__restrict__ base_aff)
{
    using F = typename C::F;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint64_t *p = pts + 15ull * i;
    uint64_t l[5];
    jac J;
    for (int j = 0; j < 5; j++)
        l[j] = p[j];
    fe_from_r51<F>(J.X, l);
    for (int j = 0; j < 5; j++)
        l[j] = p[5 + j];
    fe_from_r51<F>(J.Y, l);
    for (int j = 0; j < 5; j++)
        l[j] = p[10 + j];
    fe_from_r51<F>(J.Z, l);
    aff a;
    if (fe_is_zero<F>(J.Z))
    {
        fe_zero(a.x);
        fe_zero(a.y);
    }
    else
        jac_to_aff<F>(a, J);
    store_aff<F>(base_aff, i, a);
}

// Tables from the affine bases: thread (i, k) computes (k+1) * P_i.  Optionally also writes the
// reference's helios_affine layout (x[5], y[5] radix-2^51, canonical).
template <class C, int ENTRIES>
__global__ void __launch_bounds__(256) k_fixed_tables(const uint32_t *__restrict__ base_aff, uint32_t n,
                                                      uint32_t *__restrict__ tables, uint64_t *__restrict__ ref_out)
{
    using F = typename C::F;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * ENTRIES)
        return;
    const uint32_t i = t / ENTRIES, k = t % ENTRIES;
    aff base, a;
    load_aff(base, base_aff, i);
    bool pt = !(fe_is_zero<F>(base.x) && fe_is_zero<F>(base.y));
    if (pt)
    {
        xyzz b, m;
        xyzz_from_aff<F>(b, base);
        xyzz_mul_small<F>(m, b, k + 1);
        pt = xyzz_to_aff<F>(a, m);
    }
    if (!pt)
    {
        fe_zero(a.x);
        fe_zero(a.y);
    }
    store_aff<F>(tables, t, a);
    if (ref_out)
    {
        uint64_t *o = ref_out + 10ull * t;
        fe_to_r51<F>(o, a.x);
        fe_to_r51<F>(o + 5, a.y);
    }
}

// Comb tables, step 1: the tooth bases Q_j = 2^(16 j) P_i as canonical affine entries (all-zero
// for the identity), one thread per (base i, tooth j).
template <class C>
__global__ void __launch_bounds__(256) k_comb_teeth(const uint32_t *__restrict__ base_aff, uint32_t n,
                                                    uint32_t *__restrict__ teeth)
{
    using F = typename C::F;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * 16)
        return;
    const uint32_t i = t >> 4, j = t & 15;
    aff b, a;
    load_aff(b, base_aff, i);
    bool pt = !(fe_is_zero<F>(b.x) && fe_is_zero<F>(b.y));
    if (pt)
    {
        xyzz q;
        xyzz_from_aff<F>(q, b);
        for (uint32_t d = 0; d < 16 * j; d++)
            xyzz_dbl<F>(q, q);
        pt = xyzz_to_aff<F>(a, q);
    }
    if (!pt)
    {
        fe_zero(a.x);
        fe_zero(a.y);
    }
    store_aff<F>(teeth, t, a);
}

// Comb tables, step 2: entry e of base i, Q_15 + sum_{j<15} (bit j of e ? Q_j : -Q_j), is
// A_(e mod 256) + B_(e >> 8) with
//   A_l = sum_{j<8} (bit j of l ? Q_j : -Q_j)                     (256 per base)
//   B_h = Q_15 + sum_{j<7} (bit j of h ? Q_(8+j) : -Q_(8+j))       (128 per base)
// The half tables come from quarter tables the same way (step 2a below).
constexpr uint32_t COMB_HALF = 384;

// COMB_RUN table entries B + A_k, A_k = A[a0 + a_step k] (canonical affine, all-zero = identity),
// written to out[o0 + o_step k] for the k whose bit is set in `valid`.  One inversion serves all
// the denominators x_B - x_A of the terms that are not the identity (Montgomery's trick):
//   lambda = (y_B - y_A) / (x_B - x_A),  x = lambda^2 - x_A - x_B,  y = lambda (x_A - x) - y_A;
// an identity term gives the other one.  Should a denominator be zero (B = +-A_k), every entry takes
// the general addition.
constexpr int COMB_RUN = 16;
static_assert(COMB_RUN <= 32 && 32 % COMB_RUN == 0, "runs split the 32 low half entries evenly");
__device__ __forceinline__ bool aff_is_zero_words(const aff &a)
{
    uint32_t o = 0;
#pragma unroll
    for (int j = 0; j < NL; j++)
        o |= a.x.v[j] | a.y.v[j];
    return o == 0;
}

template <class F>
__device__ __forceinline__ void pair_sums(const aff &b, const uint32_t *__restrict__ A, uint32_t a0, uint32_t a_step,
                                          uint32_t *__restrict__ out, uint32_t o0, uint32_t o_step, uint32_t valid)
{
    if (aff_is_zero_words(b))
    {
        // B is the identity: the entries are the A_k
#pragma unroll
        for (int k = 0; k < COMB_RUN; k++)
            if ((valid >> k) & 1u)
            {
                aff a;
                load_aff(a, A, a0 + a_step * k);
                store_aff<F>(out, o0 + o_step * k, a);
            }
        return;
    }
    fe one;
    fe_zero(one);
    one.v[0] = 1;
    fe c[COMB_RUN]; // prefix products of the denominators (1 for an identity A_k)
    uint32_t az = 0; // identity A_k
#pragma unroll
    for (int k = 0; k < COMB_RUN; k++)
    {
        aff a;
        load_aff(a, A, a0 + a_step * k);
        fe d;
        fe_sub<F>(d, b.x, a.x);
        if (aff_is_zero_words(a))
        {
            az |= 1u << k;
            d = one;
        }
        if (k == 0)
            c[0] = d;
        else
            fe_mul<F>(c[k], c[k - 1], d);
    }
    if (fe_is_zero<F>(c[COMB_RUN - 1]))
    {
        for (int k = 0; k < COMB_RUN; k++)
        {
            if (!((valid >> k) & 1u))
                continue;
            aff a, o;
            load_aff(a, A, a0 + a_step * k);
            xyzz acc;
            bool acc_id = true;
            xyzz_identity<F>(acc);
            if (!aff_is_zero_words(a))
                xyzz_madd<F>(acc, acc_id, a);
            xyzz_madd<F>(acc, acc_id, b);
            if (acc_id || !xyzz_to_aff<F>(o, acc))
            {
                fe_zero(o.x);
                fe_zero(o.y);
            }
            store_aff<F>(out, o0 + o_step * k, o);
        }
        return;
    }
    fe inv;
    fe_inv<F>(inv, c[COMB_RUN - 1]);
#pragma unroll
    for (int k = COMB_RUN - 1; k >= 0; k--)
    {
        aff a;
        load_aff(a, A, a0 + a_step * k);
        const bool zk = (az >> k) & 1u;
        fe ik, t0, lam;
        if (k > 0)
        {
            fe_mul<F>(ik, inv, c[k - 1]);
            fe_sub<F>(t0, b.x, a.x);
            if (zk)
                t0 = one;
            fe_mul<F>(inv, inv, t0);
        }
        else
            ik = inv;
        if ((valid >> k) & 1u)
        {
            aff o;
            fe_sub<F>(t0, b.y, a.y);
            fe_mul<F>(lam, t0, ik);
            fe_sq<F>(t0, lam);
            fe_sub2<F>(o.x, t0, a.x, b.x);
            fe_sub<F>(t0, a.x, o.x);
            fe_mul<F>(o.y, lam, t0);
            fe_sub<F>(o.y, o.y, a.y);
            store_aff<F>(out, o0 + o_step * k, zk ? b : o);
        }
    }
}

// Step 2b: thread (i, h, r) forms the COMB_RUN entries (h << 8) | (r + S k) of base i as B_h +
// A_(r + S k), S = 256 / COMB_RUN; adjacent threads write adjacent entries.  For a point of prime order no half entry
// is the identity and no denominator is zero: A_l = +-B_h would need sum_{j<8} +-2^(16 j) =
// +-(2^240 + sum_{8<=j<15} +-2^(16 j)) modulo the order, with both sides in (-2^241, 2^241).  An
// identity base has all-zero half tables and gets all-zero entries.
template <class C>
__global__ void __launch_bounds__(256) k_comb_tables_runs(const uint32_t *__restrict__ halves, uint32_t n,
                                                          uint32_t *__restrict__ tables)
{
    using F = typename C::F;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    constexpr uint32_t S = 256 / COMB_RUN;
    if (t >= n * (128u * S))
        return;
    const uint32_t i = t / (128u * S), h = (t / S) % 128u, r = t % S;
    aff b;
    load_aff(b, halves, COMB_HALF * i + 256 + h);
    pair_sums<F>(b, halves, COMB_HALF * i + r, S, tables, (i << 15) | (h << 8) | r, S,
                 COMB_RUN == 32 ? 0xffffffffu : (1u << COMB_RUN) - 1u);
}

// Step 2a: the half tables from four quarter tables per base (COMB_QUARTER entries, one thread
// each, from the tooth bases):
//   QA0_u = sum_{j<4} +-Q_j,  QA1_v = sum_{4<=j<8} +-Q_j,  QB0_u = sum_{j<4} +-Q_(8+j)  (16 each),
//   QB1_v = Q_15 + sum_{j<3} +-Q_(12+j)  (8),
// so A_l = QA0_(l mod 16) + QA1_(l >> 4) and B_h = QB0_(h mod 16) + QB1_(h >> 4).  For a point of
// prime order the two terms are never equal or opposite (their multiples of P differ in size:
// below 2^49 against at least 2^111, and below 2^177 against at least 2^239).
constexpr uint32_t COMB_QUARTER = 56;
template <class C>
__global__ void __launch_bounds__(256) k_comb_quarters(const uint32_t *__restrict__ teeth, uint32_t n,
                                                       uint32_t *__restrict__ quarters)
{
    using F = typename C::F;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * COMB_QUARTER)
        return;
    const uint32_t i = t / COMB_QUARTER, u = t % COMB_QUARTER;
    const uint32_t part = u < 48 ? u >> 4 : 3, bits = u < 48 ? (u & 15u) : u - 48;
    const int first = (int)(4 * part), count = part < 3 ? 4 : 3; // tooth bases of this quarter
    aff q, a;
    load_aff(q, teeth, 16 * i + 15);
    bool pt = !aff_is_zero_words(q);
    if (pt)
    {
        xyzz acc;
        bool acc_id = true;
        xyzz_identity<F>(acc);
        if (part == 3)
            xyzz_madd<F>(acc, acc_id, q);
        for (int j = 0; j < count; j++)
        {
            load_aff(q, teeth, 16 * i + first + j);
            if (!((bits >> j) & 1u))
                fe_neg<F>(q.y, q.y);
            xyzz_madd<F>(acc, acc_id, q);
        }
        pt = !acc_id && xyzz_to_aff<F>(a, acc);
    }
    if (!pt)
    {
        fe_zero(a.x);
        fe_zero(a.y);
    }
    store_aff<F>(quarters, t, a);
}

// Step 2a, continued: thread (i, v) of base i forms 16 half entries, A_(16 v + k) for v < 16 and
// B_(16 (v - 16) + k) for 16 <= v < 24.
template <class C>
__global__ void __launch_bounds__(256) k_comb_halves(const uint32_t *__restrict__ quarters, uint32_t n,
                                                     uint32_t *__restrict__ halves)
{
    using F = typename C::F;
    static_assert(COMB_RUN == 16, "one run per quarter-table row");
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * 24u)
        return;
    const uint32_t i = t / 24u, v = t % 24u;
    const bool hi = v >= 16;
    const uint32_t q0 = COMB_QUARTER * i;
    aff b;
    load_aff(b, quarters, hi ? q0 + 48 + (v - 16) : q0 + 16 + v);
    pair_sums<F>(b, quarters, hi ? q0 + 32 : q0, 1, halves, COMB_HALF * i + (hi ? 256 + 16 * (v - 16) : 16 * v), 1,
                 0xffffu);
}

// Windowed tables (entries (k+1) P, k < E = 2^(WB-1), WB >= 8) by the same construction: with
// m = k + 1 = 32 a + b (0 <= b < 32, 0 <= a <= E / 32), entry m is H_a + L_b over the half tables
// L_b = b P and H_a = 32 a P (the identity at b = 0 and a = 0), WIN_HALF(E) entries per base,
// L first.  For a point of prime order the denominators of the sums with a, b > 0 are nonzero
// (32 a != +-b modulo the order); a sum with a = 0 or b = 0 is its other term.
__host__ __device__ constexpr uint32_t win_half(uint32_t E) { return 32 + E / 32 + 1; }

// Step 1: one thread per (base, half entry)
template <class C, int E>
__global__ void __launch_bounds__(256) k_win_halves(const uint32_t *__restrict__ base_aff, uint32_t n,
                                                    uint32_t *__restrict__ halves)
{
    using F = typename C::F;
    constexpr uint32_t NH = win_half(E);
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * NH)
        return;
    const uint32_t i = t / NH, u = t % NH, k = u < 32 ? u : 32 * (u - 32);
    aff base, a;
    load_aff(base, base_aff, i);
    bool pt = k != 0 && !aff_is_zero_words(base);
    if (pt)
    {
        xyzz b, m;
        xyzz_from_aff<F>(b, base);
        xyzz_mul_small<F>(m, b, k);
        pt = xyzz_to_aff<F>(a, m);
    }
    if (!pt)
    {
        fe_zero(a.x);
        fe_zero(a.y);
    }
    store_aff<F>(halves, t, a);
}

// Step 2: thread (i, a, r) forms the entries m = 32 a + r + S j (j < COMB_RUN, S = 32 / COMB_RUN) in
// [1, E] of base i; adjacent threads write adjacent entries.
template <class C, int E>
__global__ void __launch_bounds__(256) k_win_tables_runs(const uint32_t *__restrict__ halves, uint32_t n,
                                                         uint32_t *__restrict__ tables)
{
    using F = typename C::F;
    constexpr uint32_t NH = win_half(E), NA = E / 32 + 1;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    constexpr uint32_t S = 32 / COMB_RUN;
    if (t >= n * NA * S)
        return;
    const uint32_t i = t / (NA * S), a = (t / S) % NA, r = t % S;
    uint32_t valid = 0;
#pragma unroll
    for (int j = 0; j < COMB_RUN; j++)
    {
        const uint32_t m = 32 * a + r + S * j;
        valid |= (m >= 1 && m <= (uint32_t)E) ? 1u << j : 0u;
    }
    if (!valid)
        return;
    aff h;
    load_aff(h, halves, NH * i + 32 + a);
    // entry m lives at index m - 1; for m = 0 (a = r = 0, j = 0: not valid) the index is unused
    pair_sums<F>(h, halves, NH * i + r, S, tables, E * i + 32 * a + r - 1, S, valid);
}

// Tables given in the reference's helios_affine layout (possibly lazy limbs): canonicalise.
// (0, 0) stays the all-zero identity entry.
template <class C>
__global__ void __launch_bounds__(256) k_fixed_tables_from_ref(const uint64_t *__restrict__ ref_in, uint32_t entries,
                                                               uint32_t *__restrict__ tables)
{
    using F = typename C::F;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= entries)
        return;
    const uint64_t *p = ref_in + 10ull * t;
    uint64_t l[5];
    aff a;
    fe v;
    for (int j = 0; j < 5; j++)
        l[j] = p[j];
    fe_from_r51<F>(v, l);
    fe_canon<F>(a.x, v);
    for (int j = 0; j < 5; j++)
        l[j] = p[5 + j];
    fe_from_r51<F>(v, l);
    fe_canon<F>(a.y, v);
    store_aff<F>(tables, t, a);
}

// Canonical scalar (< order < 2^255) -> the 16 column patterns of its signed-binary comb form
// (FixedW<FIXED_COMB16>): s' = s, or s + order when s is even, so s' is odd and below 2^256; then
// s' = sum_i (2 u_i - 1) 2^i for u = floor(s' / 2) + 2^255.  Pattern r (16 bits, halfword r of the
// output) holds bit 16 j + r of u at bit j.
template <class O> HS_HD void comb_recode(uint32_t k[8])
{
    if (!(k[0] & 1u))
    {
        uint64_t t = 0;
#pragma unroll
        for (int i = 0; i < 8; i++)
        {
            t = (uint64_t)k[i] + O::mw(i) + (t >> 32);
            k[i] = (uint32_t)t;
        }
    }
    uint32_t u[8];
#pragma unroll
    for (int i = 0; i < 7; i++)
        u[i] = (k[i] >> 1) | (k[i + 1] << 31);
    u[7] = (k[7] >> 1) | 0x80000000u;
#pragma unroll
    for (int i = 0; i < 8; i++)
        k[i] = 0;
#pragma unroll
    for (int r = 0; r < 16; r++)
    {
        uint32_t pr = 0;
#pragma unroll
        for (int j = 0; j < 16; j++)
            pr |= ((u[j >> 1] >> (16 * (j & 1) + r)) & 1u) << j;
        k[r >> 1] |= pr << (16 * (r & 1));
    }
}

// Scalars -> canonical words mod the group order, with the reference's window-8 substitution
// when `w8_compat` is set (each row then matches helios_msm_vartime over the same bases); with
// `comb`, then recoded to comb column patterns.
template <class C>
__global__ void __launch_bounds__(256) k_fixed_prep(const uint32_t *__restrict__ in, size_t count,
                                                    uint32_t *__restrict__ out, uint32_t w8_compat, uint32_t comb)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count)
        return;
    const uint4 *sp = reinterpret_cast<const uint4 *>(in + 8 * t);
    const uint4 a = sp[0], b = sp[1];
    const uint32_t s[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    uint32_t k[8];
    scalar_prepare<typename C::O>(k, s, w8_compat != 0);
    if (comb)
        comb_recode<typename C::O>(k);
    uint4 *op = reinterpret_cast<uint4 *>(out + 8 * t);
    op[0] = make_uint4(k[0], k[1], k[2], k[3]);
    op[1] = make_uint4(k[4], k[5], k[6], k[7]);
}

// The signed WB-bit digit of window w of a canonical scalar (< 2^255), in [-2^(WB-1), 2^(WB-1)],
// from scalar words wi = fixed_digit_word<WB>(w) (lo) and wi + 1 (hi, 0 past the top).
template <int WB> __device__ __forceinline__ int fixed_digit_word(int w)
{
    return w == 0 ? 0 : (WB * w - 1) >> 5;
}

template <int WB> __device__ __forceinline__ int fixed_digit_from(uint32_t lo, uint32_t hi, int w)
{
    constexpr uint32_t mask = (2u << WB) - 1; // the window's WB bits and the bit below it
    uint32_t v;
    if (w == 0)
        v = (lo << 1) & mask;
    else
    {
        const int off = (WB * w - 1) & 31;
        v = (uint32_t)((((uint64_t)lo | ((uint64_t)hi << 32)) >> off) & mask);
    }
    const int raw = (int)(v >> 1), cin = (int)(v & 1u);
    return raw + cin - ((raw >> (WB - 1)) << WB);
}

// Comb column r (halfword r of the prepared row): +-(entry + 1) of the pattern, never 0.
template <> __device__ __forceinline__ int fixed_digit_word<FIXED_COMB16>(int r)
{
    return r >> 1;
}

template <> __device__ __forceinline__ int fixed_digit_from<FIXED_COMB16>(uint32_t lo, uint32_t, int r)
{
    const uint32_t p = (lo >> (16 * (r & 1))) & 0xffffu;
    return (p & 0x8000u) ? (int)(p & 0x7fffu) + 1 : -(int)((~p & 0x7fffu) + 1);
}

// One wavefront per (MSM k, CPW consecutive windows w): S[k * windows + w] = sum_i T_i[d_kiw]
// (Jacobian).
// 64 / CPW lanes share a window (lane l takes bases l, l + 64 / CPW, ...), then an LDS tree sum
// per window; more windows per wavefront mean a shorter tree per addition.
template <class C, int WB, int CPW>
__global__ void __launch_bounds__(64) k_fixed_window_sums(const uint32_t *__restrict__ tables,
                                                          const uint32_t *__restrict__ scal, uint32_t n,
                                                          jac *__restrict__ S)
{
    using F = typename C::F;
    constexpr int WINDOWS = FixedW<WB>::windows, ENTRIES = FixedW<WB>::entries;
    constexpr uint32_t LANES = 64 / CPW, GROUPS = (WINDOWS + CPW - 1) / CPW;
    static_assert(64 % CPW == 0, "windows per wavefront must divide 64");
    __shared__ xyzz sh[64];
    const uint32_t k = blockIdx.x / GROUPS, lane = threadIdx.x % LANES;
    const int w = (int)(blockIdx.x % GROUPS) * CPW + (int)(threadIdx.x / LANES);
    const uint32_t nn = w < WINDOWS ? n : 0u; // lanes past the last window sum nothing
    const uint32_t *sk = scal + (size_t)k * n * 8;
    xyzz acc;
    bool acc_id = true;
    xyzz_identity<F>(acc);
    // software pipeline over this lane's bases: the scalar words two bases ahead and the table
    // entry one base ahead are in flight during the current addition
    const int wi = fixed_digit_word<WB>(w);
    uint32_t lo = 0, hi = 0;
    auto words = [&](uint32_t ii) {
        const uint32_t *r = sk + 8ull * ii;
        lo = r[wi];
        hi = wi + 1 < 8 ? r[wi + 1] : 0u;
    };
    int dn = 0;
    uint4 qn[4] = {};
    auto entry = [&](uint32_t ii) {
        dn = fixed_digit_from<WB>(lo, hi, w);
        if (dn != 0)
        {
            const uint4 *p =
                reinterpret_cast<const uint4 *>(tables + 16ull * (ii * ENTRIES + (uint32_t)((dn < 0 ? -dn : dn) - 1)));
            qn[0] = p[0];
            qn[1] = p[1];
            qn[2] = p[2];
            qn[3] = p[3];
        }
    };
    if (lane < nn)
    {
        words(lane);
        entry(lane);
    }
    if (lane + LANES < nn)
        words(lane + LANES);
    for (uint32_t i = lane; i < nn; i += LANES)
    {
        const int d = dn;
        const uint4 q0 = qn[0], q1 = qn[1], q2 = qn[2], q3 = qn[3];
        if (i + LANES < nn)
        {
            entry(i + LANES);
            if (i + 2 * LANES < nn)
                words(i + 2 * LANES);
        }
        if (d == 0)
            continue;
        const uint32_t xw[8] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w};
        const uint32_t yw[8] = {q2.x, q2.y, q2.z, q2.w, q3.x, q3.y, q3.z, q3.w};
        uint32_t any = 0;
#pragma unroll
        for (int j = 0; j < 8; j++)
            any |= xw[j] | yw[j];
        if (any == 0)
            continue; // identity base (all-zero entry)
        aff q;
        fe_from_words(q.x, xw);
        fe_from_words(q.y, yw);
        // a negative digit adds -T: y -> 65 m - y limb by limb (lazy limbs below 2^30, which
        // xyzz_madd accepts for the base's y)
        const uint32_t keep = d < 0 ? 0u : 0xffffffffu;
#pragma unroll
        for (int j = 0; j < NL; j++)
            q.y.v[j] = (q.y.v[j] & keep) | ((F::bias2(j) - q.y.v[j]) & ~keep);
        xyzz_madd<F>(acc, acc_id, q);
    }
    if (acc_id)
        xyzz_identity<F>(acc);
    store_xyzz(sh + threadIdx.x, acc);
    __syncthreads();
    for (uint32_t h = LANES / 2; h > 0; h >>= 1)
    {
        if (lane < h)
        {
            xyzz a, b;
            load_xyzz(a, sh + threadIdx.x);
            load_xyzz(b, sh + threadIdx.x + h);
            xyzz_add<F>(a, a, b);
            store_xyzz(sh + threadIdx.x, a);
        }
        __syncthreads();
    }
    if (lane == 0 && w < WINDOWS)
    {
        // stored in Jacobian form, so the serial Horner chain needs no conversion
        xyzz r;
        load_xyzz(r, sh + threadIdx.x);
        jac J;
        xyzz_to_jac<F>(J, r);
        S[(size_t)k * WINDOWS + w] = J;
    }
}

// One quad of lanes per MSM: Horner over the window sums (step shared doublings per window),
// then the 32-byte encoding and the reference struct (affine, Z = 1; identity 1:1:0).
template <class C, int WB>
__global__ void __launch_bounds__(64) k_fixed_horner(const jac *__restrict__ S, uint32_t K,
                                                     uint32_t *__restrict__ out32, uint64_t *__restrict__ outjac)
{
    using F = typename C::F;
    constexpr int WINDOWS = FixedW<WB>::windows;
    const uint32_t k = (blockIdx.x * blockDim.x + threadIdx.x) >> 2;
    if (k >= K)
        return;
    const jac *sk = S + (size_t)k * WINDOWS;
    jac acc, t;
    acc = sk[WINDOWS - 1];
    for (int w = WINDOWS - 2; w >= 0; w--)
    {
        if constexpr (FixedW<WB>::step == 1)
            jac_dbl_q<F>(acc); // 4 product levels; the pipelined form needs 5 for one doubling
        else
            jac_dbl_n_q<F>(acc, FixedW<WB>::step);
        t = sk[w];
        jac_add_q<F>(acc, acc, t);
    }
    if ((threadIdx.x & 3) != 0)
        return;
    aff a;
    const bool pt = jac_to_aff<F>(a, acc);
    uint32_t enc[8];
    encode_aff<F>(enc, a, pt);
    if (out32)
        for (int j = 0; j < 8; j++)
            out32[8ull * k + j] = enc[j];
    if (outjac)
        store_r51_aff<F>(outjac + 15ull * k, a, pt);
}

} // namespace hs
