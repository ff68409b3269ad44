// Test infrastructure: the product's field and curve arithmetic headers compiled for the host
// CPU, so the exact same limb code can be checked against Python big integers and the oracle
// without a GPU.  Built by tests/native/Makefile; never part of the product library.
#include <stdint.h>
#include <string.h>
#define HS_HD inline
#include "../../helioselene_amd/csrc/field.h"

using namespace hs;

static void ldw(uint32_t w[8], const uint8_t *b)
{
    for (int i = 0; i < 8; i++)
        w[i] = (uint32_t)b[4 * i] | ((uint32_t)b[4 * i + 1] << 8) | ((uint32_t)b[4 * i + 2] << 16) |
               ((uint32_t)b[4 * i + 3] << 24);
}
static void stw(uint8_t *b, const uint32_t w[8])
{
    for (int i = 0; i < 8; i++)
        for (int j = 0; j < 4; j++)
            b[4 * i + j] = (uint8_t)(w[i] >> (8 * j));
}
// 32 bytes (any value < 2^256) -> limbs
static void ld(fe &r, const uint8_t *b)
{
    uint32_t w[8];
    ldw(w, b);
    fe_from_words(r, w);
}
// limbs -> canonical 32 bytes (field k decides the modulus)
static int g_field = 0;
static void st(uint8_t *b, const fe &a)
{
    uint32_t w[8];
    if (g_field == 0)
        fe_to_words_canon<Fp>(w, a);
    else
        fe_to_words_canon<Fq>(w, a);
    stw(b, w);
}

template <class F> static void op(int o, uint8_t *out, const uint8_t *pa, const uint8_t *pb)
{
    fe a, b, r;
    ld(a, pa);
    ld(b, pb);
    switch (o)
    {
    case 0: fe_mul<F>(r, a, b); break;
    case 1: fe_sq<F>(r, a); break;
    case 2: fe_add<F>(r, a, b); break;
    case 3: fe_sub<F>(r, a, b); break;
    case 4: fe_inv<F>(r, a); break;
    case 5: fe_canon<F>(r, a); break;
    case 6: fe_neg<F>(r, a); break;
    case 7: fe_mul_small<F>(r, a, pb[0] | (pb[1] << 8)); break;
    case 8: fe_inv_fermat<F>(r, a); break;
    default: fe_zero(r);
    }
    st(out, r);
}

// raw (non-canonical) limbs -> value as 36 bytes (checks the < 2^261 + small bound)
static void st_raw(uint8_t *b, const fe &a)
{
    unsigned __int128 acc = 0;
    int bits = 0, o = 0;
    memset(b, 0, 40);
    for (int i = 0; i < NL; i++)
    {
        // add a.v[i] << (29 i) into the little-endian byte array
        unsigned __int128 x = (unsigned __int128)a.v[i] << ((29 * i) % 8);
        int byte = (29 * i) / 8;
        unsigned carry = 0;
        for (int k = byte; k < 40; k++)
        {
            unsigned v = b[k] + (unsigned)(x & 0xff) + carry;
            b[k] = (uint8_t)v;
            carry = v >> 8;
            x >>= 8;
            if (!x && !carry)
                break;
        }
    }
    (void)acc; (void)bits; (void)o;
}

extern "C" void hst_fe_op_raw(int field, int o, uint8_t *out40, const uint8_t *pa, const uint8_t *pb)
{
    fe a, b, r;
    ld(a, pa);
    ld(b, pb);
    if (field == 0)
    {
        switch (o) { case 0: fe_mul<Fp>(r, a, b); break; case 1: fe_sq<Fp>(r, a); break; case 2: fe_add<Fp>(r, a, b); break; default: fe_sub<Fp>(r, a, b); }
    }
    else
    {
        switch (o) { case 0: fe_mul<Fq>(r, a, b); break; case 1: fe_sq<Fq>(r, a); break; case 2: fe_add<Fq>(r, a, b); break; default: fe_sub<Fq>(r, a, b); }
    }
    st_raw(out40, r);
    uint32_t mx = 0;
    for (int i = 0; i < NL; i++)
        mx = r.v[i] > mx ? r.v[i] : mx;
    out40[39] = mx > (1u << 29) ? 1 : 0; // limb bound violated
}

// ops on raw limbs (any normalised limbs, up to 2^29 each): result value (40 bytes, byte 39 = limb
// bound flag); o = 4: the zero test, o = 5: the zero test for fe_norm / reduce_cols outputs
extern "C" int hst_fe_op_limbs(int field, int o, uint8_t *out40, const uint32_t *la, const uint32_t *lb)
{
    fe a, b, r;
    for (int i = 0; i < NL; i++)
    {
        a.v[i] = la[i];
        b.v[i] = lb[i];
    }
    if (o == 4)
        return field == 0 ? (int)fe_is_zero<Fp>(a) : (int)fe_is_zero<Fq>(a);
    if (o == 5)
        return field == 0 ? (int)fe_is_zero_norm<Fp>(a) : (int)fe_is_zero_norm<Fq>(a);
    if (o >= 16 && o < 32) // a - k b, k = o - 16
    {
        if (field == 0)
            fe_sub_small_mul<Fp>(r, a, b, (uint32_t)(o - 16));
        else
            fe_sub_small_mul<Fq>(r, a, b, (uint32_t)(o - 16));
        st_raw(out40, r);
        uint32_t mx = 0;
        for (int i = 0; i < NL; i++)
            mx = r.v[i] > mx ? r.v[i] : mx;
        out40[39] = mx > (1u << 29) ? 1 : 0;
        return 0;
    }
    if (field == 0)
    {
        switch (o) { case 0: fe_mul<Fp>(r, a, b); break; case 1: fe_sq<Fp>(r, a); break; case 2: fe_add<Fp>(r, a, b); break; default: fe_sub<Fp>(r, a, b); }
    }
    else
    {
        switch (o) { case 0: fe_mul<Fq>(r, a, b); break; case 1: fe_sq<Fq>(r, a); break; case 2: fe_add<Fq>(r, a, b); break; default: fe_sub<Fq>(r, a, b); }
    }
    st_raw(out40, r);
    uint32_t mx = 0;
    for (int i = 0; i < NL; i++)
        mx = r.v[i] > mx ? r.v[i] : mx;
    out40[39] = mx > (1u << 29) ? 1 : 0;
    return 0;
}

// fe_mul_sum2(a, b, c, d) on raw limbs (out40 as hst_fe_op_limbs); with lazy != 0 the operands
// are formed as in the point formulas: b = fe_sub_lazy2(lb, lb2), c = fe_neg_lazy2(lc)
template <class F>
static void mul_sum2_limbs(uint8_t *out40, const uint32_t *la, const uint32_t *lb, const uint32_t *lb2,
                           const uint32_t *lc, const uint32_t *ld, int lazy)
{
    fe a, b, b2, c, d, r;
    for (int i = 0; i < NL; i++)
    {
        a.v[i] = la[i];
        b.v[i] = lb[i];
        b2.v[i] = lb2[i];
        c.v[i] = lc[i];
        d.v[i] = ld[i];
    }
    if (lazy)
    {
        fe_sub_lazy2<F>(b, b, b2);
        fe_neg_lazy2<F>(c, c);
    }
    fe_mul_sum2<F>(r, a, b, c, d);
    st_raw(out40, r);
    uint32_t mx = 0;
    for (int i = 0; i < NL; i++)
        mx = r.v[i] > mx ? r.v[i] : mx;
    out40[39] = mx > (1u << 29) ? 1 : 0;
}

extern "C" void hst_fe_mul_sum2(int field, uint8_t *out40, const uint32_t *la, const uint32_t *lb, const uint32_t *lb2,
                                const uint32_t *lc, const uint32_t *ld, int lazy)
{
    if (field == 0)
        mul_sum2_limbs<Fp>(out40, la, lb, lb2, lc, ld, lazy);
    else
        mul_sum2_limbs<Fq>(out40, la, lb, lb2, lc, ld, lazy);
}

// mode 0: a b + fe_neg_lazy2(c); mode 1: a^2 + fe_neg3_lazy(c, d, e) (out40 as hst_fe_op_limbs)
template <class F>
static void mul_add_limbs(uint8_t *out40, const uint32_t *la, const uint32_t *lb, const uint32_t *lc,
                          const uint32_t *ld, const uint32_t *le, int mode)
{
    fe a, b, c, d, e, t, r;
    for (int i = 0; i < NL; i++)
    {
        a.v[i] = la[i];
        b.v[i] = lb[i];
        c.v[i] = lc[i];
        d.v[i] = ld[i];
        e.v[i] = le[i];
    }
    if (mode == 0)
    {
        fe_neg_lazy2<F>(t, c);
        fe_mul_add<F>(r, a, b, t);
    }
    else
    {
        fe_neg3_lazy<F>(t, c, d, e);
        fe_sq_add<F>(r, a, t);
    }
    st_raw(out40, r);
    uint32_t mx = 0;
    for (int i = 0; i < NL; i++)
        mx = r.v[i] > mx ? r.v[i] : mx;
    out40[39] = mx > (1u << 29) ? 1 : 0;
}

extern "C" void hst_fe_mul_add(int field, uint8_t *out40, const uint32_t *la, const uint32_t *lb, const uint32_t *lc,
                               const uint32_t *ld, const uint32_t *le, int mode)
{
    if (field == 0)
        mul_add_limbs<Fp>(out40, la, lb, lc, ld, le, mode);
    else
        mul_add_limbs<Fq>(out40, la, lb, lc, ld, le, mode);
}

extern "C" void hst_fe_op(int field, int o, uint8_t *out, const uint8_t *a, const uint8_t *b)
{
    g_field = field;
    if (field == 0)
        op<Fp>(o, out, a, b);
    else
        op<Fq>(o, out, a, b);
}

extern "C" void hst_fe_from_r51(int field, uint8_t *out, const uint64_t *l)
{
    g_field = field;
    fe r;
    if (field == 0)
        fe_from_r51<Fp>(r, l);
    else
        fe_from_r51<Fq>(r, l);
    st(out, r);
}

extern "C" void hst_fe_to_r51(int field, uint64_t *l, const uint8_t *in)
{
    fe a;
    ld(a, in);
    if (field == 0)
        fe_to_r51<Fp>(l, a);
    else
        fe_to_r51<Fq>(l, a);
}

#include "../../helioselene_amd/csrc/curve.h"

// point checks on affine inputs given as canonical (x, y) 64-byte blobs
template <class F> static void pop(int o, uint8_t *out, const uint8_t *pa, const uint8_t *pb, uint32_t k)
{
    aff a, b, r;
    ld(a.x, pa);
    ld(a.y, pa + 32);
    ld(b.x, pb);
    ld(b.y, pb + 32);
    xyzz A, B, R;
    bool aid = false;
    xyzz_from_aff<F>(A, a);
    xyzz_from_aff<F>(B, b);
    switch (o)
    {
    case 0: R = A; xyzz_madd<F>(R, aid, b); if (aid) xyzz_identity<F>(R); break;
    case 1: xyzz_add<F>(R, A, B); break;
    case 2: xyzz_dbl<F>(R, A); break;
    case 3: xyzz_mdbl<F>(R, a); break;
    case 4: xyzz_mul_small<F>(R, A, k); break;
    case 5: { xyzz t; xyzz_dbl<F>(t, A); xyzz_add<F>(R, t, B); break; } // non-unit ZZ input
    default: xyzz_identity<F>(R);
    }
    bool ok = xyzz_to_aff<F>(r, R);
    st(out, r.x);
    st(out + 32, r.y);
    out[64] = ok;
}

// sum of n affine points (x || y, 64 bytes each; all-zero = identity, skipped) accumulated with
// xyzz_madd in order, as the bucket kernels do; neg[i] != 0 adds -P_i
template <class F> static void madd_chain(uint8_t *out, const uint8_t *pts, const uint8_t *neg, size_t n)
{
    xyzz acc;
    bool aid = true;
    xyzz_identity<F>(acc);
    for (size_t i = 0; i < n; i++)
    {
        int nz = 0;
        for (int j = 0; j < 64; j++)
            nz |= pts[64 * i + j];
        if (!nz)
            continue;
        aff q;
        ld(q.x, pts + 64 * i);
        ld(q.y, pts + 64 * i + 32);
        if (neg[i])
            fe_neg_lazy2<F>(q.y, q.y); // as the bucket kernels negate
        xyzz_madd<F>(acc, aid, q);
    }
    if (aid)
        xyzz_identity<F>(acc);
    aff r;
    bool ok = xyzz_to_aff<F>(r, acc);
    st(out, r.x);
    st(out + 32, r.y);
    out[64] = ok;
}

// The bucket walk's form of the same sum: each point arrives with its sign applied (y or a
// normalised -y, as the base records store them), the first non-identity point starts the run
// (xyzz_from_aff) and the rest go through xyzz_madd_hot, as k_bucket_flat does.
template <class F> static void walk_chain(uint8_t *out, const uint8_t *pts, const uint8_t *neg, size_t n)
{
    xyzz acc;
    bool aid = true, started = false;
    xyzz_identity<F>(acc);
    for (size_t i = 0; i < n; i++)
    {
        int nz = 0;
        for (int j = 0; j < 64; j++)
            nz |= pts[64 * i + j];
        if (!nz)
            continue;
        aff q;
        ld(q.x, pts + 64 * i);
        ld(q.y, pts + 64 * i + 32);
        if (neg[i])
            fe_neg<F>(q.y, q.y);
        if (!started)
        {
            xyzz_from_aff<F>(acc, q);
            aid = false;
            started = true;
            continue;
        }
        xyzz_madd_hot<F>(acc, aid, q);
    }
    if (aid)
        xyzz_identity<F>(acc);
    aff r;
    bool ok = xyzz_to_aff<F>(r, acc);
    st(out, r.x);
    st(out + 32, r.y);
    out[64] = ok;
}

extern "C" void hst_walk_chain(int field, uint8_t *out, const uint8_t *pts, const uint8_t *neg, size_t n)
{
    g_field = field;
    if (field == 0)
        walk_chain<Fp>(out, pts, neg, n);
    else
        walk_chain<Fq>(out, pts, neg, n);
}

extern "C" void hst_madd_chain(int field, uint8_t *out, const uint8_t *pts, const uint8_t *neg, size_t n)
{
    g_field = field;
    if (field == 0)
        madd_chain<Fp>(out, pts, neg, n);
    else
        madd_chain<Fq>(out, pts, neg, n);
}

extern "C" void hst_pt_op(int field, int o, uint8_t *out, const uint8_t *a, const uint8_t *b, uint32_t k)
{
    g_field = field;
    if (field == 0)
        pop<Fp>(o, out, a, b, k);
    else
        pop<Fq>(o, out, a, b, k);
}

#include "../../helioselene_amd/csrc/encode.h"


// compressed -> affine (x || y); returns 0 ok, 1 identity, -1 invalid
extern "C" int hst_decompress(int curve, uint8_t *out, const uint8_t *in)
{
    g_field = curve;
    uint32_t w[8];
    ldw(w, in);
    aff a;
    int rc = curve == 0 ? decompress<Fp, HeliosK>(a, w) : decompress<Fq, SeleneK>(a, w);
    st(out, a.x);
    st(out + 32, a.y);
    return rc;
}

extern "C" void hst_sswu(int curve, uint8_t *out, const uint8_t *u)
{
    g_field = curve;
    uint32_t w[8];
    ldw(w, u);
    aff a;
    if (curve == 0)
        sswu<Fp, HeliosK>(a, w);
    else
        sswu<Fq, SeleneK>(a, w);
    st(out, a.x);
    st(out + 32, a.y);
}

// k doublings of an affine point, then + q; exercises xyzz_dbl / xyzz_add with ZZ != 1
extern "C" void hst_dbl_chain(int field, uint8_t *out, const uint8_t *pa, const uint8_t *pb, int k)
{
    g_field = field;
    aff a, b, r;
    ld(a.x, pa); ld(a.y, pa + 32); ld(b.x, pb); ld(b.y, pb + 32);
    bool ok;
    if (field == 0)
    {
        xyzz A, B; xyzz_from_aff<Fp>(A, a); xyzz_from_aff<Fp>(B, b);
        for (int i = 0; i < k; i++) xyzz_dbl<Fp>(A, A);
        xyzz_dbl<Fp>(B, B);
        xyzz_add<Fp>(A, A, B);
        ok = xyzz_to_aff<Fp>(r, A);
    }
    else
    {
        xyzz A, B; xyzz_from_aff<Fq>(A, a); xyzz_from_aff<Fq>(B, b);
        for (int i = 0; i < k; i++) xyzz_dbl<Fq>(A, A);
        xyzz_dbl<Fq>(B, B);
        xyzz_add<Fq>(A, A, B);
        ok = xyzz_to_aff<Fq>(r, A);
    }
    st(out, r.x); st(out + 32, r.y); out[64] = ok;
}

#include "../../helioselene_amd/csrc/scalar.h"

// prepared (reduced, optionally window-8 adjusted) scalar and its width-c digits (int32, signed)
extern "C" int hst_recode(int curve, int c, int compat, const uint8_t *s32, uint8_t *prepared, int32_t *digits)
{
    uint32_t w[8];
    ldw(w, s32);
    uint32_t k[8];
    if (curve == 0)
        scalar_prepare<Fq>(k, w, compat != 0);
    else
        scalar_prepare<Fp>(k, w, compat != 0);
    stw(prepared, k);
    const int W = (256 + c - 1) / c;
    uint32_t carry = 0;
    for (int i = 0; i < W; i++)
    {
        uint32_t d = recode_digit(k, c, i, carry);
        digits[i] = (d >> 31) ? -(int32_t)(d & 0x7fffffffu) : (int32_t)d;
        if ((d >> 31) && (d & 0x7fffffffu) == 0)
            return -1; // a signed zero must never be produced
    }
    return W;
}

// Jacobian chain: to_jac(xyzz of a) doubled k times, plus to_jac(xyzz(b) doubled once)
extern "C" void hst_jac_chain(int field, uint8_t *out, const uint8_t *pa, const uint8_t *pb, int k)
{
    g_field = field;
    aff a, b, r;
    ld(a.x, pa); ld(a.y, pa + 32); ld(b.x, pb); ld(b.y, pb + 32);
    bool ok;
    if (field == 0)
    {
        xyzz A, B; jac JA, JB; xyzz_from_aff<Fp>(A, a); xyzz_from_aff<Fp>(B, b);
        xyzz_dbl<Fp>(A, A); xyzz_dbl<Fp>(B, B); // non-trivial ZZ, ZZZ
        xyzz_to_jac<Fp>(JA, A); xyzz_to_jac<Fp>(JB, B);
        for (int i = 0; i < k; i++) jac_dbl<Fp>(JA, JA);
        jac_add<Fp>(JA, JA, JB);
        ok = jac_to_aff<Fp>(r, JA);
    }
    else
    {
        xyzz A, B; jac JA, JB; xyzz_from_aff<Fq>(A, a); xyzz_from_aff<Fq>(B, b);
        xyzz_dbl<Fq>(A, A); xyzz_dbl<Fq>(B, B);
        xyzz_to_jac<Fq>(JA, A); xyzz_to_jac<Fq>(JB, B);
        for (int i = 0; i < k; i++) jac_dbl<Fq>(JA, JA);
        jac_add<Fq>(JA, JA, JB);
        ok = jac_to_aff<Fq>(r, JA);
    }
    st(out, r.x); st(out + 32, r.y); out[64] = ok;
}
