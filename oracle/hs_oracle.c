/*
 * hs_oracle.c -- TEST INFRASTRUCTURE (the checker).  A plain-C restatement of the reference's
 * variable-time MSM path for Helios and Selene.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product library never does.
 *
 * What each part restates (reference file:line):
 *   field F_p / F_q   fp/include/fp_ops.h:42-91, fp/include/x64/fp51_inline.h:46-243 (mod 2^255-19);
 *                     fq/include/fq_ops.h:50-101, fq/include/x64/fq51.h:37-86, fq51_inline.h:1086-1225
 *                     (mod q = 2^255-gamma).  Values are kept canonical in 4x64-bit limbs; the
 *                     reference's lazy radix-2^51 limbs denote the same residues.
 *   inversion         fp/src/x64/fp_invert.cpp:32-67 (Fermat a^(p-2)); fq_invert.cpp:41-72 (divsteps);
 *                     both are "a^-1, 0 -> 0", restated here as a^(m-2).
 *   sqrt              fp/src/x64/fp_sqrt.cpp:67+ (Atkin, p = 5 mod 8: beta = z^((p+3)/8), times
 *                     sqrt(-1) when beta^2 = -z); fq/src/x64/fq_sqrt.cpp:45 (z^((q+1)/4), q = 3 mod 4).
 *   points            helios/src/x64/helios_dbl.cpp:32-88 (dbl-2001-b, a=-3),
 *                     helios_add.cpp:32-107 (add-2007-bl), helios_madd.cpp:32-106 (madd-2007-bl),
 *                     helios/include/helios_add.h:44-87 and helios_msm_vartime.cpp:59-104 (edge cases),
 *                     helios/include/helios_ops.h:44-124; Selene twins in selene/.
 *   encoding          helios_tobytes.cpp:42-66, helios_frombytes.cpp:59-109, selene_frombytes.cpp:52-111.
 *   MSM               helios/src/x64/helios_msm_vartime.cpp:110-416 (Straus n<=16 with signed
 *                     4-bit digits; Pippenger with window 5..11, signed w-bit digits whose final
 *                     carry is dropped, 2^(w-1) buckets, running-sum reduction, Horner combine).
 *   Pedersen          helios/include/helios_pedersen.h:55-86.
 *   fixed-base MSM    helios/include/helios_msm_fixed.h:60-155, helios_scalarmult_fixed.h:54-89,
 *                     helios_batch_affine.h:55-98, fp/include/fp_batch_invert.h:61-114.
 *   SSWU              helios/src/x64/helios_map_to_curve.cpp:55-198 (Z = 7),
 *                     selene/src/x64/selene_map_to_curve.cpp:59-73 (Z = -4): RFC 9380 6.6.2.
 *   reduce_wide       helios/include/helios_scalar.h:113-152: int(64 bytes LE) mod order.
 *   PRNG              src/tests_fuzz.cpp:103-157 (xoshiro256** seeded by splitmix64).
 */
#include "hs_oracle.h"

#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;
typedef struct
{
    uint64_t v[4];
} fe;

typedef struct
{
    fe m;          /* modulus 2^255 - c */
    uint64_t c[2]; /* c (19 or gamma) */
    uint64_t c2[2];
    fe b;     /* curve constant b */
    fe gx, gy; /* generator */
    fe order;  /* group order of the curve over this field */
    fe sswu_z; /* SSWU Z */
} field_t;

static field_t FLD[2];
static int g_init = 0;

static void load32(fe *r, const uint8_t *b)
{
    for (int i = 0; i < 4; i++)
    {
        uint64_t w = 0;
        for (int j = 7; j >= 0; j--)
            w = (w << 8) | b[8 * i + j];
        r->v[i] = w;
    }
}

static void store32(uint8_t *b, const fe *a)
{
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 8; j++)
            b[8 * i + j] = (uint8_t)(a->v[i] >> (8 * j));
}

static int fe_cmp(const fe *a, const fe *b)
{
    for (int i = 3; i >= 0; i--)
    {
        if (a->v[i] < b->v[i])
            return -1;
        if (a->v[i] > b->v[i])
            return 1;
    }
    return 0;
}

static int fe_iszero(const fe *a) { return (a->v[0] | a->v[1] | a->v[2] | a->v[3]) == 0; }

static uint64_t add4(fe *r, const fe *a, const fe *b)
{
    u128 c = 0;
    for (int i = 0; i < 4; i++)
    {
        c += (u128)a->v[i] + b->v[i];
        r->v[i] = (uint64_t)c;
        c >>= 64;
    }
    return (uint64_t)c;
}

static uint64_t sub4(fe *r, const fe *a, const fe *b)
{
    uint64_t borrow = 0;
    for (int i = 0; i < 4; i++)
    {
        u128 d = (u128)a->v[i] - b->v[i] - borrow;
        r->v[i] = (uint64_t)d;
        borrow = (uint64_t)(d >> 127);
    }
    return borrow;
}

/* Reduce an nl-limb value (nl <= 8) modulo m = 2^255 - c, using 2^256 = 2c (mod m). */
static void fe_reduce(const field_t *F, fe *r, const uint64_t *t, int nl)
{
    uint64_t x[12];
    memset(x, 0, sizeof x);
    memcpy(x, t, (size_t)nl * 8);
    int n = nl;
    while (n > 4 && x[n - 1] == 0)
        n--;
    while (n > 4)
    {
        uint64_t y[12];
        memset(y, 0, sizeof y);
        memcpy(y, x, 32);
        int m = n - 4;
        for (int i = 0; i < m; i++)
        {
            u128 carry = 0;
            for (int j = 0; j < 2; j++)
            {
                u128 v = (u128)x[4 + i] * F->c2[j] + y[i + j] + carry;
                y[i + j] = (uint64_t)v;
                carry = v >> 64;
            }
            for (int k = i + 2; carry && k < 12; k++)
            {
                u128 v = (u128)y[k] + carry;
                y[k] = (uint64_t)v;
                carry = v >> 64;
            }
        }
        memcpy(x, y, sizeof y);
        n = 12;
        while (n > 4 && x[n - 1] == 0)
            n--;
    }
    fe a = {{x[0], x[1], x[2], x[3]}};
    while (a.v[3] >> 63)
    {
        a.v[3] &= 0x7fffffffffffffffULL;
        fe cc = {{F->c[0], F->c[1], 0, 0}};
        add4(&a, &a, &cc);
    }
    if (fe_cmp(&a, &F->m) >= 0)
        sub4(&a, &a, &F->m);
    *r = a;
}

static void fe_from_bytes_reduced(const field_t *F, fe *r, const uint8_t b[32])
{
    fe t;
    load32(&t, b);
    fe_reduce(F, r, t.v, 4);
}

static void fe_add(const field_t *F, fe *r, const fe *a, const fe *b)
{
    fe t;
    add4(&t, a, b); /* a, b < 2^255 so no carry out */
    if (fe_cmp(&t, &F->m) >= 0)
        sub4(&t, &t, &F->m);
    *r = t;
}

static void fe_sub(const field_t *F, fe *r, const fe *a, const fe *b)
{
    fe t;
    if (sub4(&t, a, b))
        add4(&t, &t, &F->m);
    *r = t;
}

static void fe_neg(const field_t *F, fe *r, const fe *a)
{
    fe z = {{0, 0, 0, 0}};
    fe_sub(F, r, &z, a);
}

static void fe_mul(const field_t *F, fe *r, const fe *a, const fe *b)
{
    uint64_t t[8];
    memset(t, 0, sizeof t);
    for (int i = 0; i < 4; i++)
    {
        u128 carry = 0;
        for (int j = 0; j < 4; j++)
        {
            u128 v = (u128)a->v[i] * b->v[j] + t[i + j] + carry;
            t[i + j] = (uint64_t)v;
            carry = v >> 64;
        }
        t[i + 4] = (uint64_t)carry;
    }
    fe_reduce(F, r, t, 8);
}

static void fe_sq(const field_t *F, fe *r, const fe *a) { fe_mul(F, r, a, a); }

static void fe_small(fe *r, uint64_t v)
{
    r->v[0] = v;
    r->v[1] = r->v[2] = r->v[3] = 0;
}

static void fe_pow(const field_t *F, fe *r, const fe *a, const fe *e)
{
    fe acc;
    fe_small(&acc, 1);
    for (int i = 255; i >= 0; i--)
    {
        fe_sq(F, &acc, &acc);
        if ((e->v[i / 64] >> (i % 64)) & 1)
            fe_mul(F, &acc, &acc, a);
    }
    *r = acc;
}

static void fe_inv(const field_t *F, fe *r, const fe *a)
{
    fe e = F->m, two;
    fe_small(&two, 2);
    sub4(&e, &e, &two);
    fe_pow(F, r, a, &e); /* 0 -> 0 */
}

static int fe_isodd(const fe *a) { return (int)(a->v[0] & 1); }

static int fe_eq(const fe *a, const fe *b) { return fe_cmp(a, b) == 0; }

/* sqrt: returns 0 and a root, or -1 and 0 when a is a non-residue. */
static int fe_sqrt(int k, fe *r, const fe *z)
{
    const field_t *F = &FLD[k];
    fe e, one, t;
    fe_small(&one, 1);
    if (k == 0)
    {
        /* p = 5 mod 8: beta = z^((p+3)/8) */
        e = F->m;
        fe three;
        fe_small(&three, 3);
        add4(&e, &e, &three);
        for (int i = 0; i < 3; i++)
        {
            e.v[0] = (e.v[0] >> 1) | (e.v[1] << 63);
            e.v[1] = (e.v[1] >> 1) | (e.v[2] << 63);
            e.v[2] = (e.v[2] >> 1) | (e.v[3] << 63);
            e.v[3] >>= 1;
        }
        fe beta, bsq, nz;
        fe_pow(F, &beta, z, &e);
        fe_sq(F, &bsq, &beta);
        fe_neg(F, &nz, z);
        if (fe_eq(&bsq, z))
        {
            *r = beta;
            return 0;
        }
        if (fe_eq(&bsq, &nz))
        {
            /* sqrt(-1) = 2^((p-1)/4) */
            fe two, em;
            fe_small(&two, 2);
            em = F->m;
            sub4(&em, &em, &one);
            for (int i = 0; i < 2; i++)
            {
                em.v[0] = (em.v[0] >> 1) | (em.v[1] << 63);
                em.v[1] = (em.v[1] >> 1) | (em.v[2] << 63);
                em.v[2] = (em.v[2] >> 1) | (em.v[3] << 63);
                em.v[3] >>= 1;
            }
            fe_pow(F, &t, &two, &em);
            fe_mul(F, r, &beta, &t);
            return 0;
        }
        fe_small(r, 0);
        return -1;
    }
    /* q = 3 mod 4: y = z^((q+1)/4) */
    e = F->m;
    add4(&e, &e, &one);
    for (int i = 0; i < 2; i++)
    {
        e.v[0] = (e.v[0] >> 1) | (e.v[1] << 63);
        e.v[1] = (e.v[1] >> 1) | (e.v[2] << 63);
        e.v[2] = (e.v[2] >> 1) | (e.v[3] << 63);
        e.v[3] >>= 1;
    }
    fe y, ysq;
    fe_pow(F, &y, z, &e);
    fe_sq(F, &ysq, &y);
    *r = y;
    return fe_eq(&ysq, z) ? 0 : -1;
}

static void hexle(fe *r, const char *hex)
{
    /* hex: 64 chars, little-endian byte order (as in the reference test-vector JSON) */
    uint8_t b[32];
    for (int i = 0; i < 32; i++)
    {
        unsigned v = 0;
        for (int j = 0; j < 2; j++)
        {
            char ch = hex[2 * i + j];
            v = v * 16 + (unsigned)((ch >= '0' && ch <= '9') ? ch - '0' : (ch | 32) - 'a' + 10);
        }
        b[i] = (uint8_t)v;
    }
    load32(r, b);
}

static void init_fields(void)
{
    if (g_init)
        return;
    /* F_p: p = 2^255 - 19 */
    field_t *P = &FLD[0];
    P->m.v[0] = 0xffffffffffffffedULL;
    P->m.v[1] = P->m.v[2] = 0xffffffffffffffffULL;
    P->m.v[3] = 0x7fffffffffffffffULL;
    P->c[0] = 19;
    P->c[1] = 0;
    /* F_q: q = Helios order (helios_constants.h:57-64) */
    field_t *Q = &FLD[1];
    hexle(&Q->m, "9fc7277972d2b66e586b65b72c787fbfffffffffffffffffffffffffffffff7f");
    /* gamma = 2^255 - q, which is < 2^128 */
    {
        fe two255 = {{0, 0, 0, 0x8000000000000000ULL}}, g;
        sub4(&g, &two255, &Q->m);
        Q->c[0] = g.v[0];
        Q->c[1] = g.v[1];
    }
    for (int k = 0; k < 2; k++)
    {
        field_t *F = &FLD[k];
        F->c2[0] = F->c[0] << 1;
        F->c2[1] = (F->c[1] << 1) | (F->c[0] >> 63);
    }
    /* curve constants (test_vectors JSON "parameters"; helios_constants.h, selene_constants.h) */
    hexle(&P->b, "d43ad7ede19eb42235bf1341386f3f043b7bbb3e6ba794beb870eab039c7e822");
    hexle(&P->gx, "0300000000000000000000000000000000000000000000000000000000000000");
    hexle(&P->gy, "f43e18e339e643d2dca586c5dd3b0059075f2050836692bd1c72c07ad9747b53");
    P->order = Q->m;
    fe_small(&P->sswu_z, 7);
    hexle(&Q->b, "58455acb8309de680e90ae06df4b94f379e2ff95a5bb517fc176586913771270");
    hexle(&Q->gx, "0100000000000000000000000000000000000000000000000000000000000000");
    hexle(&Q->gy, "d2fdd3a1a60a1e74379bf0c898b18b935f825c457731c95792ca5cb827d9197a");
    Q->order = P->m;
    {
        fe four;
        fe_small(&four, 4);
        g_init = 1; /* fe_neg only needs m */
        fe_neg(Q, &Q->sswu_z, &four);
    }
    g_init = 1;
}

/* ------------------------------------------------------------------ points (Jacobian) */
typedef struct
{
    fe X, Y, Z;
} jac;

typedef struct
{
    fe x, y;
} aff;

static void pt_identity(jac *r)
{
    fe_small(&r->X, 1);
    fe_small(&r->Y, 1);
    fe_small(&r->Z, 0);
}

static int pt_isid(const jac *p) { return fe_iszero(&p->Z); }

/* dbl-2001-b, a = -3 (helios_dbl.cpp:32-88) */
static void pt_dbl(const field_t *F, jac *r, const jac *p)
{
    fe delta, gamma, beta, alpha, t0, t1, t2, X3, Y3, Z3;
    fe_sq(F, &delta, &p->Z);
    fe_sq(F, &gamma, &p->Y);
    fe_mul(F, &beta, &p->X, &gamma);
    fe_sub(F, &t0, &p->X, &delta);
    fe_add(F, &t1, &p->X, &delta);
    fe_mul(F, &alpha, &t0, &t1);
    fe_add(F, &t0, &alpha, &alpha);
    fe_add(F, &alpha, &t0, &alpha);
    fe_sq(F, &X3, &alpha);
    fe_add(F, &t0, &beta, &beta);
    fe_add(F, &t0, &t0, &t0);
    fe_sub(F, &X3, &X3, &t0);
    fe_sub(F, &X3, &X3, &t0);
    fe_add(F, &t1, &p->Y, &p->Z);
    fe_sq(F, &t2, &t1);
    fe_sub(F, &t2, &t2, &gamma);
    fe_sub(F, &Z3, &t2, &delta);
    fe_sub(F, &t1, &t0, &X3);
    fe_mul(F, &t2, &alpha, &t1);
    fe_sq(F, &t0, &gamma);
    fe_add(F, &t0, &t0, &t0);
    fe_add(F, &t0, &t0, &t0);
    fe_sub(F, &Y3, &t2, &t0);
    fe_sub(F, &Y3, &Y3, &t0);
    r->X = X3;
    r->Y = Y3;
    r->Z = Z3;
}

/* add-2007-bl raw (helios_add.cpp:32-107) */
static void pt_add_raw(const field_t *F, jac *r, const jac *p, const jac *q)
{
    fe Z1Z1, Z2Z2, U1, U2, S1, S2, H, I, J, rr, V, t0, t1, X3, Y3, Z3;
    fe_sq(F, &Z1Z1, &p->Z);
    fe_sq(F, &Z2Z2, &q->Z);
    fe_mul(F, &U1, &p->X, &Z2Z2);
    fe_mul(F, &U2, &q->X, &Z1Z1);
    fe_mul(F, &t0, &q->Z, &Z2Z2);
    fe_mul(F, &S1, &p->Y, &t0);
    fe_mul(F, &t0, &p->Z, &Z1Z1);
    fe_mul(F, &S2, &q->Y, &t0);
    fe_sub(F, &H, &U2, &U1);
    fe_add(F, &t0, &H, &H);
    fe_sq(F, &I, &t0);
    fe_mul(F, &J, &H, &I);
    fe_sub(F, &rr, &S2, &S1);
    fe_add(F, &rr, &rr, &rr);
    fe_mul(F, &V, &U1, &I);
    fe_sq(F, &X3, &rr);
    fe_sub(F, &X3, &X3, &J);
    fe_add(F, &t0, &V, &V);
    fe_sub(F, &X3, &X3, &t0);
    fe_sub(F, &t0, &V, &X3);
    fe_mul(F, &t1, &rr, &t0);
    fe_mul(F, &t0, &S1, &J);
    fe_add(F, &t0, &t0, &t0);
    fe_sub(F, &Y3, &t1, &t0);
    fe_add(F, &t0, &p->Z, &q->Z);
    fe_sq(F, &t1, &t0);
    fe_sub(F, &t1, &t1, &Z1Z1);
    fe_sub(F, &t1, &t1, &Z2Z2);
    fe_mul(F, &Z3, &t1, &H);
    r->X = X3;
    r->Y = Y3;
    r->Z = Z3;
}

/* madd-2007-bl raw, no edge cases (helios_madd.cpp:32-106) */
static void pt_madd_raw(const field_t *F, jac *r, const jac *p, const aff *q)
{
    fe Z1Z1, U2, S2, H, HH, I, J, rr, V, t0, t1, X3, Y3, Z3;
    fe_sq(F, &Z1Z1, &p->Z);
    fe_mul(F, &U2, &q->x, &Z1Z1);
    fe_mul(F, &t0, &p->Z, &Z1Z1);
    fe_mul(F, &S2, &q->y, &t0);
    fe_sub(F, &H, &U2, &p->X);
    fe_sq(F, &HH, &H);
    fe_add(F, &I, &HH, &HH);
    fe_add(F, &I, &I, &I);
    fe_mul(F, &J, &H, &I);
    fe_sub(F, &rr, &S2, &p->Y);
    fe_add(F, &rr, &rr, &rr);
    fe_mul(F, &V, &p->X, &I);
    fe_sq(F, &X3, &rr);
    fe_sub(F, &X3, &X3, &J);
    fe_add(F, &t0, &V, &V);
    fe_sub(F, &X3, &X3, &t0);
    fe_sub(F, &t0, &V, &X3);
    fe_mul(F, &t1, &rr, &t0);
    fe_mul(F, &t0, &p->Y, &J);
    fe_add(F, &t0, &t0, &t0);
    fe_sub(F, &Y3, &t1, &t0);
    fe_add(F, &t0, &p->Z, &H);
    fe_sq(F, &t1, &t0);
    fe_sub(F, &t1, &t1, &Z1Z1);
    fe_sub(F, &Z3, &t1, &HH);
    r->X = X3;
    r->Y = Y3;
    r->Z = Z3;
}

/* add with identity / doubling / inverse handling (helios_msm_vartime.cpp:59-104,
 * helios/include/helios_add.h:44-87) */
static void pt_add(const field_t *F, jac *r, const jac *p, const jac *q)
{
    if (pt_isid(p))
    {
        *r = *q;
        return;
    }
    if (pt_isid(q))
    {
        *r = *p;
        return;
    }
    fe z1z1, z2z2, u1, u2;
    fe_sq(F, &z1z1, &p->Z);
    fe_sq(F, &z2z2, &q->Z);
    fe_mul(F, &u1, &p->X, &z2z2);
    fe_mul(F, &u2, &q->X, &z1z1);
    if (fe_eq(&u1, &u2))
    {
        fe s1, s2, t;
        fe_mul(F, &t, &q->Z, &z2z2);
        fe_mul(F, &s1, &p->Y, &t);
        fe_mul(F, &t, &p->Z, &z1z1);
        fe_mul(F, &s2, &q->Y, &t);
        if (fe_eq(&s1, &s2))
            pt_dbl(F, r, p);
        else
            pt_identity(r);
        return;
    }
    pt_add_raw(F, r, p, q);
}

static void pt_neg(const field_t *F, jac *r, const jac *p)
{
    r->X = p->X;
    fe_neg(F, &r->Y, &p->Y);
    r->Z = p->Z;
}

static void pt_to_affine(const field_t *F, aff *r, const jac *p)
{
    fe zi, zi2, zi3;
    fe_inv(F, &zi, &p->Z);
    fe_sq(F, &zi2, &zi);
    fe_mul(F, &zi3, &zi2, &zi);
    fe_mul(F, &r->x, &p->X, &zi2);
    fe_mul(F, &r->y, &p->Y, &zi3);
}

static void pt_tobytes(const field_t *F, uint8_t out[32], const jac *p)
{
    if (pt_isid(p))
    {
        memset(out, 0, 32);
        return;
    }
    aff a;
    pt_to_affine(F, &a, p);
    store32(out, &a.x);
    out[31] |= (uint8_t)(fe_isodd(&a.y) << 7);
}

/* helios_frombytes.cpp:59-109 / selene_frombytes.cpp:52-111; 0^32 -> identity (test loader). */
static int pt_frombytes(int k, jac *r, const uint8_t in[32])
{
    const field_t *F = &FLD[k];
    int allz = 1;
    for (int i = 0; i < 32; i++)
        allz &= in[i] == 0;
    if (allz)
    {
        pt_identity(r);
        return 0;
    }
    unsigned parity = in[31] >> 7;
    uint8_t xb[32];
    memcpy(xb, in, 32);
    xb[31] &= 0x7f;
    fe x;
    load32(&x, xb);
    if (fe_cmp(&x, &F->m) >= 0)
        return -1;
    fe x2, x3, rhs, three_x, y;
    fe_sq(F, &x2, &x);
    fe_mul(F, &x3, &x2, &x);
    fe_add(F, &three_x, &x, &x);
    fe_add(F, &three_x, &three_x, &x);
    fe_sub(F, &rhs, &x3, &three_x);
    fe_add(F, &rhs, &rhs, &F->b);
    if (fe_sqrt(k, &y, &rhs) != 0)
        return -1;
    if ((unsigned)fe_isodd(&y) != parity)
        fe_neg(F, &y, &y);
    r->X = x;
    r->Y = y;
    fe_small(&r->Z, 1);
    return 0;
}

/* Exact scalar multiplication (double-and-add); s is any 256-bit integer. */
static void pt_smul(const field_t *F, jac *r, const uint8_t s[32], const jac *p)
{
    jac acc;
    pt_identity(&acc);
    for (int i = 255; i >= 0; i--)
    {
        pt_dbl(F, &acc, &acc);
        if ((s[i / 8] >> (i % 8)) & 1)
            pt_add(F, &acc, &acc, p);
    }
    *r = acc;
}

/* ------------------------------------------------------------------ recoding */
/* encode_signed_w4 (helios_msm_vartime.cpp:110-125) */
static void encode_w4(int16_t d[64], const uint8_t *s)
{
    int carry = 0;
    for (int i = 0; i < 31; i++)
    {
        carry += s[i];
        int carry2 = (carry + 8) >> 4;
        d[2 * i] = (int16_t)(carry - (carry2 << 4));
        carry = (carry2 + 8) >> 4;
        d[2 * i + 1] = (int16_t)(carry2 - (carry << 4));
    }
    carry += s[31];
    int carry2 = (carry + 8) >> 4;
    d[62] = (int16_t)(carry - (carry2 << 4));
    d[63] = (int16_t)carry2;
}

/* encode_signed_wbit (helios_msm_vartime.cpp:127-162): ceil(256/w) digits in [-2^(w-1), 2^(w-1)),
 * the carry out of the last digit is not kept.  Returns the digit count; *dropped gets that carry. */
static int encode_wbit(int16_t *d, const uint8_t *s, int w, int *dropped)
{
    const int half = 1 << (w - 1), mask = (1 << w) - 1, nd = (256 + w - 1) / w;
    int carry = 0;
    for (int i = 0; i < nd; i++)
    {
        int bit = i * w, byte = bit / 8, off = bit % 8;
        int raw = 0;
        if (byte < 32)
            raw = s[byte] >> off;
        if (byte + 1 < 32 && off + w > 8)
            raw |= (int)s[byte + 1] << (8 - off);
        if (byte + 2 < 32 && off + w > 16)
            raw |= (int)s[byte + 2] << (16 - off);
        int val = (raw & mask) + carry;
        carry = val >> w;
        val &= mask;
        if (val >= half)
        {
            val -= 1 << w;
            carry = 1;
        }
        d[i] = (int16_t)val;
    }
    if (dropped)
        *dropped = carry;
    return nd;
}

int hso_recode(int16_t *digits, const uint8_t s[32], int w) { return encode_wbit(digits, s, w, NULL); }

int hso_recode_drops_carry(const uint8_t s[32], int w)
{
    int16_t d[64];
    int dropped = 0;
    encode_wbit(d, s, w, &dropped);
    return dropped;
}

/* pippenger_window_size (helios_msm_vartime.cpp:248-263) */
int hso_ref_window(size_t n)
{
    if (n <= 16)
        return 0;
    if (n < 96)
        return 5;
    if (n < 288)
        return 6;
    if (n < 864)
        return 7;
    if (n < 2592)
        return 8;
    if (n < 7776)
        return 9;
    if (n < 23328)
        return 10;
    return 11;
}

/* ------------------------------------------------------------------ MSM (reference algorithm) */
static int msm_straus(const field_t *F, jac *res, const uint8_t *scalars, const jac *pts, size_t n)
{
    int16_t *dg = malloc(n * 64 * sizeof(int16_t));
    jac *tab = malloc(n * 8 * sizeof(jac));
    if (!dg || !tab)
    {
        free(dg);
        free(tab);
        return -1;
    }
    int rc = 0;
    for (size_t i = 0; i < n; i++)
    {
        encode_w4(dg + i * 64, scalars + 32 * i);
        jac *T = tab + i * 8;
        T[0] = pts[i];
        pt_dbl(F, &T[1], &pts[i]);
        for (int j = 1; j < 7; j++)
            pt_add(F, &T[j + 1], &T[j], &pts[i]);
    }
    jac acc;
    pt_identity(&acc);
    for (int d = 63; d >= 0; d--)
    {
        for (int k = 0; k < 4; k++)
            pt_dbl(F, &acc, &acc);
        for (size_t i = 0; i < n; i++)
        {
            int dv = dg[i * 64 + d];
            if (dv == 0)
                continue;
            int a = dv > 0 ? dv : -dv;
            if (a > 8)
            {
                rc = -2; /* the reference reads past its 8-entry table here (scalar >= ~0xF8*2^248) */
                continue;
            }
            jac pt = tab[i * 8 + (a - 1)];
            if (dv < 0)
                pt_neg(F, &pt, &pt);
            pt_add(F, &acc, &acc, &pt);
        }
    }
    *res = acc;
    free(dg);
    free(tab);
    return rc;
}

static int msm_pippenger(const field_t *F, jac *res, const uint8_t *scalars, const jac *pts, size_t n)
{
    const int w = hso_ref_window(n);
    const size_t nb = (size_t)1 << (w - 1);
    const size_t nw = (size_t)((256 + w - 1) / w);
    int16_t *dg = malloc(n * nw * sizeof(int16_t));
    jac *bk = malloc(nb * sizeof(jac));
    if (!dg || !bk)
    {
        free(dg);
        free(bk);
        return -1;
    }
    for (size_t i = 0; i < n; i++)
        encode_wbit(dg + i * nw, scalars + 32 * i, w, NULL);
    jac total;
    pt_identity(&total);
    for (size_t win = nw; win-- > 0;)
    {
        for (int d = 0; d < w; d++)
            pt_dbl(F, &total, &total);
        for (size_t b = 0; b < nb; b++)
            pt_identity(&bk[b]);
        for (size_t i = 0; i < n; i++)
        {
            int dv = dg[i * nw + win];
            if (dv == 0)
                continue;
            jac e = pts[i];
            size_t b = (size_t)((dv > 0 ? dv : -dv) - 1);
            if (dv < 0)
                pt_neg(F, &e, &e);
            pt_add(F, &bk[b], &bk[b], &e);
        }
        jac running, partial;
        pt_identity(&running);
        pt_identity(&partial);
        for (size_t j = nb; j-- > 0;)
        {
            pt_add(F, &running, &running, &bk[j]);
            pt_add(F, &partial, &partial, &running);
        }
        pt_add(F, &total, &total, &partial);
    }
    *res = total;
    free(dg);
    free(bk);
    return 0;
}

static int decode_all(int k, jac **out, const uint8_t *points, size_t n)
{
    *out = malloc((n ? n : 1) * sizeof(jac));
    if (!*out)
        return -1;
    for (size_t i = 0; i < n; i++)
        if (pt_frombytes(k, &(*out)[i], points + 32 * i) != 0)
        {
            free(*out);
            *out = NULL;
            return -1;
        }
    return 0;
}

int hso_msm(int c, uint8_t out[32], const uint8_t *scalars, const uint8_t *points, size_t n)
{
    init_fields();
    const field_t *F = &FLD[c];
    jac *p, r;
    if (decode_all(c, &p, points, n) != 0)
        return -1;
    int rc = 0;
    if (n == 0)
        pt_identity(&r);
    else if (n <= 16)
        rc = msm_straus(F, &r, scalars, p, n);
    else
        rc = msm_pippenger(F, &r, scalars, p, n);
    pt_tobytes(F, out, &r);
    free(p);
    return rc;
}

int hso_msm_naive(int c, uint8_t out[32], const uint8_t *scalars, const uint8_t *points, size_t n)
{
    init_fields();
    const field_t *F = &FLD[c];
    jac *p, acc, t;
    if (decode_all(c, &p, points, n) != 0)
        return -1;
    pt_identity(&acc);
    for (size_t i = 0; i < n; i++)
    {
        pt_smul(F, &t, scalars + 32 * i, &p[i]);
        pt_add(F, &acc, &acc, &t);
    }
    pt_tobytes(F, out, &acc);
    free(p);
    return 0;
}

int hso_pedersen(int c, uint8_t out[32], const uint8_t blinding[32], const uint8_t H[32], const uint8_t *values,
                 const uint8_t *gens, size_t n)
{
    uint8_t *s = malloc(32 * (n + 1)), *g = malloc(32 * (n + 1));
    if (!s || !g)
    {
        free(s);
        free(g);
        return -1;
    }
    memcpy(s, blinding, 32);
    memcpy(g, H, 32);
    if (n)
    {
        memcpy(s + 32, values, 32 * n);
        memcpy(g + 32, gens, 32 * n);
    }
    int rc = hso_msm(c, out, s, g, n + 1);
    free(s);
    free(g);
    return rc;
}

/* ------------------------------------------------------------------ fixed-base */
/* fp_batch_invert (fp/include/fp_batch_invert.h:61-114), including its zero handling */
static void batch_invert(const field_t *F, fe *out, const fe *in, size_t n)
{
    if (n == 0)
        return;
    if (n == 1)
    {
        fe_inv(F, &out[0], &in[0]);
        return;
    }
    fe *acc = malloc(n * sizeof(fe));
    acc[0] = in[0];
    for (size_t i = 1; i < n; i++)
    {
        if (!fe_iszero(&in[i]))
            fe_mul(F, &acc[i], &acc[i - 1], &in[i]);
        else
            acc[i] = acc[i - 1];
    }
    fe inv;
    fe_inv(F, &inv, &acc[n - 1]);
    for (size_t i = n - 1; i > 0; i--)
    {
        if (fe_iszero(&in[i]))
            fe_small(&out[i], 0);
        else
        {
            fe tmp = in[i];
            fe_mul(F, &out[i], &inv, &acc[i - 1]);
            fe_mul(F, &inv, &inv, &tmp);
        }
    }
    if (!fe_iszero(&in[0]))
        out[0] = inv;
    else
        fe_small(&out[0], 0);
    free(acc);
}

void hso_batch_invert(int k, uint8_t *out, const uint8_t *in, size_t n)
{
    init_fields();
    const field_t *F = &FLD[k];
    fe *a = malloc((n ? n : 1) * sizeof(fe)), *b = malloc((n ? n : 1) * sizeof(fe));
    for (size_t i = 0; i < n; i++)
        fe_from_bytes_reduced(F, &a[i], in + 32 * i);
    batch_invert(F, b, a, n);
    for (size_t i = 0; i < n; i++)
        store32(out + 32 * i, &b[i]);
    free(a);
    free(b);
}

/* helios_scalarmult_fixed_precompute + helios_batch_to_affine */
static void fixed_table(const field_t *F, aff T[16], const jac *p)
{
    jac j[16];
    fe z[16], zi[16];
    j[0] = *p;
    pt_dbl(F, &j[1], p);
    for (int i = 2; i < 16; i++)
        pt_add(F, &j[i], &j[i - 1], p);
    for (int i = 0; i < 16; i++)
        z[i] = j[i].Z;
    batch_invert(F, zi, z, 16);
    for (int i = 0; i < 16; i++)
    {
        fe zi2, zi3;
        fe_sq(F, &zi2, &zi[i]);
        fe_mul(F, &zi3, &zi2, &zi[i]);
        fe_mul(F, &T[i].x, &j[i].X, &zi2);
        fe_mul(F, &T[i].y, &j[i].Y, &zi3);
        if (fe_iszero(&j[i].Z))
        {
            fe_small(&T[i].x, 0);
            fe_small(&T[i].y, 0);
        }
    }
}

/* helios_scalar_recode_signed5 (helios_scalarmult_fixed.h:69-89) */
static void recode5(int8_t d[52], const uint8_t *s)
{
    int carry = 0;
    for (int i = 0; i < 51; i++)
    {
        int bo = 5 * i, bi = bo >> 3, bp = bo & 7;
        unsigned word = s[bi];
        if (bi + 1 < 32)
            word |= (unsigned)s[bi + 1] << 8;
        int val = (int)((word >> bp) & 0x1fu) + carry;
        carry = (val + 16) >> 5;
        d[i] = (int8_t)(val - (carry << 5));
    }
    d[51] = (int8_t)(((s[31] >> 7) & 1) + carry);
}

/* one step of helios_msm_fixed's inner loop: r += sign*T[|d|-1] via raw madd, with the
 * reference's identity handling (helios_msm_fixed.h:86-108, 124-147) */
static void fixed_step(const field_t *F, jac *r, const aff *T, int d)
{
    if (d == 0)
        return;
    int a = d < 0 ? -d : d;
    aff sel = T[a - 1];
    if (d < 0)
        fe_neg(F, &sel.y, &sel.y);
    if (!pt_isid(r))
    {
        jac t;
        pt_madd_raw(F, &t, r, &sel);
        *r = t;
    }
    else
    {
        r->X = sel.x;
        r->Y = sel.y;
        fe_small(&r->Z, 1);
    }
}

int hso_fixed_table(int c, uint8_t out[16 * 64], const uint8_t p[32])
{
    init_fields();
    const field_t *F = &FLD[c];
    jac P;
    if (pt_frombytes(c, &P, p) != 0)
        return -1;
    aff T[16];
    fixed_table(F, T, &P);
    for (int i = 0; i < 16; i++)
    {
        store32(out + 64 * i, &T[i].x);
        store32(out + 64 * i + 32, &T[i].y);
    }
    return 0;
}

int hso_msm_fixed(int c, uint8_t out[32], const uint8_t *scalars, const uint8_t *points, size_t n)
{
    init_fields();
    const field_t *F = &FLD[c];
    jac *p, r;
    if (decode_all(c, &p, points, n) != 0)
        return -1;
    pt_identity(&r);
    if (n == 0)
    {
        pt_tobytes(F, out, &r);
        free(p);
        return 0;
    }
    aff *T = malloc(n * 16 * sizeof(aff));
    int8_t *dg = malloc(n * 52);
    for (size_t j = 0; j < n; j++)
    {
        fixed_table(F, T + 16 * j, &p[j]);
        recode5(dg + 52 * j, scalars + 32 * j);
    }
    if (n == 1)
    {
        /* helios_scalarmult_fixed: same digit walk for one base */
        for (int i = 51; i >= 0; i--)
        {
            if (i != 51)
                for (int k = 0; k < 5; k++)
                    pt_dbl(F, &r, &r);
            fixed_step(F, &r, T, dg[i]);
        }
    }
    else
    {
        for (size_t j = 0; j < n; j++)
            fixed_step(F, &r, T + 16 * j, dg[j * 52 + 51]);
        for (int i = 50; i >= 0; i--)
        {
            for (int k = 0; k < 5; k++)
                pt_dbl(F, &r, &r);
            for (size_t j = 0; j < n; j++)
                fixed_step(F, &r, T + 16 * j, dg[j * 52 + i]);
        }
    }
    pt_tobytes(F, out, &r);
    free(T);
    free(dg);
    free(p);
    return 0;
}

/* ------------------------------------------------------------------ SSWU (RFC 9380 6.6.2) */
static void sswu(int k, jac *r, const fe *u)
{
    const field_t *F = &FLD[k];
    fe A, three, B = F->b, Z = F->sswu_z;
    fe_small(&three, 3);
    fe_neg(F, &A, &three);
    fe u2, Zu2, tv, den, x1, gx1, x2, gx2, y, t, one;
    fe_small(&one, 1);
    fe_sq(F, &u2, u);
    fe_mul(F, &Zu2, &Z, &u2);
    fe_sq(F, &tv, &Zu2);
    fe_add(F, &den, &tv, &Zu2); /* Z^2 u^4 + Z u^2 */
    if (fe_iszero(&den))
    {
        /* x1 = B / (Z A) */
        fe za, zai;
        fe_mul(F, &za, &Z, &A);
        fe_inv(F, &zai, &za);
        fe_mul(F, &x1, &B, &zai);
    }
    else
    {
        /* x1 = (-B/A) (1 + 1/den) */
        fe ai, nba, di;
        fe_inv(F, &ai, &A);
        fe_mul(F, &nba, &B, &ai);
        fe_neg(F, &nba, &nba);
        fe_inv(F, &di, &den);
        fe_add(F, &di, &di, &one);
        fe_mul(F, &x1, &nba, &di);
    }
    fe_sq(F, &t, &x1);
    fe_mul(F, &gx1, &t, &x1);
    fe_mul(F, &t, &A, &x1);
    fe_add(F, &gx1, &gx1, &t);
    fe_add(F, &gx1, &gx1, &B);
    fe_mul(F, &x2, &Zu2, &x1);
    fe_sq(F, &t, &x2);
    fe_mul(F, &gx2, &t, &x2);
    fe_mul(F, &t, &A, &x2);
    fe_add(F, &gx2, &gx2, &t);
    fe_add(F, &gx2, &gx2, &B);
    fe x;
    if (fe_sqrt(k, &y, &gx1) == 0)
        x = x1;
    else
    {
        fe_sqrt(k, &y, &gx2);
        x = x2;
    }
    if (fe_isodd(u) != fe_isodd(&y))
        fe_neg(F, &y, &y);
    r->X = x;
    r->Y = y;
    fe_small(&r->Z, 1);
}

void hso_map_to_curve(int c, uint8_t *out, const uint8_t *u, size_t n)
{
    init_fields();
    for (size_t i = 0; i < n; i++)
    {
        fe uf;
        uint8_t b[32];
        memcpy(b, u + 32 * i, 32);
        b[31] &= 0x7f; /* fp_frombytes masks bit 255 (fp_frombytes.cpp:47) */
        fe_from_bytes_reduced(&FLD[c], &uf, b);
        jac r;
        sswu(c, &r, &uf);
        pt_tobytes(&FLD[c], out + 32 * i, &r);
    }
}

void hso_map_to_curve2(int c, uint8_t out[32], const uint8_t u0[32], const uint8_t u1[32])
{
    init_fields();
    uint8_t a[32], b[32];
    hso_map_to_curve(c, a, u0, 1);
    hso_map_to_curve(c, b, u1, 1);
    hso_point_add(c, out, a, b);
}

/* ------------------------------------------------------------------ small exported helpers */
void hso_fe_mul(int k, uint8_t out[32], const uint8_t a[32], const uint8_t b[32])
{
    init_fields();
    fe x, y;
    fe_from_bytes_reduced(&FLD[k], &x, a);
    fe_from_bytes_reduced(&FLD[k], &y, b);
    fe_mul(&FLD[k], &x, &x, &y);
    store32(out, &x);
}

void hso_fe_sq(int k, uint8_t out[32], const uint8_t a[32]) { hso_fe_mul(k, out, a, a); }

void hso_fe_add(int k, uint8_t out[32], const uint8_t a[32], const uint8_t b[32])
{
    init_fields();
    fe x, y;
    fe_from_bytes_reduced(&FLD[k], &x, a);
    fe_from_bytes_reduced(&FLD[k], &y, b);
    fe_add(&FLD[k], &x, &x, &y);
    store32(out, &x);
}

void hso_fe_sub(int k, uint8_t out[32], const uint8_t a[32], const uint8_t b[32])
{
    init_fields();
    fe x, y;
    fe_from_bytes_reduced(&FLD[k], &x, a);
    fe_from_bytes_reduced(&FLD[k], &y, b);
    fe_sub(&FLD[k], &x, &x, &y);
    store32(out, &x);
}

void hso_fe_inv(int k, uint8_t out[32], const uint8_t a[32])
{
    init_fields();
    fe x;
    fe_from_bytes_reduced(&FLD[k], &x, a);
    fe_inv(&FLD[k], &x, &x);
    store32(out, &x);
}

int hso_fe_sqrt(int k, uint8_t out[32], const uint8_t a[32])
{
    init_fields();
    fe x, y;
    fe_from_bytes_reduced(&FLD[k], &x, a);
    int rc = fe_sqrt(k, &y, &x);
    store32(out, &y);
    return rc;
}

int hso_point_add(int c, uint8_t out[32], const uint8_t a[32], const uint8_t b[32])
{
    init_fields();
    jac p, q, r;
    if (pt_frombytes(c, &p, a) || pt_frombytes(c, &q, b))
        return -1;
    pt_add(&FLD[c], &r, &p, &q);
    pt_tobytes(&FLD[c], out, &r);
    return 0;
}

int hso_point_dbl(int c, uint8_t out[32], const uint8_t a[32])
{
    init_fields();
    jac p, r;
    if (pt_frombytes(c, &p, a))
        return -1;
    if (pt_isid(&p))
        r = p;
    else
        pt_dbl(&FLD[c], &r, &p);
    pt_tobytes(&FLD[c], out, &r);
    return 0;
}

int hso_point_neg(int c, uint8_t out[32], const uint8_t a[32])
{
    init_fields();
    jac p, r;
    if (pt_frombytes(c, &p, a))
        return -1;
    pt_neg(&FLD[c], &r, &p);
    pt_tobytes(&FLD[c], out, &r);
    return 0;
}

int hso_point_valid(int c, const uint8_t a[32])
{
    init_fields();
    jac p;
    return pt_frombytes(c, &p, a) == 0;
}

int hso_scalar_mul(int c, uint8_t out[32], const uint8_t s[32], const uint8_t pb[32])
{
    init_fields();
    jac p, r;
    if (pt_frombytes(c, &p, pb))
        return -1;
    pt_smul(&FLD[c], &r, s, &p);
    pt_tobytes(&FLD[c], out, &r);
    return 0;
}

void hso_generator(int c, uint8_t out[32])
{
    init_fields();
    jac g;
    g.X = FLD[c].gx;
    g.Y = FLD[c].gy;
    fe_small(&g.Z, 1);
    pt_tobytes(&FLD[c], out, &g);
}

void hso_reduce_wide(int c, uint8_t out[32], const uint8_t wide[64])
{
    /* int(wide, LE) mod order; the order of curve c is the modulus of the other field */
    init_fields();
    const field_t *O = &FLD[1 - c];
    uint64_t t[8];
    for (int i = 0; i < 8; i++)
    {
        uint64_t w = 0;
        for (int j = 7; j >= 0; j--)
            w = (w << 8) | wide[8 * i + j];
        t[i] = w;
    }
    fe r;
    fe_reduce(O, &r, t, 8);
    store32(out, &r);
}

int hso_decompress(int c, uint8_t out[64], const uint8_t p[32])
{
    init_fields();
    jac P;
    if (pt_frombytes(c, &P, p) != 0)
        return -1;
    if (pt_isid(&P))
    {
        memset(out, 0, 64);
        return 1;
    }
    store32(out, &P.X);
    store32(out + 32, &P.Y);
    return 0;
}

static uint64_t splitmix64(uint64_t *st)
{
    uint64_t z = (*st += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

static uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

void hso_xoshiro_fill(uint64_t seed, uint8_t *buf, size_t len)
{
    uint64_t s[4], sm = seed;
    for (int i = 0; i < 4; i++)
        s[i] = splitmix64(&sm);
    size_t i = 0;
    while (i < len)
    {
        uint64_t result = rotl64(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl64(s[3], 45);
        size_t k = len - i < 8 ? len - i : 8;
        memcpy(buf + i, &result, k);
        i += k;
    }
}
