// Short-Weierstrass point arithmetic (a = -3) for Helios (over Fp) and Selene (over Fq).
//
// Coordinates used on the device:
//   aff   (x, y)              bases after ingest; the identity is carried as a flag, not a point
//   xyzz  (X, Y, ZZ, ZZZ)     x = X/ZZ, y = Y/ZZZ, ZZ^3 = ZZZ^2; ZZ = 0 is the identity.  Buckets,
//                             running sums and window sums live here: the mixed add costs 8M+2S
//                             and needs no inversion.
//   The reference keeps Jacobian points throughout (helios/include/helios.h:41-46) with
//   add-2007-bl / dbl-2001-b / madd-2007-bl (helios/src/x64/helios_add.cpp:32-107,
//   helios_dbl.cpp:32-88, helios_madd.cpp:32-106) and the edge-case wrapper of
//   helios/include/helios_add.h:44-87.  The group law is the same, so every representation
//   yields the same affine point and the same 32-byte encoding (helios_tobytes.cpp:42-66).
//
// Formulas (EFD, g1p/shortw/xyzz): madd-2008-s, add-2008-s, dbl-2008-s-1, mdbl-2008-s-1.  The
// equal-x cases (P == Q -> doubling, P == -Q -> identity) are handled explicitly, as the
// reference's helios_add_safe does (helios_msm_vartime.cpp:59-104).
#pragma once
#include "field.h"

namespace hs
{

struct aff
{
    fe x, y;
};

struct xyzz
{
    fe X, Y, ZZ, ZZZ;
};

template <class F> HS_HD void xyzz_identity(xyzz &r)
{
    fe_one(r.X);
    fe_one(r.Y);
    fe_zero(r.ZZ);
    fe_zero(r.ZZZ);
}

template <class F> HS_HD bool xyzz_is_identity(const xyzz &p) { return fe_is_zero<F>(p.ZZ); }

template <class F> HS_HD void xyzz_from_aff(xyzz &r, const aff &a)
{
    r.X = a.x;
    r.Y = a.y;
    fe_one(r.ZZ);
    fe_one(r.ZZZ);
}

// 3*(x - z)(x + z): the a = -3 tangent numerator 3x^2 - 3z^2
template <class F> HS_HD void tangent_m(fe &m, const fe &x, const fe &z)
{
    fe t0, t1;
    fe_sub<F>(t0, x, z);
    fe_add<F>(t1, x, z);
    fe_mul<F>(m, t0, t1);
    fe_add<F>(t0, m, m);
    fe_add<F>(m, t0, m);
}

// dbl-2008-s-1 with a = -3 (M = 3X^2 - 3ZZ^2 = 3(X-ZZ)(X+ZZ)); 7M + 2S
template <class F> HS_HD void xyzz_dbl(xyzz &r, const xyzz &p)
{
    fe U, V, W, S, M, t;
    fe_add<F>(U, p.Y, p.Y);
    fe_sq<F>(V, U);
    fe_mul<F>(W, U, V);
    fe_mul<F>(S, p.X, V);
    tangent_m<F>(M, p.X, p.ZZ);
    fe X3, Y3;
    fe_sq<F>(X3, M);
    fe_sub<F>(X3, X3, S);
    fe_sub<F>(X3, X3, S);
    fe_sub<F>(t, S, X3);
    fe_mul<F>(Y3, M, t);
    fe_mul<F>(t, W, p.Y);
    fe_sub<F>(Y3, Y3, t);
    fe_mul<F>(r.ZZ, V, p.ZZ);
    fe_mul<F>(r.ZZZ, W, p.ZZZ);
    r.X = X3;
    r.Y = Y3;
}

// mdbl-2008-s-1 with a = -3: doubling of an affine point into xyzz
template <class F> HS_HD void xyzz_mdbl(xyzz &r, const aff &p)
{
    fe one, U, V, W, S, M, t;
    fe_one(one);
    fe_add<F>(U, p.y, p.y);
    fe_sq<F>(V, U);
    fe_mul<F>(W, U, V);
    fe_mul<F>(S, p.x, V);
    tangent_m<F>(M, p.x, one);
    fe X3, Y3;
    fe_sq<F>(X3, M);
    fe_sub<F>(X3, X3, S);
    fe_sub<F>(X3, X3, S);
    fe_sub<F>(t, S, X3);
    fe_mul<F>(Y3, M, t);
    fe_mul<F>(t, W, p.y);
    fe_sub<F>(Y3, Y3, t);
    r.X = X3;
    r.Y = Y3;
    r.ZZ = V;
    r.ZZZ = W;
}

// acc += q (affine), madd-2008-s, with identity / doubling / inverse handling.
// `acc_id` carries acc's identity state so the hot loop needs no zero test on ZZ.
template <class F> HS_HD void xyzz_madd(xyzz &acc, bool &acc_id, const aff &q)
{
    if (acc_id)
    {
        xyzz_from_aff<F>(acc, q);
        acc_id = false;
        return;
    }
    fe U2, S2, P, R;
    fe_mul<F>(U2, q.x, acc.ZZ);
    fe_mul<F>(S2, q.y, acc.ZZZ);
    fe_sub<F>(P, U2, acc.X);
    fe_sub<F>(R, S2, acc.Y);
    if (fe_is_zero<F>(P))
    {
        if (fe_is_zero<F>(R))
            xyzz_mdbl<F>(acc, q); // acc == q
        else
            acc_id = true; // acc == -q
        return;
    }
    fe PP, PPP, Q, t;
    fe_sq<F>(PP, P);
    fe_mul<F>(PPP, P, PP);
    fe_mul<F>(Q, acc.X, PP);
    fe X3, Y3;
    fe_sq<F>(X3, R);
    fe_sub<F>(X3, X3, PPP);
    fe_sub<F>(X3, X3, Q);
    fe_sub<F>(X3, X3, Q);
    fe_sub<F>(t, Q, X3);
    fe_mul<F>(Y3, R, t);
    fe_mul<F>(t, acc.Y, PPP);
    fe_sub<F>(Y3, Y3, t);
    fe_mul<F>(acc.ZZ, acc.ZZ, PP);
    fe_mul<F>(acc.ZZZ, acc.ZZZ, PPP);
    acc.X = X3;
    acc.Y = Y3;
}

// r = p + q (both xyzz), add-2008-s, with identity / doubling / inverse handling.
template <class F> HS_HD void xyzz_add(xyzz &r, const xyzz &p, const xyzz &q)
{
    if (xyzz_is_identity<F>(p))
    {
        r = q;
        return;
    }
    if (xyzz_is_identity<F>(q))
    {
        r = p;
        return;
    }
    fe U1, U2, S1, S2, P, R;
    fe_mul<F>(U1, p.X, q.ZZ);
    fe_mul<F>(U2, q.X, p.ZZ);
    fe_mul<F>(S1, p.Y, q.ZZZ);
    fe_mul<F>(S2, q.Y, p.ZZZ);
    fe_sub<F>(P, U2, U1);
    fe_sub<F>(R, S2, S1);
    if (fe_is_zero<F>(P))
    {
        if (fe_is_zero<F>(R))
            xyzz_dbl<F>(r, p);
        else
            xyzz_identity<F>(r);
        return;
    }
    fe PP, PPP, Q, t;
    fe_sq<F>(PP, P);
    fe_mul<F>(PPP, P, PP);
    fe_mul<F>(Q, U1, PP);
    fe X3, Y3;
    fe_sq<F>(X3, R);
    fe_sub<F>(X3, X3, PPP);
    fe_sub<F>(X3, X3, Q);
    fe_sub<F>(X3, X3, Q);
    fe_sub<F>(t, Q, X3);
    fe_mul<F>(Y3, R, t);
    fe_mul<F>(t, S1, PPP);
    fe_sub<F>(Y3, Y3, t);
    fe ZZ3, ZZZ3;
    fe_mul<F>(ZZ3, p.ZZ, q.ZZ);
    fe_mul<F>(ZZ3, ZZ3, PP);
    fe_mul<F>(ZZZ3, p.ZZZ, q.ZZZ);
    fe_mul<F>(ZZZ3, ZZZ3, PPP);
    r.X = X3;
    r.Y = Y3;
    r.ZZ = ZZ3;
    r.ZZZ = ZZZ3;
}

template <class F> HS_HD void xyzz_neg(xyzz &r, const xyzz &p)
{
    r = p;
    fe_neg<F>(r.Y, p.Y);
}

// k * p for a small non-negative k (binary, most significant bit first)
template <class F> HS_HD void xyzz_mul_small(xyzz &r, const xyzz &p, uint32_t k)
{
    xyzz acc;
    xyzz_identity<F>(acc);
    if (k == 0 || xyzz_is_identity<F>(p))
    {
        r = acc;
        return;
    }
    int top = 31;
    while (!((k >> top) & 1))
        top--;
    acc = p;
    for (int i = top - 1; i >= 0; i--)
    {
        xyzz_dbl<F>(acc, acc);
        if ((k >> i) & 1)
            xyzz_add<F>(acc, acc, p);
    }
    r = acc;
}

// xyzz -> affine (canonical); returns false for the identity.
template <class F> HS_HD bool xyzz_to_aff(aff &r, const xyzz &p)
{
    if (xyzz_is_identity<F>(p))
    {
        fe_zero(r.x);
        fe_zero(r.y);
        return false;
    }
    // one inversion: i = (ZZ*ZZZ)^-1, 1/ZZ = i*ZZZ, 1/ZZZ = i*ZZ
    fe t, i, izz, izzz;
    fe_mul<F>(t, p.ZZ, p.ZZZ);
    fe_inv<F>(i, t);
    fe_mul<F>(izz, i, p.ZZZ);
    fe_mul<F>(izzz, i, p.ZZ);
    fe_mul<F>(t, p.X, izz);
    fe_canon<F>(r.x, t);
    fe_mul<F>(t, p.Y, izzz);
    fe_canon<F>(r.y, t);
    return true;
}

// 32-byte compressed encoding (reference helios_tobytes.cpp:42-66): canonical x little-endian,
// bit 255 = parity of canonical y, identity = 32 zero bytes.
template <class F> HS_HD void encode_aff(uint32_t out[8], const aff &a, bool is_point)
{
    uint32_t xw[8], yw[8];
    fe_to_words_canon<F>(xw, a.x);
    fe_to_words_canon<F>(yw, a.y);
#pragma unroll
    for (int i = 0; i < 8; i++)
        out[i] = is_point ? xw[i] : 0u;
    if (is_point)
        out[7] |= (yw[0] & 1u) << 31;
}

} // namespace hs
