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
    fe_add_lazy<F>(t1, x, z); // multiplied by the normalised t0 only
    fe_mul<F>(t0, t0, t1);
    fe_mul_small<F>(m, t0, 3);
}

// dbl-2008-s-1 with a = -3 (M = 3X^2 - 3ZZ^2 = 3(X-ZZ)(X+ZZ)); 7M + 2S
template <class F> HS_HD void xyzz_dbl(xyzz &r, const xyzz &p)
{
    fe U, V, W, S, M, t, nY;
    fe_add<F>(U, p.Y, p.Y);
    fe_sq<F>(V, U);
    fe_mul<F>(W, U, V);
    fe_mul<F>(S, p.X, V);
    tangent_m<F>(M, p.X, p.ZZ);
    fe X3, Y3, z;
    fe_zero(z);
    fe_neg3_lazy<F>(t, S, S, z);
    fe_sq_add<F>(X3, M, t); // M^2 - 2S
    fe_sub_lazy2<F>(t, S, X3);
    fe_neg_lazy2<F>(nY, p.Y);
    fe_mul_sum2<F>(Y3, M, t, nY, W); // M (S - X3) - W Y
    fe_mul<F>(r.ZZ, V, p.ZZ);
    fe_mul<F>(r.ZZZ, W, p.ZZZ);
    r.X = X3;
    r.Y = Y3;
}

// mdbl-2008-s-1 with a = -3: doubling of an affine point into xyzz
template <class F> HS_HD void xyzz_mdbl(xyzz &r, const aff &p)
{
    fe one, U, V, W, S, M, t, nY;
    fe_one(one);
    fe_add<F>(U, p.y, p.y);
    fe_sq<F>(V, U);
    fe_mul<F>(W, U, V);
    fe_mul<F>(S, p.x, V);
    tangent_m<F>(M, p.x, one);
    fe X3, Y3, z;
    fe_zero(z);
    fe_neg3_lazy<F>(t, S, S, z);
    fe_sq_add<F>(X3, M, t); // M^2 - 2S
    fe_sub_lazy2<F>(t, S, X3);
    fe_neg_lazy2<F>(nY, p.y);
    fe_mul_sum2<F>(Y3, M, t, nY, W); // M (S - X3) - W y
    r.X = X3;
    r.Y = Y3;
    r.ZZ = V;
    r.ZZZ = W;
}

// acc += q (affine), madd-2008-s, with identity / doubling / inverse handling.
// `acc_id` carries acc's identity state so the hot loop needs no zero test on ZZ.  q.x is
// normalised; q.y may also be a lazy negation (fe_neg_lazy2, limbs < 2^30): on the common path it
// is only a product operand, and the identity / doubling paths normalise it first.
template <class F> HS_HD void xyzz_madd(xyzz &acc, bool &acc_id, const aff &q)
{
    if (acc_id)
    {
        xyzz_from_aff<F>(acc, q);
        fe_norm<F>(acc.Y);
        acc_id = false;
        return;
    }
    // P = x2 ZZ1 - X1 and R = y2 ZZZ1 - Y1, each one product with the negation as addend
    fe nX, nY, P, R;
    fe_neg_lazy2<F>(nX, acc.X);
    fe_neg_lazy2<F>(nY, acc.Y);
    fe_mul_add<F>(P, q.x, acc.ZZ, nX);
    fe_mul_add<F>(R, q.y, acc.ZZZ, nY);
    fe_opaque(P);
    fe_opaque(R);
    if (fe_is_zero_norm<F>(P))
    {
        if (fe_is_zero_norm<F>(R))
        {
            aff d = q;
            fe_norm<F>(d.y);
            xyzz_mdbl<F>(acc, d); // acc == q
        }
        else
            acc_id = true; // acc == -q
        return;
    }
    fe PP, PPP, Q, t;
    fe_sq<F>(PP, P);
    fe_mul<F>(PPP, P, PP);
    fe_mul<F>(Q, acc.X, PP);
    fe X3, Y3;
    fe_neg3_lazy<F>(t, PPP, Q, Q);
    fe_sq_add<F>(X3, R, t); // R^2 - PPP - 2Q
    fe_sub_lazy2<F>(t, Q, X3);
    fe_mul_sum2<F>(Y3, R, t, nY, PPP); // R (Q - X3) - Y1 PPP
    fe_mul<F>(acc.ZZ, acc.ZZ, PP);
    fe_mul<F>(acc.ZZZ, acc.ZZZ, PPP);
    acc.X = X3;
    acc.Y = Y3;
}

// The same addition for the bucket walk, laid out for SIMD lanes that nearly always take the
// common path: the madd-2008-s result is formed unconditionally and written to acc; the rare cases
// (acc was the identity, acc == q, acc == -q) are fixed up afterwards in a branch the whole
// wavefront takes only when one of its lanes needs it, and the fix-up selects limb by limb.  A
// divergent branch that redefines acc would make the compiler keep the common result in other
// registers and copy it back on every step.
template <class F> HS_HD void xyzz_madd_hot(xyzz &acc, bool &acc_id, const aff &q)
{
    fe nX, nY, P, R;
    fe_neg_lazy2<F>(nX, acc.X);
    fe_neg_lazy2<F>(nY, acc.Y);
    fe_mul_add<F>(P, q.x, acc.ZZ, nX);
    fe_mul_add<F>(R, q.y, acc.ZZZ, nY);
    fe_opaque(P);
    fe_opaque(R);
    // only the cheap exclusion runs here; the canonical zero test waits in the rare branch
    const bool rare = acc_id || fe_maybe_zero_norm(P);
    fe PP, PPP, Q, t;
    fe_sq<F>(PP, P);
    fe_mul<F>(PPP, P, PP);
    fe_mul<F>(Q, acc.X, PP);
    fe X3, Y3;
    fe_neg3_lazy<F>(t, PPP, Q, Q);
    fe_sq_add<F>(X3, R, t); // R^2 - PPP - 2Q
    fe_sub_lazy2<F>(t, Q, X3);
    fe_mul_sum2<F>(Y3, R, t, nY, PPP); // R (Q - X3) - Y1 PPP
    fe_mul<F>(acc.ZZ, acc.ZZ, PP);
    fe_mul<F>(acc.ZZZ, acc.ZZZ, PPP);
    acc.X = X3;
    acc.Y = Y3;
#ifdef __HIP_DEVICE_COMPILE__
    const bool any_rare = __builtin_amdgcn_ballot_w64(rare) != 0;
#else
    const bool any_rare = rare;
#endif
    if (any_rare)
    {
        // the empty asm pins these copies inside the branch, so nothing computed from them is
        // hoisted onto the common path
        aff d = q;
        fe_opaque(d.x);
        fe_opaque(d.y);
        fe_opaque(P);
        fe_opaque(R);
        fe_norm<F>(d.y);
        const bool zero_p = rare && !acc_id && fe_is_zero_canon<F>(P);
        const bool same = zero_p && fe_is_zero_norm<F>(R);
        xyzz alt;
        if (same)
            xyzz_mdbl<F>(alt, d); // acc == q
        else
            xyzz_from_aff<F>(alt, d);
        const bool use = same || (rare && acc_id);
#pragma unroll
        for (int i = 0; i < NL; i++)
        {
            acc.X.v[i] = use ? alt.X.v[i] : acc.X.v[i];
            acc.Y.v[i] = use ? alt.Y.v[i] : acc.Y.v[i];
            acc.ZZ.v[i] = use ? alt.ZZ.v[i] : acc.ZZ.v[i];
            acc.ZZZ.v[i] = use ? alt.ZZZ.v[i] : acc.ZZZ.v[i];
        }
        // acc == -q leaves the identity; a start from the identity leaves q
        acc_id = (zero_p && !same) ? true : (acc_id ? false : acc_id);
    }
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
    // P = U2 - U1 and R = S2 - S1, the second product of each taking the negation as addend
    fe U1, S1, nU1, nS1, P, R;
    fe_mul<F>(U1, p.X, q.ZZ);
    fe_mul<F>(S1, p.Y, q.ZZZ);
    fe_neg_lazy2<F>(nU1, U1);
    fe_neg_lazy2<F>(nS1, S1);
    fe_mul_add<F>(P, q.X, p.ZZ, nU1);
    fe_mul_add<F>(R, q.Y, p.ZZZ, nS1);
    fe_opaque(P);
    fe_opaque(R);
    if (fe_is_zero_norm<F>(P))
    {
        if (fe_is_zero_norm<F>(R))
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
    fe_neg3_lazy<F>(t, PPP, Q, Q);
    fe_sq_add<F>(X3, R, t); // R^2 - PPP - 2Q
    fe_sub_lazy2<F>(t, Q, X3);
    fe_mul_sum2<F>(Y3, R, t, nS1, PPP); // R (Q - X3) - S1 PPP
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

// ---------------------------------------------------------------- Jacobian (x = X/Z^2, y = Y/Z^3)
// Used for long doubling chains (window combine), where dbl-2001-b (3M + 5S, dependency depth
// four) beats the xyzz doubling.
struct jac
{
    fe X, Y, Z;
};

template <class F> HS_HD void jac_identity(jac &r)
{
    fe_one(r.X);
    fe_one(r.Y);
    fe_zero(r.Z);
}

// xyzz -> Jacobian without inversion: Z' = ZZ*ZZZ, X' = X*ZZ*ZZZ^2, Y' = Y*ZZ^3*ZZZ^2
// (X'/Z'^2 = X/ZZ and Y'/Z'^3 = Y/ZZZ for any ZZ, ZZZ).
template <class F> HS_HD void xyzz_to_jac(jac &r, const xyzz &p)
{
    fe zz2, zzz2, t;
    fe_mul<F>(r.Z, p.ZZ, p.ZZZ);
    fe_sq<F>(zzz2, p.ZZZ);
    fe_mul<F>(t, p.X, p.ZZ);
    fe_mul<F>(r.X, t, zzz2);
    fe_sq<F>(zz2, p.ZZ);
    fe_mul<F>(t, zz2, p.ZZ);
    fe_mul<F>(t, t, zzz2);
    fe_mul<F>(r.Y, p.Y, t);
}

// dbl-2001-b, a = -3 (reference helios/src/x64/helios_dbl.cpp:32-88); the identity (Z = 0)
// maps to itself.
template <class F> HS_HD void jac_dbl(jac &r, const jac &p)
{
    fe delta, gamma, beta, alpha, t0, t1, X3, Y3, Z3;
    fe_sq<F>(delta, p.Z);
    fe_sq<F>(gamma, p.Y);
    fe_mul<F>(beta, p.X, gamma);
    fe_sub<F>(t0, p.X, delta);
    fe_add<F>(t1, p.X, delta);
    fe_mul<F>(alpha, t0, t1);
    fe_mul_small<F>(alpha, alpha, 3);
    fe_sq<F>(X3, alpha);
    fe_mul_small<F>(t0, beta, 8);
    fe_sub<F>(X3, X3, t0);
    fe_add<F>(t1, p.Y, p.Z);
    fe_sq<F>(Z3, t1);
    fe_sub<F>(Z3, Z3, gamma);
    fe_sub<F>(Z3, Z3, delta);
    fe_mul_small<F>(t0, beta, 4);
    fe_sub<F>(t0, t0, X3);
    fe_mul<F>(Y3, alpha, t0);
    fe_sq<F>(t1, gamma);
    fe_mul_small<F>(t1, t1, 8);
    fe_sub<F>(Y3, Y3, t1);
    r.X = X3;
    r.Y = Y3;
    r.Z = Z3;
}

// add-2007-bl with the identity / equal / opposite cases (helios_add.cpp:32-107,
// helios/include/helios_add.h:44-87)
template <class F> HS_HD void jac_add(jac &r, const jac &p, const jac &q)
{
    if (fe_is_zero<F>(p.Z))
    {
        r = q;
        return;
    }
    if (fe_is_zero<F>(q.Z))
    {
        r = p;
        return;
    }
    fe Z1Z1, Z2Z2, U1, U2, S1, S2, H, I, J, rr, V, t0, t1;
    fe_sq<F>(Z1Z1, p.Z);
    fe_sq<F>(Z2Z2, q.Z);
    fe_mul<F>(U1, p.X, Z2Z2);
    fe_mul<F>(U2, q.X, Z1Z1);
    fe_mul<F>(t0, q.Z, Z2Z2);
    fe_mul<F>(S1, p.Y, t0);
    fe_mul<F>(t0, p.Z, Z1Z1);
    fe_mul<F>(S2, q.Y, t0);
    fe_sub<F>(H, U2, U1);
    fe_sub<F>(rr, S2, S1);
    if (fe_is_zero<F>(H))
    {
        if (fe_is_zero<F>(rr))
            jac_dbl<F>(r, p);
        else
            jac_identity<F>(r);
        return;
    }
    fe_add<F>(t0, H, H);
    fe_sq<F>(I, t0);
    fe_mul<F>(J, H, I);
    fe_add<F>(rr, rr, rr);
    fe_mul<F>(V, U1, I);
    fe X3, Y3, Z3;
    fe_sq<F>(X3, rr);
    fe_sub<F>(X3, X3, J);
    fe_add<F>(t0, V, V);
    fe_sub<F>(X3, X3, t0);
    fe_sub<F>(t0, V, X3);
    fe_mul<F>(t1, rr, t0);
    fe_mul<F>(t0, S1, J);
    fe_add<F>(t0, t0, t0);
    fe_sub<F>(Y3, t1, t0);
    fe_add<F>(t0, p.Z, q.Z);
    fe_sq<F>(t1, t0);
    fe_sub<F>(t1, t1, Z1Z1);
    fe_sub<F>(t1, t1, Z2Z2);
    fe_mul<F>(Z3, t1, H);
    r.X = X3;
    r.Y = Y3;
    r.Z = Z3;
}

// Jacobian -> affine (canonical); returns false for the identity
template <class F> HS_HD bool jac_to_aff(aff &r, const jac &p)
{
    if (fe_is_zero<F>(p.Z))
    {
        fe_zero(r.x);
        fe_zero(r.y);
        return false;
    }
    fe zi, zi2, zi3, t;
    fe_inv<F>(zi, p.Z);
    fe_sq<F>(zi2, zi);
    fe_mul<F>(zi3, zi2, zi);
    fe_mul<F>(t, p.X, zi2);
    fe_canon<F>(r.x, t);
    fe_mul<F>(t, p.Y, zi3);
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
