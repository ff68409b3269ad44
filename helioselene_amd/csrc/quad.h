// Point arithmetic spread over a quad of lanes, for the latency-bound serial chains (Horner
// combines, window-sum conversions).
//
// A single lane walking 256 doublings is issue-bound: every field product is ~200 dependent-ish
// VALU instructions and no other wave hides them.  Here the four lanes of a quad hold the same
// point; each lane computes one of up to four independent products of a formula level, and a DPP
// quad_perm broadcast (one VALU instruction per limb) hands every result to the whole quad.
// dbl-2001-b then takes 4 product levels instead of 8 products, add-2007-bl 5 instead of 16.
// All four lanes of a quad must be active; results are identical in the four lanes.
#pragma once
#include "curve.h"

namespace hs
{

template <int K> __device__ __forceinline__ void fe_qbcast(fe &r, const fe &a)
{
#pragma unroll
    for (int i = 0; i < NL; i++)
        r.v[i] = (uint32_t)__builtin_amdgcn_mov_dpp((int)a.v[i], K * 0x55, 0xf, 0xf, false);
}

// lane q of the quad takes a_q.  Per-limb selects on the lane's two quad-index bits (v_cndmask with
// loop-invariant lane masks: three per limb for four sources, two for three, one for two); a
// select between the structs themselves would be lowered to a select of their stack addresses.
__device__ __forceinline__ void fe_qsel(fe &r, const fe &a0, const fe &a1, const fe &a2, const fe &a3)
{
    const bool lo = (threadIdx.x & 1u) != 0, hi = (threadIdx.x & 2u) != 0;
#pragma unroll
    for (int i = 0; i < NL; i++)
    {
        const uint32_t x = lo ? a1.v[i] : a0.v[i];
        const uint32_t y = lo ? a3.v[i] : a2.v[i];
        r.v[i] = hi ? y : x;
    }
}

// lanes 0, 1, 2, 3 take a0, a1, a2, a2
__device__ __forceinline__ void fe_qsel3(fe &r, const fe &a0, const fe &a1, const fe &a2)
{
    const bool lo = (threadIdx.x & 1u) != 0, hi = (threadIdx.x & 2u) != 0;
#pragma unroll
    for (int i = 0; i < NL; i++)
    {
        const uint32_t x = lo ? a1.v[i] : a0.v[i];
        r.v[i] = hi ? a2.v[i] : x;
    }
}

// lanes 0, 1, 2, 3 take a0, a1, a1, a1
__device__ __forceinline__ void fe_qsel2(fe &r, const fe &a0, const fe &a1)
{
    const bool z = (threadIdx.x & 3u) == 0;
#pragma unroll
    for (int i = 0; i < NL; i++)
        r.v[i] = z ? a0.v[i] : a1.v[i];
}

// four independent products a_k * b_k, one per lane of the quad
template <class F>
__device__ __forceinline__ void fe_mul_q4(fe &r0, fe &r1, fe &r2, fe &r3, const fe &a0, const fe &b0, const fe &a1,
                                          const fe &b1, const fe &a2, const fe &b2, const fe &a3, const fe &b3)
{
    fe a, b, r;
    fe_qsel(a, a0, a1, a2, a3);
    fe_qsel(b, b0, b1, b2, b3);
    fe_mul<F>(r, a, b);
    fe_qbcast<0>(r0, r);
    fe_qbcast<1>(r1, r);
    fe_qbcast<2>(r2, r);
    fe_qbcast<3>(r3, r);
}

template <class F>
__device__ __forceinline__ void fe_mul_q3(fe &r0, fe &r1, fe &r2, const fe &a0, const fe &b0, const fe &a1,
                                          const fe &b1, const fe &a2, const fe &b2)
{
    fe a, b, r;
    fe_qsel3(a, a0, a1, a2);
    fe_qsel3(b, b0, b1, b2);
    fe_mul<F>(r, a, b);
    fe_qbcast<0>(r0, r);
    fe_qbcast<1>(r1, r);
    fe_qbcast<2>(r2, r);
}

// three independent squares a_k^2, one per lane (lanes 2 and 3 both square a2)
template <class F>
__device__ __forceinline__ void fe_sq_q3(fe &r0, fe &r1, fe &r2, const fe &a0, const fe &a1, const fe &a2)
{
    fe a, r;
    fe_qsel3(a, a0, a1, a2);
    fe_sq<F>(r, a);
    fe_qbcast<0>(r0, r);
    fe_qbcast<1>(r1, r);
    fe_qbcast<2>(r2, r);
}

template <class F>
__device__ __forceinline__ void fe_mul_q2(fe &r0, fe &r1, const fe &a0, const fe &b0, const fe &a1, const fe &b1)
{
    fe a, b, r;
    fe_qsel2(a, a0, a1);
    fe_qsel2(b, b0, b1);
    fe_mul<F>(r, a, b);
    fe_qbcast<0>(r0, r);
    fe_qbcast<1>(r1, r);
}

// dbl-2001-b (a = -3), same formula as jac_dbl: levels {Z^2, Y^2, (Y+Z)^2}, {X gamma,
// (X - delta)(X + delta), gamma^2}, {alpha^2}, {alpha (4 beta - X3)}.
template <class F> __device__ __forceinline__ void jac_dbl_q(jac &p)
{
    // Operand bounds: Y + Z and X + delta stay un-normalised (limbs <= 2^30), which fe_mul accepts
    // for both operands; 4 beta - X3 is a lazy subtraction, multiplied by the normalised alpha.
    fe yz, delta, gamma, yz2, xm, xp, beta, t, g2, alpha, X3, Y3, Z3, t0;
    fe_add_lazy<F>(yz, p.Y, p.Z);
    fe_sq_q3<F>(delta, gamma, yz2, p.Z, p.Y, yz);
    fe_sub<F>(xm, p.X, delta);
    fe_add_lazy<F>(xp, p.X, delta);
    fe_mul_q3<F>(beta, t, g2, p.X, gamma, xm, xp, gamma, gamma);
    fe_mul_small<F>(alpha, t, 3);
    fe_sq<F>(X3, alpha);
    fe_mul_small<F>(t0, beta, 8);
    fe_sub<F>(X3, X3, t0);
    fe_sub2<F>(Z3, yz2, gamma, delta);
    fe_mul_small<F>(t0, beta, 4);
    fe_sub_lazy<F>(t0, t0, X3);
    fe_mul<F>(Y3, alpha, t0);
    fe_mul_small<F>(t0, g2, 8);
    fe_sub<F>(Y3, Y3, t0);
    p.X = X3;
    p.Y = Y3;
    p.Z = Z3;
}

// n doublings of p (dbl-2001-b, same results as n calls of jac_dbl_q), pipelined across the
// doublings: a doubling of (X, Y, Z) with delta = Z^2 and alpha = 3 (X - delta)(X + delta)
// already known takes three product levels,
//   A: gamma = Y^2, (Y + Z)^2, alpha^2                 -> Z3 = (Y + Z)^2 - gamma - delta
//   B: beta = X gamma, gamma^2, delta3 = Z3^2          -> X3 = alpha^2 - 8 beta
//   C: alpha (4 beta - X3), (X3 - delta3)(X3 + delta3) -> Y3, alpha3
// since delta and alpha of the next doubling come out of levels B and C.  Two levels prepare the
// first delta and alpha: 2 + 3n levels instead of 4n.
template <class F> __device__ __forceinline__ void jac_dbl_n_q(jac &p, int n)
{
    if (n <= 0)
        return;
    fe d, a, xm, xp, t;
    fe_sq<F>(d, p.Z);
    fe_sub<F>(xm, p.X, d);
    fe_add_lazy<F>(xp, p.X, d);
    fe_mul<F>(t, xm, xp);
    fe_mul_small<F>(a, t, 3);
    for (int k = 0; k < n; k++)
    {
        // Operand bounds as in jac_dbl_q: Y + Z and X3 + delta3 are un-normalised sums (limbs
        // <= 2^30), 4 beta - X3 a lazy subtraction multiplied by the normalised alpha.
        fe yz, g, yz2, a2, Z3, b, g2, d3, X3, Y3, t3;
        fe_add_lazy<F>(yz, p.Y, p.Z);
        fe_sq_q3<F>(g, yz2, a2, p.Y, yz, a);
        fe_sub2<F>(Z3, yz2, g, d);
        fe_mul_q3<F>(b, g2, d3, p.X, g, g, g, Z3, Z3);
        fe_sub_small_mul<F>(X3, a2, b, 8);
        fe_mul_small<F>(t, b, 4);
        fe_sub_lazy<F>(t, t, X3);
        fe_sub<F>(xm, X3, d3);
        fe_add_lazy<F>(xp, X3, d3);
        fe_mul_q2<F>(Y3, t3, a, t, xm, xp);
        fe_sub_small_mul<F>(Y3, Y3, g2, 8);
        fe_mul_small<F>(a, t3, 3);
        d = d3;
        p.X = X3;
        p.Y = Y3;
        p.Z = Z3;
    }
}

// xyzz -> Jacobian (Z' = ZZ ZZZ, X' = X ZZ ZZZ^2, Y' = (Y ZZ) Z'^2), three levels
template <class F> __device__ __forceinline__ void xyzz_to_jac_q(jac &r, const xyzz &p)
{
    fe Z, zzz2, t, u, z2;
    fe_mul_q4<F>(Z, zzz2, t, u, p.ZZ, p.ZZZ, p.ZZZ, p.ZZZ, p.X, p.ZZ, p.Y, p.ZZ);
    fe_mul_q2<F>(r.X, z2, t, zzz2, Z, Z);
    fe_mul<F>(r.Y, u, z2);
    r.Z = Z;
}

// add-2007-bl with the identity / equal / opposite cases, same results as jac_add
template <class F> __device__ __forceinline__ void jac_add_q(jac &r, const jac &p, const jac &q)
{
    if (fe_is_zero_norm<F>(p.Z))
    {
        r = q;
        return;
    }
    if (fe_is_zero_norm<F>(q.Z))
    {
        r = p;
        return;
    }
    fe Z1Z1, Z2Z2, zs, zs2, U1, U2, a, b, S1, S2, H, rr, I, Zt, Z3, J, V, r2, X3, Y3, t0, t1;
    fe_add_lazy<F>(zs, p.Z, q.Z); // squared below: limbs <= 2^30 are fine for both operands
    fe_sq_q3<F>(Z1Z1, Z2Z2, zs2, p.Z, q.Z, zs);
    fe_mul_q4<F>(U1, U2, a, b, p.X, Z2Z2, q.X, Z1Z1, q.Z, Z2Z2, p.Z, Z1Z1);
    fe_sub<F>(H, U2, U1);
    fe_add_lazy<F>(t0, H, H);
    fe_sub2<F>(Zt, zs2, Z1Z1, Z2Z2);
    fe_mul_q4<F>(S1, S2, I, Z3, p.Y, a, q.Y, b, t0, t0, Zt, H);
    fe_sub<F>(rr, S2, S1);
    if (fe_is_zero_norm<F>(H))
    {
        if (fe_is_zero_norm<F>(rr))
        {
            r = p;
            jac_dbl_q<F>(r);
        }
        else
            jac_identity<F>(r);
        return;
    }
    fe_add<F>(rr, rr, rr);
    fe_mul_q3<F>(J, V, r2, H, I, U1, I, rr, rr);
    fe_sub3<F>(X3, r2, J, V, V);  // rr^2 - J - 2V
    fe_sub_lazy<F>(t0, V, X3);    // multiplied by the normalised rr only
    fe_mul_q2<F>(t1, Y3, rr, t0, S1, J);
    fe_sub2<F>(Y3, t1, Y3, Y3);   // rr (V - X3) - 2 S1 J
    r.X = X3;
    r.Y = Y3;
    r.Z = Z3;
}

} // namespace hs
