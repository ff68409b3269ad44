// Jacobian doubling chains and additions with every field element spread over a DPP row
// (rowfield.h): the wavefront's four rows compute up to four independent products of a formula
// level at once and hand the results to every row (two permlane swaps).  Same formulas and
// operand order as quad.h's jac_dbl_n_q / jac_add_q (dbl-2001-b with a = -3, add-2007-bl with the
// identity / equal / opposite cases), so the results are the same residues.  A whole wavefront
// works on one chain: all 64 lanes must be active.
#pragma once
#include "curve.h"
#include "rowfield.h"

namespace hs
{
namespace row
{

struct rjac
{
    rfe X, Y, Z;
};

template <class F> HS_RW void rjac_load(const RowK<F> &k, rjac &r, const jac *p)
{
    r.X = rload<F>(k, &p->X);
    r.Y = rload<F>(k, &p->Y);
    r.Z = rload<F>(k, &p->Z);
}

// row 0's point, normalised to field.h's format (limbs <= 2^29), as a wave-uniform value
template <class F> HS_RW void rjac_get(jac &r, const rjac &a)
{
    rfe_get(r.X, a.X);
    rfe_get(r.Y, a.Y);
    rfe_get(r.Z, a.Z);
    fe_norm<F>(r.X);
    fe_norm<F>(r.Y);
    fe_norm<F>(r.Z);
}

template <class F> HS_RW void rjac_identity(const RowK<F> &k, rjac &r)
{
    r.X.v = sel(k.is0, (u32v)1u, (u32v)0u);
    r.Y = r.X;
    r.Z.v = 0u;
}

// n doublings, pipelined across the doublings as jac_dbl_n_q: with delta = Z^2 and
// alpha = 3 (X - delta)(X + delta) known, a doubling is three product levels
//   A: gamma = Y^2, (Y + Z)^2, alpha^2
//   B: beta = X gamma, gamma^2, delta3 = Z3^2
//   C: alpha (4 beta - X3), (X3 - delta3)(X3 + delta3)
template <class F> HS_RW void rjac_dbl_n(const RowK<F> &k, rjac &p, int n)
{
    if (n <= 0)
        return;
    rfe d = rsq<F>(k, p.Z);
    rfe t = rmul<F>(k, rsub<F>(k, p.X, d), radd_lazy(p.X, d));
    rfe a = rmul_small<F>(k, t, 3);
    for (int j = 0; j < n; j++)
    {
        rfe g, yz2, a2, b, g2, d3, Y3, t3;
        const rfe yz = radd_lazy(p.Y, p.Z);
        rdist3(g, yz2, a2, rsq<F>(k, rsel3<F>(k, p.Y, yz, a)));
        const rfe Z3 = rsub2<F>(k, yz2, g, d);
        rdist3(b, g2, d3, rmul<F>(k, rsel3<F>(k, p.X, g, Z3), rsel3<F>(k, g, g, Z3)));
        const rfe X3 = rsub_small_mul<F>(k, a2, b, 8);
        t = rsub_lazy<F>(k, rmul_small<F>(k, b, 4), X3);
        const rfe xm = rsub<F>(k, X3, d3), xp = radd_lazy(X3, d3);
        rdist2(Y3, t3, rmul<F>(k, rsel2<F>(k, a, xm), rsel2<F>(k, t, xp)));
        Y3 = rsub_small_mul<F>(k, Y3, g2, 8);
        a = rmul_small<F>(k, t3, 3);
        d = d3;
        p.X = X3;
        p.Y = Y3;
        p.Z = Z3;
    }
}

// r = p + q (add-2007-bl), five product levels; the identity, P == Q and P == -Q cases as jac_add_q
template <class F> HS_RW void rjac_add(const RowK<F> &k, rjac &r, const rjac &p, const rjac &q)
{
    if (ris_zero<F>(p.Z))
    {
        r = q;
        return;
    }
    if (ris_zero<F>(q.Z))
    {
        r = p;
        return;
    }
    rfe Z1Z1, Z2Z2, zs2, U1, U2, a, b, S1, S2, I, Z3, J, V, r2, t1, Y3;
    rdist3(Z1Z1, Z2Z2, zs2, rsq<F>(k, rsel3<F>(k, p.Z, q.Z, radd_lazy(p.Z, q.Z))));
    rdist4(U1, U2, a, b, rmul<F>(k, rsel4<F>(k, p.X, q.X, q.Z, p.Z), rsel4<F>(k, Z2Z2, Z1Z1, Z2Z2, Z1Z1)));
    const rfe H = rsub<F>(k, U2, U1);
    const rfe t0 = radd_lazy(H, H);
    const rfe Zt = rsub2<F>(k, zs2, Z1Z1, Z2Z2);
    rdist4(S1, S2, I, Z3, rmul<F>(k, rsel4<F>(k, p.Y, q.Y, t0, Zt), rsel4<F>(k, a, b, t0, H)));
    rfe rr = rsub<F>(k, S2, S1);
    if (ris_zero<F>(H))
    {
        if (ris_zero<F>(rr))
        {
            r = p;
            rjac_dbl_n<F>(k, r, 1);
        }
        else
            rjac_identity<F>(k, r);
        return;
    }
    rr = radd<F>(k, rr, rr);
    rdist3(J, V, r2, rmul<F>(k, rsel3<F>(k, H, U1, rr), rsel3<F>(k, I, I, rr)));
    const rfe X3 = rsub3<F>(k, r2, J, V, V);
    rdist2(t1, Y3, rmul<F>(k, rsel2<F>(k, rr, S1), rsel2<F>(k, rsub_lazy<F>(k, V, X3), J)));
    r.X = X3;
    r.Y = rsub2<F>(k, t1, Y3, Y3);
    r.Z = Z3;
}

// Horner over window sums win[w_lo .. w_hi), highest first: acc = 2^c acc + S_w, starting from
// win[w_hi - 1] (first) or from *start; returns the total (row form)
template <class F>
HS_RW void rhorner(const RowK<F> &k, rjac &acc, const jac *win, int w_lo, int w_hi, int c, bool first, const jac *start)
{
    int w = w_hi - 1;
    if (first)
    {
        rjac_load<F>(k, acc, win + w);
        w--;
    }
    else
        rjac_load<F>(k, acc, start);
    for (; w >= w_lo; w--)
    {
        // doublings of the identity are skipped: windows above every scalar's top digit are empty
        if (!ris_zero<F>(acc.Z))
            rjac_dbl_n<F>(k, acc, c);
        rjac t;
        rjac_load<F>(k, t, win + w);
        rjac_add<F>(k, acc, acc, t);
    }
}

} // namespace row
} // namespace hs
