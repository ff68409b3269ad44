// Test infrastructure: the product's row-distributed field and curve arithmetic
// (helioselene_amd/csrc/rowfield.h, rowcurve.h) compiled for the host over the lane model of
// row_emu.h, beside the single-lane arithmetic of field.h / curve.h it must agree with.  Built by
// tests/native/Makefile; never part of the product library.
#include <stdint.h>
#include <string.h>
#define HS_HD inline
#include "row_emu.h"
#include "../../helioselene_amd/csrc/rowcurve.h"

namespace hs
{
namespace row
{
long g_emu_wrap = 0;
}
} // namespace hs

using namespace hs;
using namespace hs::row;

// four rows of nine limbs (row-major) -> one row element per row, zeros in lanes 9..15
static rfe put(const uint32_t *v)
{
    rfe r;
    for (int i = 0; i < 64; i++)
        r.v.l[i] = (i & 15) < 9 ? v[9 * (i >> 4) + (i & 15)] : 0u;
    return r;
}

// out: four rows of nine limbs; stats[0] = wraps, stats[1] = largest limb in lanes 0..8,
// stats[2] = nonzero lanes 9..15 found
static void get(uint32_t *out, uint32_t *stats, const rfe &a)
{
    for (int i = 0; i < 64; i++)
    {
        const int l = i & 15;
        if (l < 9)
        {
            out[9 * (i >> 4) + l] = a.v.l[i];
            stats[1] = a.v.l[i] > stats[1] ? a.v.l[i] : stats[1];
        }
        else if (a.v.l[i])
            stats[2]++;
    }
}

template <class F> static rfe op(int o, const rfe &a, const rfe &b, const rfe &c, const rfe &d, uint32_t s)
{
    RowK<F> k;
    rowk_init(k);
    switch (o)
    {
    case 0: return rmul<F>(k, a, b);
    case 1: return rsq<F>(k, a);
    case 2: return rsub<F>(k, a, b);
    case 3: return radd<F>(k, a, b);
    case 4: return rsub2<F>(k, a, b, c);
    case 5: return rsub3<F>(k, a, b, c, d);
    case 6: return rmul_small<F>(k, a, s);
    case 7: return rsub_small_mul<F>(k, a, b, s);
    case 8: return rmul<F>(k, a, rsub_lazy<F>(k, b, c));      // RN x lazy difference
    case 9: return rmul<F>(k, radd_lazy(a, b), radd_lazy(c, d)); // lazy sum x lazy sum
    case 10: return rnorm<F>(k, a.v);                          // raw limbs < 2^32
    case 11: return rsq<F>(k, radd_lazy(a, b));
    default: return a;
    }
}

// one operation per row on per-row inputs (nine limbs each, any values the operation admits)
extern "C" void hst_row_op(int field, int o, uint32_t *out, uint32_t *stats, const uint32_t *a, const uint32_t *b,
                           const uint32_t *c, const uint32_t *d, uint32_t s)
{
    g_emu_wrap = 0;
    const rfe A = put(a), B = put(b), C = put(c), D = put(d);
    const rfe r = field == 0 ? op<Fp>(o, A, B, C, D, s) : op<Fq>(o, A, B, C, D, s);
    stats[0] = (uint32_t)g_emu_wrap;
    stats[1] = 0;
    stats[2] = 0;
    get(out, stats, r);
}

// all four rows must agree after a distribution; row 0's point as 27 limbs
static void get_jac(uint32_t *out, uint32_t *stats, const rjac &p)
{
    uint32_t tmp[36];
    const rfe *c[3] = {&p.X, &p.Y, &p.Z};
    for (int k = 0; k < 3; k++)
    {
        get(tmp, stats, *c[k]);
        for (int r = 1; r < 4; r++)
            for (int l = 0; l < 9; l++)
                if (tmp[9 * r + l] != tmp[l])
                    stats[3]++;
        memcpy(out + 9 * k, tmp, 36);
    }
}

template <class F> static void curve_ops(int o, uint32_t *out, const uint32_t *p, const uint32_t *q, int n)
{
    RowK<F> k;
    rowk_init(k);
    rjac P, Q, R;
    rjac_load<F>(k, P, (const jac *)p);
    rjac_load<F>(k, Q, (const jac *)q);
    if (o == 0)
    {
        R = P;
        rjac_dbl_n<F>(k, R, n);
    }
    else if (o == 1)
        rjac_add<F>(k, R, P, Q);
    else
    {
        // Horner over n window sums at p (q: c in q[0]), highest first
        rhorner<F>(k, R, (const jac *)p, 0, n, (int)q[0], true, nullptr);
    }
    uint32_t st[4] = {0, 0, 0, 0};
    get_jac(out, st, R);
    out[27] = (uint32_t)g_emu_wrap;
    out[28] = st[1];
    out[29] = st[2];
    out[30] = st[3];
    jac J;
    rjac_get<F>(J, R);
    memcpy(out + 31, &J, sizeof J);
}

// o = 0: n doublings of p; 1: p + q; 2: Horner over n points at p with c = q[0].  p, q: jac
// structs (27 limbs each, limbs <= 2^29).  out: 27 raw limbs of row 0, wraps, largest limb,
// nonzero lanes 9..15, rows disagreeing, then the 27 limbs of the normalised wave-uniform copy.
extern "C" void hst_row_curve(int field, int o, uint32_t *out, const uint32_t *p, const uint32_t *q, int n)
{
    g_emu_wrap = 0;
    if (field == 0)
        curve_ops<Fp>(o, out, p, q, n);
    else
        curve_ops<Fq>(o, out, p, q, n);
}

// the same operations on one lane with curve.h (jac_dbl, jac_add): 27 limbs out
extern "C" void hst_lane_curve(int field, int o, uint32_t *out, const uint32_t *p, const uint32_t *q, int n)
{
    jac P, Q, R;
    memcpy(&P, p, sizeof P);
    memcpy(&Q, q, sizeof Q);
    if (field == 0)
    {
        if (o == 0)
        {
            R = P;
            for (int i = 0; i < n; i++)
                jac_dbl<Fp>(R, R);
        }
        else if (o == 1)
            jac_add<Fp>(R, P, Q);
        else
        {
            const jac *w = (const jac *)p;
            R = w[n - 1];
            for (int i = n - 2; i >= 0; i--)
            {
                if (!fe_is_zero_norm<Fp>(R.Z))
                    for (uint32_t j = 0; j < q[0]; j++)
                        jac_dbl<Fp>(R, R);
                jac_add<Fp>(R, R, w[i]);
            }
        }
    }
    else
    {
        if (o == 0)
        {
            R = P;
            for (int i = 0; i < n; i++)
                jac_dbl<Fq>(R, R);
        }
        else if (o == 1)
            jac_add<Fq>(R, P, Q);
        else
        {
            const jac *w = (const jac *)p;
            R = w[n - 1];
            for (int i = n - 2; i >= 0; i--)
            {
                if (!fe_is_zero_norm<Fq>(R.Z))
                    for (uint32_t j = 0; j < q[0]; j++)
                        jac_dbl<Fq>(R, R);
                jac_add<Fq>(R, R, w[i]);
            }
        }
    }
    memcpy(out, &R, sizeof R);
}
