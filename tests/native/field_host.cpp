// Test infrastructure: the product's field and curve arithmetic headers compiled for the host
// CPU, so the exact same limb code can be checked against Python big integers and the oracle
// without a GPU.  Built by tests/native/Makefile; never part of the product library.
#include <stdint.h>
#include <string.h>
#define HS_HD static inline
#include "../../helioselene_amd/csrc/field.h"

using namespace hs;

static void ld(fe &r, const uint8_t *b)
{
    for (int i = 0; i < 8; i++)
        r.v[i] = (uint32_t)b[4 * i] | ((uint32_t)b[4 * i + 1] << 8) | ((uint32_t)b[4 * i + 2] << 16) |
                 ((uint32_t)b[4 * i + 3] << 24);
}
static void st(uint8_t *b, const fe &a)
{
    for (int i = 0; i < 8; i++)
        for (int j = 0; j < 4; j++)
            b[4 * i + j] = (uint8_t)(a.v[i] >> (8 * j));
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
    default: memset(&r, 0, sizeof r);
    }
    st(out, r);
}

extern "C" void hst_fe_op(int field, int o, uint8_t *out, const uint8_t *a, const uint8_t *b)
{
    if (field == 0)
        op<Fp>(o, out, a, b);
    else
        op<Fq>(o, out, a, b);
}

extern "C" void hst_fe_from_r51(int field, uint8_t *out, const uint64_t *l)
{
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

extern "C" void hst_pt_op(int field, int o, uint8_t *out, const uint8_t *a, const uint8_t *b, uint32_t k)
{
    if (field == 0)
        pop<Fp>(o, out, a, b, k);
    else
        pop<Fq>(o, out, a, b, k);
}
