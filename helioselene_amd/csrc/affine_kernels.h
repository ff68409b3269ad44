// Bucket accumulation with batched affine additions (a selectable alternative to k_bucket_flat's
// xyzz mixed additions).
//
// Each lane runs K chains; chain j of thread t sums chunk t K + j of the sorted entries, so the
// chunk layout, the partial-sum slots and the merge kernels are those of k_bucket_flat with run
// length Lc.  A step adds one entry to every live chain:
//   add       (x, y) += (x2, y2):   l = (y2 - y) / (x2 - x), x3 = l^2 - x - x2, y3 = l (x - x3) - y
//   doubling  (x2, y2) == (x, y):   l = (3 x^2 - 3) / (2 y)                       (a = -3)
//   cancel    (x2, y2) == (x, -y):  the chain's sum becomes the identity
//   start     the chain's sum is the identity: it becomes (x2, y2)
// The 64 K denominators of a step (1 for the cases without one) are inverted together by
// Montgomery's trick across the wavefront: per-lane prefix products over its K chains, an inclusive
// prefix scan and an exclusive suffix scan of the lane totals over the 64 lanes (shuffles), one
// field inversion of the wave's total, then each denominator's inverse from its neighbours'
// products (the reference's batch inversion, fp/include/fp_batch_invert.h:61-114, and batch
// affine conversion, helios/include/helios_batch_affine.h:55-98, follow the same trick serially).
// Cost per addition: 5M + 1S plus 1/(64 K) of the scans and the inversion's instruction stream,
// which the wave issues once for all of its lanes.
#pragma once
#include "msm_kernels.h"

namespace hs
{

__device__ __forceinline__ void fe_shfl_up(fe &r, const fe &a, int o)
{
#pragma unroll
    for (int i = 0; i < NL; i++)
        r.v[i] = __shfl_up(a.v[i], o, 64);
}

__device__ __forceinline__ void fe_shfl_down(fe &r, const fe &a, int o)
{
#pragma unroll
    for (int i = 0; i < NL; i++)
        r.v[i] = __shfl_down(a.v[i], o, 64);
}

__device__ __forceinline__ void fe_shfl(fe &r, const fe &a, int lane)
{
#pragma unroll
    for (int i = 0; i < NL; i++)
        r.v[i] = __shfl(a.v[i], lane, 64);
}

// The inverses of x_j of every lane (j < K): Montgomery's trick over the whole wavefront.  All 64
// lanes must be active; x_j != 0.
template <class F, int K> __device__ __forceinline__ void wave_batch_inverse(fe (&inv)[K], const fe (&x)[K])
{
    const int lane = threadIdx.x & 63;
    fe q[K]; // per-lane prefix products
    q[0] = x[0];
#pragma unroll
    for (int j = 1; j < K; j++)
        fe_mul<F>(q[j], q[j - 1], x[j]);
    // inclusive prefix product of the lane totals, and the exclusive suffix product
    fe pre = q[K - 1], suf, t;
    fe_one(suf);
    fe suf_in = q[K - 1]; // inclusive suffix, built below
#pragma unroll
    for (int o = 1; o < 64; o <<= 1)
    {
        fe_shfl_up(t, pre, o);
        if (lane >= o)
            fe_mul<F>(pre, pre, t);
        fe_shfl_down(t, suf_in, o);
        if (lane + o < 64)
            fe_mul<F>(suf_in, suf_in, t);
    }
    fe excl_pre;
    fe_shfl_up(excl_pre, pre, 1);
    fe_shfl_down(suf, suf_in, 1);
    if (lane == 0)
        fe_one(excl_pre);
    if (lane == 63)
        fe_one(suf);
    fe total, itot, il;
    fe_shfl(total, pre, 63);
    fe_inv<F>(itot, total);
    // inverse of this lane's total = 1 / (prefix excluding it * own * suffix excluding it) * (the rest)
    fe_mul<F>(t, excl_pre, suf);
    fe_mul<F>(il, itot, t);
    // walk back over the lane's own K values
#pragma unroll
    for (int j = K - 1; j > 0; j--)
    {
        fe_mul<F>(inv[j], il, q[j - 1]);
        fe_mul<F>(il, il, x[j]);
    }
    inv[0] = il;
}

enum : uint32_t
{
    AFF_NONE = 0,
    AFF_ADD = 1,
    AFF_DBL = 2,
    AFF_START = 3,
    AFF_CANCEL = 4
};

// One chain: the entries [pos, end) of one chunk, summed bucket by bucket (k_bucket_flat's run
// bookkeeping on an affine sum).
struct AffChain
{
    uint32_t pos, end, g, bucket_end, run_start, t;
    bool first, live, acc_id;
    aff acc;
};

template <class C, int K>
__global__ void __launch_bounds__(64) k_bucket_affine(const uint32_t *__restrict__ offsets,
                                                      const uint32_t *__restrict__ sorted, uint32_t g_lo, uint32_t g_hi,
                                                      uint32_t Lc, const uint32_t *__restrict__ chunk_bucket,
                                                      const uint32_t *__restrict__ pts, xyzz *__restrict__ buckets,
                                                      xyzz *__restrict__ bnd)
{
    using F = typename C::F;
    const uint32_t E = offsets[g_hi], base = offsets[g_lo];
    AffChain ch[K];
#pragma unroll
    for (int j = 0; j < K; j++)
    {
        AffChain &c = ch[j];
        c.t = (blockIdx.x * 64 + threadIdx.x) * K + j;
        const uint64_t beg = (uint64_t)base + (uint64_t)c.t * Lc;
        c.live = beg < E;
        c.pos = c.live ? (uint32_t)beg : E;
        c.end = c.live ? (uint32_t)min((uint64_t)E, beg + Lc) : E;
        c.g = c.live ? chunk_bucket[c.t] : 0;
        c.bucket_end = c.live ? offsets[c.g + 1] : 0;
        c.run_start = c.pos;
        c.first = true;
        c.acc_id = true;
        fe_zero(c.acc.x);
        fe_zero(c.acc.y);
    }
    // every lane steps Lc times (chains past their end idle), so the wave stays converged for the
    // cross-lane products
    for (uint32_t s = 0; s < Lc; s++)
    {
        aff p[K];
        fe den[K], num[K];
        // case code of each chain, 3 bits each (kept in one register: as an array of K codes, the
        // Selene kernels with K >= 2 read back clobbered codes after the batch inversion)
        uint32_t kinds = 0;
#pragma unroll
        for (int j = 0; j < K; j++)
        {
            AffChain &c = ch[j];
            fe_one(den[j]);
            fe_zero(num[j]);
            fe_zero(p[j].x);
            fe_zero(p[j].y);
            if (!c.live || c.pos >= c.end)
                continue;
            const uint32_t e = sorted[c.pos];
            if (e == ENTRY_SKIP)
                continue;
            EntryBase b;
            load_entry_base(b, pts, e);
            entry_point(p[j], b);
            if (c.acc_id)
            {
                kinds |= AFF_START << (3 * j);
                continue;
            }
            fe dx, dy;
            fe_sub<F>(dx, p[j].x, c.acc.x);
            fe_sub<F>(dy, p[j].y, c.acc.y);
            if (!fe_is_zero_norm<F>(dx))
            {
                kinds |= AFF_ADD << (3 * j);
                den[j] = dx;
                num[j] = dy;
            }
            else if (fe_is_zero_norm<F>(dy))
            {
                kinds |= AFF_DBL << (3 * j); // l = 3 (x - 1)(x + 1) / 2y
                fe_add<F>(den[j], c.acc.y, c.acc.y);
                tangent_m<F>(num[j], c.acc.x, fe{{1, 0, 0, 0, 0, 0, 0, 0, 0}});
            }
            else
                kinds |= AFF_CANCEL << (3 * j);
        }
        fe inv[K];
        wave_batch_inverse<F, K>(inv, den);
#pragma unroll
        for (int j = 0; j < K; j++)
        {
            AffChain &c = ch[j];
            const uint32_t kj = (kinds >> (3 * j)) & 7u;
            if (kj == AFF_ADD || kj == AFF_DBL)
            {
                fe l, l2, x3, y3, t, x2;
                fe_mul<F>(l, num[j], inv[j]);
                fe_sq<F>(l2, l);
                fe zero;
                fe_zero(zero);
                // x3 = l^2 - x - x2 with x2 = x for a doubling (per-limb select, not a select of structs)
                const bool add = kj == AFF_ADD;
#pragma unroll
                for (int i = 0; i < NL; i++)
                    x2.v[i] = add ? p[j].x.v[i] : c.acc.x.v[i];
                fe_sub3<F>(x3, l2, c.acc.x, x2, zero);
                fe_sub<F>(t, c.acc.x, x3);
                fe_mul<F>(y3, l, t);
                fe_sub<F>(y3, y3, c.acc.y);
                c.acc.x = x3;
                c.acc.y = y3;
            }
            else if (kj == AFF_START)
            {
                c.acc = p[j];
                c.acc_id = false;
            }
            else if (kj == AFF_CANCEL)
                c.acc_id = true;
            if (!c.live || c.pos >= c.end)
                continue;
            c.pos++;
            if (c.pos == c.bucket_end || c.pos == c.end)
            {
                xyzz out;
                if (c.acc_id)
                    xyzz_identity<F>(out);
                else
                    xyzz_from_aff<F>(out, c.acc);
                const bool whole = c.run_start == offsets[c.g] && c.pos == c.bucket_end;
                if (whole)
                    store_xyzz(buckets + c.g, out);
                else
                    store_xyzz(bnd + 2 * (size_t)c.t + (c.first ? 0 : 1), out);
                c.first = false;
                c.acc_id = true;
                if (c.pos < c.end)
                {
                    do
                    {
                        c.g++;
                        c.bucket_end = offsets[c.g + 1];
                    } while (c.bucket_end == c.pos);
                    c.run_start = c.pos;
                }
            }
        }
    }
}

} // namespace hs
