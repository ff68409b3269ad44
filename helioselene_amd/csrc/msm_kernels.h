// Pippenger MSM kernels for gfx950 (one plain MSM per launch sequence).
//
// Pipeline for  Q = sum s_i P_i  (reference helios/src/x64/helios_msm_vartime.cpp:265-388):
//   ingest   radix-2^51 Jacobian points (the reference struct) -> canonical affine, identity flags
//   recode   scalars -> signed width-c digits for W = ceil(256/c) windows, with a coarse
//            (window, super-bucket) histogram per block
//   sort     two-level LDS counting sort of the (window, bucket) keys: scan of the coarse
//            histogram, coarse scatter, then one block per region orders it by bucket
//   bucket   xyzz mixed additions over fixed-size runs of the sorted entries (every thread the
//            same work), then merges of the buckets cut by run edges
//   reduce   per-window running sums over segments of buckets, then a tree over segments
//   combine  Horner over windows (c doublings per window), affine + 32-byte encoding
#pragma once
#include <hip/hip_runtime.h>

#include "curve.h"
#include "scalar.h"
#include "quad.h"
#include "rowcurve.h"

namespace hs
{

struct HeliosC
{
    using F = Fp; // coordinate field
    using O = Fq; // scalar field (group order q)
    static constexpr int id = 0;
};

struct SeleneC
{
    using F = Fq;
    using O = Fp; // group order p
    static constexpr int id = 1;
};

// ---------------------------------------------------------------- memory helpers
// Fixed-base tables hold affine points as canonical little-endian words, x then y (64 bytes per
// entry); the MSM kernels' bases use the 128-byte records below; xyzz points (buckets, partial
// sums) are raw limbs, 4 x 9 words (144 bytes = 9 x 16 bytes).
__device__ __forceinline__ void load_aff(aff &a, const uint32_t *base, uint32_t idx)
{
    const uint4 *p = reinterpret_cast<const uint4 *>(base + 16ull * idx);
    uint4 q0 = p[0], q1 = p[1], q2 = p[2], q3 = p[3];
    const uint32_t xw[8] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w};
    const uint32_t yw[8] = {q2.x, q2.y, q2.z, q2.w, q3.x, q3.y, q3.z, q3.w};
    fe_from_words(a.x, xw);
    fe_from_words(a.y, yw);
}

template <class F> __device__ __forceinline__ void store_aff(uint32_t *base, uint32_t idx, const aff &a)
{
    uint32_t xw[8], yw[8];
    fe_to_words_canon<F>(xw, a.x);
    fe_to_words_canon<F>(yw, a.y);
    uint4 *p = reinterpret_cast<uint4 *>(base + 16ull * idx);
    p[0] = make_uint4(xw[0], xw[1], xw[2], xw[3]);
    p[1] = make_uint4(xw[4], xw[5], xw[6], xw[7]);
    p[2] = make_uint4(yw[0], yw[1], yw[2], yw[3]);
    p[3] = make_uint4(yw[4], yw[5], yw[6], yw[7]);
}

// Base records of the MSM kernels: 128 bytes per point, one L2 line, so a random gather moves as
// many bytes as the 64-byte canonical form, which also arrives in one 128-byte request.  The
// record holds the canonical affine point in field.h's 29-bit limbs and also -y, so the bucket walk
// neither converts words to limbs nor negates y per entry:
//   words 0..7   x limbs 0..7          words 16..23  (-y) limbs 0..7
//   words 8..15  y limbs 0..7          words 24..27  x8, y8, x8, (-y)8      (28..31 zero)
// A negative digit reads words 16..23 and 26..27 in place of 8..15 and 24..25.  The identity is
// stored as zeros (its digits are all zero, so no entry reads it).
constexpr uint32_t BASE_RECORD_WORDS = 32;
constexpr size_t BASE_RECORD_BYTES = 4 * BASE_RECORD_WORDS;

template <class F> __device__ __forceinline__ void store_base_record(uint32_t *base, uint32_t idx, const aff &a)
{
    uint32_t xw[8], yw[8], nw[8];
    fe_to_words_canon<F>(xw, a.x);
    fe_to_words_canon<F>(yw, a.y);
    // -y = m - y, or 0 for y = 0
    uint32_t yor = 0, br = 0;
#pragma unroll
    for (int i = 0; i < 8; i++)
    {
        const uint64_t t = (uint64_t)F::mw(i) - yw[i] - br;
        nw[i] = (uint32_t)t;
        br = (uint32_t)(t >> 32) & 1u;
        yor |= yw[i];
    }
#pragma unroll
    for (int i = 0; i < 8; i++)
        nw[i] = yor ? nw[i] : 0u;
    fe x, y, ny;
    fe_from_words(x, xw);
    fe_from_words(y, yw);
    fe_from_words(ny, nw);
    uint4 *p = reinterpret_cast<uint4 *>(base + (size_t)BASE_RECORD_WORDS * idx);
    p[0] = make_uint4(x.v[0], x.v[1], x.v[2], x.v[3]);
    p[1] = make_uint4(x.v[4], x.v[5], x.v[6], x.v[7]);
    p[2] = make_uint4(y.v[0], y.v[1], y.v[2], y.v[3]);
    p[3] = make_uint4(y.v[4], y.v[5], y.v[6], y.v[7]);
    p[4] = make_uint4(ny.v[0], ny.v[1], ny.v[2], ny.v[3]);
    p[5] = make_uint4(ny.v[4], ny.v[5], ny.v[6], ny.v[7]);
    p[6] = make_uint4(x.v[8], y.v[8], x.v[8], ny.v[8]);
    p[7] = make_uint4(0u, 0u, 0u, 0u);
}

__device__ __forceinline__ void store_xyzz(xyzz *dst, const xyzz &p)
{
    uint32_t w[4 * NL];
#pragma unroll
    for (int i = 0; i < NL; i++)
    {
        w[i] = p.X.v[i];
        w[NL + i] = p.Y.v[i];
        w[2 * NL + i] = p.ZZ.v[i];
        w[3 * NL + i] = p.ZZZ.v[i];
    }
    uint4 *q = reinterpret_cast<uint4 *>(dst);
#pragma unroll
    for (int i = 0; i < NL; i++)
        q[i] = make_uint4(w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]);
}

__device__ __forceinline__ void load_xyzz(xyzz &p, const xyzz *src)
{
    uint32_t w[4 * NL];
    const uint4 *q = reinterpret_cast<const uint4 *>(src);
#pragma unroll
    for (int i = 0; i < NL; i++)
    {
        uint4 t = q[i];
        w[4 * i] = t.x;
        w[4 * i + 1] = t.y;
        w[4 * i + 2] = t.z;
        w[4 * i + 3] = t.w;
    }
#pragma unroll
    for (int i = 0; i < NL; i++)
    {
        p.X.v[i] = w[i];
        p.Y.v[i] = w[NL + i];
        p.ZZ.v[i] = w[2 * NL + i];
        p.ZZZ.v[i] = w[3 * NL + i];
    }
}

// ---------------------------------------------------------------- ingest
// Reference points: helios_jacobian = {X[5], Y[5], Z[5]} radix-2^51 uint64 limbs, Z = 0 is the
// identity (helios/include/helios_ops.h:44-64).  Output: the base record of the canonical affine
// point (X/Z^2, Y/Z^3), store_base_record.
// At most two waves per SIMD (the register allocation is padded to 169 VGPRs): with several MSMs
// in flight, a full-occupancy ingest (107 VGPRs, four waves) crowds the CUs the other MSMs' bucket
// kernels run on; four-stream throughput measured 650 M points/s against 720-730 M with the cap,
// on the same box, with identical kernel code.
template <class C>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 2)))
k_ingest(const uint64_t *__restrict__ jac, uint32_t n, uint32_t *__restrict__ out, uint8_t *__restrict__ is_id)
{
    using F = typename C::F;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint64_t *p = jac + 15ull * i;
    uint64_t l[5];
    fe X, Y, Z;
#pragma unroll
    for (int k = 0; k < 5; k++)
        l[k] = p[k];
    fe_from_r51<F>(X, l);
#pragma unroll
    for (int k = 0; k < 5; k++)
        l[k] = p[5 + k];
    fe_from_r51<F>(Y, l);
#pragma unroll
    for (int k = 0; k < 5; k++)
        l[k] = p[10 + k];
    fe_from_r51<F>(Z, l);
    aff a;
    uint32_t Zc[8];
    fe_to_words_canon<F>(Zc, Z);
    uint32_t zor = 0, one = Zc[0] == 1u;
#pragma unroll
    for (int k = 0; k < 8; k++)
        zor |= Zc[k];
#pragma unroll
    for (int k = 1; k < 8; k++)
        one &= Zc[k] == 0u;
    if (zor == 0)
    {
        fe_zero(a.x);
        fe_zero(a.y);
        is_id[i] = 1;
    }
    else
    {
        if (one)
        {
            a.x = X;
            a.y = Y;
        }
        else
        {
            fe zi, zi2, zi3;
            fe_inv<F>(zi, Z);
            fe_sq<F>(zi2, zi);
            fe_mul<F>(zi3, zi2, zi);
            fe_mul<F>(a.x, X, zi2);
            fe_mul<F>(a.y, Y, zi3);
        }
        is_id[i] = 0;
    }
    store_base_record<F>(out, i, a);
}

// ---------------------------------------------------------------- exclusive scan (uint32)
// Three-phase scan: per-block scan of 2048 elements -> block sums -> add offsets.
constexpr int SCAN_BLOCK = 1024, SCAN_ITEMS = 2;

// Exclusive scan of one value per thread across the block (blockDim.x a multiple of 64, <= 1024).
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v)
{
    __shared__ uint32_t wsum[SCAN_BLOCK / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1)
    {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o)
            x += y;
    }
    if (lane == 63)
        wsum[wid] = x;
    __syncthreads();
    if (wid == 0)
    {
        uint32_t t = lane < nw ? wsum[lane] : 0u;
        uint32_t u = t;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1)
        {
            uint32_t y = __shfl_up(u, o, 64);
            if (lane >= o)
                u += y;
        }
        if (lane < nw)
            wsum[lane] = u - t;
    }
    __syncthreads();
    uint32_t r = x - v + wsum[wid];
    __syncthreads(); // wsum is reused by the next call
    return r;
}

__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_local(const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
                                                           uint32_t len, uint32_t *__restrict__ block_sums)
{
    __shared__ uint32_t tot;
    const uint32_t base = blockIdx.x * (SCAN_BLOCK * SCAN_ITEMS) + threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS], s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++)
    {
        v[k] = base + k < len ? in[base + k] : 0u;
        s += v[k];
    }
    uint32_t ex = block_exclusive_scan(s);
    if (threadIdx.x == blockDim.x - 1)
        tot = ex + s;
    __syncthreads();
    uint32_t run = ex;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++)
    {
        if (base + k < len)
            out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0)
        block_sums[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_add(uint32_t *__restrict__ out, uint32_t len,
                                                         const uint32_t *__restrict__ block_offsets)
{
    const uint32_t base = blockIdx.x * (SCAN_BLOCK * SCAN_ITEMS) + threadIdx.x * SCAN_ITEMS;
    const uint32_t off = block_offsets[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++)
        if (base + k < len)
            out[base + k] += off;
}

// ---------------------------------------------------------------- two-level counting sort
// Digits are stored as 16-bit codes, window-major: code = (|d| - 1) | (d < 0) << 15, and
// DIGIT_ZERO for a zero digit (that code would mean digit -2^15, which never occurs because
// digits lie in [-(2^(c-1)) + 1, 2^(c-1)] with c <= 16).
constexpr uint16_t DIGIT_ZERO = 0xffffu;

// Workgroup sizes of the sort passes.  The LDS-atomic loops are latency-bound and want many waves
// per workgroup: 1024-thread coarse workgroups (one per CU) measured fastest alone and with two
// MSMs per curve in flight; with exactly one MSM per curve in flight 512 threads did better.
#ifndef HS_SORT_COARSE_THREADS
#define HS_SORT_COARSE_THREADS 1024
#endif
#ifndef HS_SORT_FINE_THREADS
#define HS_SORT_FINE_THREADS 512
#endif
constexpr int SORT_COARSE_THREADS = HS_SORT_COARSE_THREADS;
constexpr int SORT_FINE_THREADS = HS_SORT_FINE_THREADS;

// Recode fused with the coarse count.  Block b recodes the points [b*T, b*T + T): the digit of
// window w of point i goes to digits[w * n + i], and the block counts its non-zero digits per
// (window, super-bucket), super-bucket = bucket >> shift, into counts[(w*nsb + sb) * nblk + b].
template <class C>
__global__ void __launch_bounds__(SORT_COARSE_THREADS) k_recode_count(
    const uint32_t *__restrict__ scalars, const uint8_t *__restrict__ is_id, uint32_t n, int c, int W, int w8_compat,
    int nsb, int shift, uint32_t T, uint32_t nblk, uint16_t *__restrict__ digits, uint32_t *__restrict__ counts)
{
    extern __shared__ uint32_t hist[];
    const uint32_t b = blockIdx.x, beg = b * T, end = min(beg + T, n), nh = (uint32_t)W * nsb;
    for (uint32_t j = threadIdx.x; j < nh; j += blockDim.x)
        hist[j] = 0;
    if (b == 0 && threadIdx.x == 0)
        counts[(size_t)nh * nblk] = 0; // the element past the last count: the scan's total
    __syncthreads();
    // the next point's scalar and identity flag are loaded while this point is recoded
    uint4 a = make_uint4(0, 0, 0, 0), bb = a;
    uint8_t idf = 0;
    if (beg + threadIdx.x < end)
    {
        const uint4 *sp = reinterpret_cast<const uint4 *>(scalars + 8ull * (beg + threadIdx.x));
        a = sp[0];
        bb = sp[1];
        idf = is_id[beg + threadIdx.x];
    }
    for (uint32_t i = beg + threadIdx.x; i < end; i += blockDim.x)
    {
        uint32_t s[8];
        s[0] = a.x; s[1] = a.y; s[2] = a.z; s[3] = a.w;
        s[4] = bb.x; s[5] = bb.y; s[6] = bb.z; s[7] = bb.w;
        const bool skip = idf != 0;
        if (i + blockDim.x < end)
        {
            const uint4 *sp = reinterpret_cast<const uint4 *>(scalars + 8ull * (i + blockDim.x));
            a = sp[0];
            bb = sp[1];
            idf = is_id[i + blockDim.x];
        }
        uint32_t k[8];
        scalar_prepare<typename C::O>(k, s, w8_compat != 0);
        uint32_t carry = 0;
        for (int w = 0; w < W; w++)
        {
            const uint32_t d = recode_digit(k, c, w, carry);
            const uint32_t mag1 = (d & 0x7fffffffu) - 1;
            const bool zero = d == 0 || skip;
            digits[(size_t)w * n + i] = zero ? DIGIT_ZERO : (uint16_t)(mag1 | ((d >> 31) << 15));
            if (!zero)
                atomicAdd(&hist[(uint32_t)w * nsb + (mag1 >> shift)], 1u);
        }
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < nh; j += blockDim.x)
        counts[(size_t)j * nblk + b] = hist[j];
}

// Coarse scatter: entry = point index (bits 0..23) | sub-bucket << 24 | sign << 31
__global__ void __launch_bounds__(SORT_COARSE_THREADS) k_coarse_scatter(const uint16_t *__restrict__ digits, uint32_t n, int W,
                                                        int nsb, int shift, uint32_t T, uint32_t nblk,
                                                        const uint32_t *__restrict__ offs,
                                                        uint32_t *__restrict__ entries)
{
    extern __shared__ uint32_t cur[];
    const uint32_t b = blockIdx.x, beg = b * T, cnt = min(T, n - beg), nh = (uint32_t)W * nsb;
    for (uint32_t j = threadIdx.x; j < nh; j += blockDim.x)
        cur[j] = offs[(size_t)j * nblk + b];
    __syncthreads();
    const uint32_t smask = (1u << shift) - 1;
    for (int w = 0; w < W; w++)
    {
        const uint16_t *dw = digits + (size_t)w * n;
        uint32_t *cw = cur + (uint32_t)w * nsb;
        for (uint32_t i = beg + threadIdx.x; i < beg + cnt; i += blockDim.x)
        {
            const uint16_t code = dw[i];
            if (code != DIGIT_ZERO)
            {
                const uint32_t mag1 = code & 0x7fffu;
                const uint32_t pos = atomicAdd(&cw[mag1 >> shift], 1u);
                entries[pos] = i | ((mag1 & smask) << 24) | ((uint32_t)(code >> 15) << 31);
            }
        }
    }
}

// Fine pass: one block per (window, super-bucket) region; orders the region by bucket, writes
// the final entries (point index | sign << 31) and the bucket offsets of the region.
__global__ void __launch_bounds__(SORT_FINE_THREADS) k_fine_sort(const uint32_t *__restrict__ entries, const uint32_t *__restrict__ offs,
                                                   uint32_t nblk, int nsb, int shift, uint32_t B, uint32_t nregions,
                                                   uint32_t *__restrict__ sorted, uint32_t *__restrict__ offsets)
{
    __shared__ uint32_t hist[128];
    __shared__ uint32_t cur[128];
    const uint32_t r = blockIdx.x;
    const uint32_t start = offs[(size_t)r * nblk], end = offs[(size_t)(r + 1) * nblk];
    const uint32_t nsub = 1u << shift;
    for (uint32_t j = threadIdx.x; j < nsub; j += blockDim.x)
        hist[j] = 0;
    __syncthreads();
    for (uint32_t e = start + threadIdx.x; e < end; e += blockDim.x)
        atomicAdd(&hist[(entries[e] >> 24) & 0x7fu], 1u);
    __syncthreads();
    if (threadIdx.x < 64)
    {
        // exclusive scan of up to 128 counters by one wave (two per lane)
        const uint32_t l = threadIdx.x;
        uint32_t a0 = 2 * l < nsub ? hist[2 * l] : 0u, a1 = 2 * l + 1 < nsub ? hist[2 * l + 1] : 0u;
        uint32_t x = a0 + a1;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1)
        {
            uint32_t y = __shfl_up(x, o, 64);
            if ((int)l >= o)
                x += y;
        }
        const uint32_t ex = x - a0 - a1;
        if (2 * l < nsub)
            cur[2 * l] = start + ex;
        if (2 * l + 1 < nsub)
            cur[2 * l + 1] = start + ex + a0;
    }
    __syncthreads();
    const uint32_t w = r / nsb, sb = r % nsb;
    for (uint32_t j = threadIdx.x; j < nsub; j += blockDim.x)
        offsets[(size_t)w * B + sb * nsub + j] = cur[j];
    if (r == nregions - 1 && threadIdx.x == 0)
        offsets[(size_t)(w + 1) * B] = end;
    __syncthreads();
    for (uint32_t e = start + threadIdx.x; e < end; e += blockDim.x)
    {
        const uint32_t v = entries[e];
        const uint32_t pos = atomicAdd(&cur[(v >> 24) & 0x7fu], 1u);
        sorted[pos] = (v & 0x00ffffffu) | (v & 0x80000000u);
    }
}

// ---------------------------------------------------------------- bucket sums
constexpr uint32_t ENTRY_SKIP = 0xffffffffu; // padding entry of the batched layout

// The base of an entry as the bucket walk needs it: x and, by the entry's sign, y or -y (18
// limbs) from its record, in flight one entry ahead so the gather overlaps the current addition.
// Padding reads point 0, which the walk then skips.
struct EntryBase
{
    uint4 x0, x1, y0, y1;
    uint2 t; // limb 8 of x and of y or -y
};

__device__ __forceinline__ void load_entry_base(EntryBase &b, const uint32_t *__restrict__ pts, uint32_t e)
{
    const uint32_t idx = e == ENTRY_SKIP ? 0u : (e & 0x7fffffffu);
    const uint32_t neg = e == ENTRY_SKIP ? 0u : e >> 31;
    const uint32_t *r = pts + (size_t)BASE_RECORD_WORDS * idx;
    b.x0 = *reinterpret_cast<const uint4 *>(r);
    b.x1 = *reinterpret_cast<const uint4 *>(r + 4);
    b.y0 = *reinterpret_cast<const uint4 *>(r + 8 + 8 * neg);
    b.y1 = *reinterpret_cast<const uint4 *>(r + 12 + 8 * neg);
    b.t = *reinterpret_cast<const uint2 *>(r + 24 + 2 * neg);
}

__device__ __forceinline__ void entry_point(aff &p, const EntryBase &b)
{
    p.x.v[0] = b.x0.x; p.x.v[1] = b.x0.y; p.x.v[2] = b.x0.z; p.x.v[3] = b.x0.w;
    p.x.v[4] = b.x1.x; p.x.v[5] = b.x1.y; p.x.v[6] = b.x1.z; p.x.v[7] = b.x1.w;
    p.x.v[8] = b.t.x;
    p.y.v[0] = b.y0.x; p.y.v[1] = b.y0.y; p.y.v[2] = b.y0.z; p.y.v[3] = b.y0.w;
    p.y.v[4] = b.y1.x; p.y.v[5] = b.y1.y; p.y.v[6] = b.y1.z; p.y.v[7] = b.y1.w;
    p.y.v[8] = b.t.y;
}

// acc += the entry's base (its sign already applied by the load).  The caller passes no padding.
template <class C> __device__ __forceinline__ void add_entry(xyzz &acc, bool &acc_id, const EntryBase &b)
{
    aff p;
    entry_point(p, b);
    xyzz_madd_hot<typename C::F>(acc, acc_id, p);
}

constexpr int BIG_THREADS = 256;

// LDS tree sum of one xyzz per thread; result in thread 0.  sh holds BIG_THREADS points.
template <class C> __device__ void block_tree_sum(xyzz &acc, xyzz *sh)
{
    using F = typename C::F;
    const int t = threadIdx.x;
    store_xyzz(sh + t, acc);
    __syncthreads();
    for (int s = blockDim.x >> 1; s > 0; s >>= 1)
    {
        if (t < s)
        {
            xyzz a, b;
            load_xyzz(a, sh + t);
            load_xyzz(b, sh + t + s);
            xyzz_add<F>(a, a, b);
            store_xyzz(sh + t, a);
        }
        __syncthreads();
    }
    if (t == 0)
        load_xyzz(acc, sh);
}

// ---------------------------------------------------------------- flat, balanced bucket sums
// The sorted entries of all windows form one array grouped by global bucket id g = w*B + b.
// Thread t sums exactly the entries [t*L, t*L + L) (every thread does the same work, whatever
// the bucket sizes), walking across bucket boundaries:
//   * a run that covers a whole bucket is that bucket's sum and is written to buckets[g];
//   * a run cut by the chunk start or end is a partial sum, written to boundary slot 2t (the
//     chunk's first run) or 2t+1 (its last run, when different from the first).
// A bucket whose entries span chunks t0 < t1 then equals slot (2 t0 + 1, or 2 t0 when it starts
// exactly at chunk t0's start) plus slots 2t for t0 < t <= t1; k_merge_span adds them up.

// The kernels below work on the buckets [g_lo, g_hi) of a window range; chunk t covers the entries
// base + [t*L, t*L + L), base = offsets[g_lo].
// chunk_bucket[t] = the bucket holding entry base + t*L; the first thread also clears *zero_word
// (the merge's count of long spans) when given
__global__ void __launch_bounds__(256) k_chunk_start(const uint32_t *__restrict__ offsets, uint32_t g_lo, uint32_t g_hi,
                                                     uint32_t L, uint32_t *__restrict__ chunk_bucket,
                                                     uint32_t *__restrict__ zero_word)
{
    if (zero_word && blockIdx.x == 0 && threadIdx.x == 0)
        *zero_word = 0;
    uint32_t g = g_lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= g_hi)
        return;
    const uint32_t base = offsets[g_lo];
    const uint32_t a = offsets[g] - base, e = offsets[g + 1] - base;
    for (uint32_t t = (a + L - 1) / L; t * L < e; t++)
        chunk_bucket[t] = g;
}

// A run of one bucket starts from its first entry's base (sign applied); a padding entry starts
// it as the identity.
template <class F> __device__ __forceinline__ void run_start(xyzz &acc, bool &acc_id, const EntryBase &b, uint32_t e)
{
    aff p;
    entry_point(p, b);
    xyzz_from_aff<F>(acc, p);
    acc_id = e == ENTRY_SKIP;
}

// State of one thread's walk over its chunk [pos, end) of sorted entries (k_bucket_flat).
struct FlatWalk
{
    uint32_t pos, end, g, g_hi, bucket_end, next_end, t, e_next;
    bool first, first_whole;
};

// One step of the walk: entry `pos` (index e, record in `cur`) either extends the current run by
// a mixed addition or, when the run has ended at pos, the run is stored and a new one starts with
// it.  The record of entry pos + 1 is issued into `nxt` first, so its gather overlaps the
// addition.  Returns false once the chunk is done (the last run stored).
template <class C>
__device__ __forceinline__ bool flat_step(FlatWalk &f, xyzz &acc, bool &acc_id, uint32_t &e, const EntryBase &cur,
                                          EntryBase &nxt, const uint32_t *__restrict__ offsets,
                                          const uint32_t *__restrict__ sorted, const uint32_t *__restrict__ pts,
                                          xyzz *__restrict__ buckets, xyzz *__restrict__ bnd)
{
    using F = typename C::F;
    const uint32_t e_cur = e;
    if (f.pos + 1 < f.end)
    {
        e = f.e_next;
        load_entry_base(nxt, pts, e);
        if (f.pos + 2 < f.end)
            f.e_next = sorted[f.pos + 2];
    }
    const bool run_done = f.pos == f.bucket_end || f.pos == f.end;
    // padding entries (batched layouts put them at the end of a window's last bucket) add nothing
    if (run_done || e_cur == ENTRY_SKIP)
    {
        if (!run_done)
        {
            f.pos++;
            return true;
        }
        if (acc_id)
            xyzz_identity<F>(acc);
        // later runs start at their bucket's first entry, so only the first run can be partial there
        const bool whole = f.pos == f.bucket_end && (!f.first || f.first_whole);
        if (whole)
            store_xyzz(buckets + f.g, acc);
        else
            store_xyzz(bnd + 2 * (size_t)f.t + (f.first ? 0 : 1), acc);
        if (f.pos == f.end)
            return false;
        f.first = false;
        // next non-empty bucket (entries are contiguous, so it starts at pos); its end was
        // prefetched, so only empty buckets cost a dependent load here
        f.g++;
        f.bucket_end = f.next_end;
        while (f.bucket_end == f.pos)
        {
            f.g++;
            f.bucket_end = offsets[f.g + 1];
        }
        f.next_end = f.g + 2 <= f.g_hi ? offsets[f.g + 2] : f.end;
        run_start<F>(acc, acc_id, cur, e_cur);
    }
    else
        add_entry<C>(acc, acc_id, cur);
    f.pos++;
    return true;
}

template <class C>
__global__ void __launch_bounds__(256) k_bucket_flat(const uint32_t *__restrict__ offsets,
                                                     const uint32_t *__restrict__ sorted, uint32_t g_lo, uint32_t g_hi,
                                                     uint32_t L,
                                                     const uint32_t *__restrict__ chunk_bucket,
                                                     const uint32_t *__restrict__ pts, xyzz *__restrict__ buckets,
                                                     xyzz *__restrict__ bnd)
{
    using F = typename C::F;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t E = offsets[g_hi], beg = offsets[g_lo] + t * L;
    if (beg >= E)
        return;
    FlatWalk f;
    f.t = t;
    f.end = min(E, beg + L);
    f.g_hi = g_hi;
    f.g = chunk_bucket[t];
    f.bucket_end = offsets[f.g + 1];
    f.next_end = f.g + 2 <= g_hi ? offsets[f.g + 2] : f.end;
    f.first = true;
    f.first_whole = offsets[f.g] == beg;
    // the first entry starts the first run; the walk then begins at beg + 1
    EntryBase qa, qb;
    uint32_t e = sorted[beg];
    load_entry_base(qa, pts, e);
    f.e_next = beg + 1 < f.end ? sorted[beg + 1] : ENTRY_SKIP;
    xyzz acc;
    bool acc_id;
    run_start<F>(acc, acc_id, qa, e);
    f.pos = beg + 1;
    // entry beg + 1's record goes to qb; from then on the two buffers alternate (unrolled by two,
    // so the in-flight record never needs a register copy)
    if (f.pos < f.end)
    {
        e = f.e_next;
        load_entry_base(qb, pts, e);
        if (f.pos + 1 < f.end)
            f.e_next = sorted[f.pos + 1];
    }
    while (true)
    {
        if (!flat_step<C>(f, acc, acc_id, e, qb, qa, offsets, sorted, pts, buckets, bnd))
            break;
        if (!flat_step<C>(f, acc, acc_id, e, qa, qb, offsets, sorted, pts, buckets, bnd))
            break;
    }
}

// Completes buckets whose entries span several chunks; spans longer than `short_max` chunks
// are queued for k_merge_long.  One thread per chunk t (entries base + [t L, t L + L), base =
// offsets[g_lo], nchunks = ceil((offsets[g_hi] - base) / L)): the thread of the chunk in which a
// cut bucket ends adds it up, so every cut bucket has exactly one thread and there are about
// half as many threads as buckets (a bucket per thread left most lanes idle).
template <class C>
__global__ void __launch_bounds__(256) k_merge_span(const uint32_t *__restrict__ offsets, uint32_t g_lo, uint32_t g_hi,
                                                    uint32_t L, const uint32_t *__restrict__ chunk_bucket,
                                                    const xyzz *__restrict__ bnd, uint32_t short_max,
                                                    xyzz *__restrict__ buckets, uint32_t *__restrict__ long_list,
                                                    uint32_t *__restrict__ long_count)
{
    using F = typename C::F;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t base = offsets[g_lo], E = offsets[g_hi] - base;
    if ((uint64_t)t * L >= E)
        return;
    // the bucket holding the chunk's first entry: cut here if it started in an earlier chunk, and
    // completed here if it ends in this chunk
    const uint32_t g = chunk_bucket[t];
    const uint32_t a = offsets[g] - base, e = offsets[g + 1] - base;
    if (a == t * L || (uint64_t)e > (uint64_t)t * L + L)
        return;
    const uint32_t t0 = a / L, t1 = t;
    if (t1 - t0 + 1 > short_max)
    {
        long_list[atomicAdd(long_count, 1u)] = g;
        return;
    }
    xyzz acc, p;
    load_xyzz(acc, bnd + 2 * (size_t)t0 + (a % L ? 1 : 0));
    for (uint32_t u = t0 + 1; u <= t1; u++)
    {
        load_xyzz(p, bnd + 2 * (size_t)u);
        xyzz_add<F>(acc, acc, p);
    }
    store_xyzz(buckets + g, acc);
}

template <class C>
__global__ void __launch_bounds__(BIG_THREADS) k_merge_long(const uint32_t *__restrict__ offsets, uint32_t g_lo, uint32_t L,
                                                            const xyzz *__restrict__ bnd,
                                                            const uint32_t *__restrict__ long_list,
                                                            const uint32_t *__restrict__ long_count,
                                                            xyzz *__restrict__ buckets)
{
    using F = typename C::F;
    __shared__ xyzz sh[BIG_THREADS];
    const uint32_t nlong = *long_count;
    for (uint32_t k = blockIdx.x; k < nlong; k += gridDim.x)
    {
        const uint32_t g = long_list[k];
        const uint32_t base = offsets[g_lo];
        const uint32_t a = offsets[g] - base, e = offsets[g + 1] - base;
        const uint32_t t0 = a / L, t1 = (e - 1) / L, cnt = t1 - t0 + 1;
        xyzz acc, p;
        xyzz_identity<F>(acc);
        for (uint32_t j = threadIdx.x; j < cnt; j += blockDim.x)
        {
            const size_t slot = j == 0 ? 2 * (size_t)t0 + (a % L ? 1 : 0) : 2 * (size_t)(t0 + j);
            load_xyzz(p, bnd + slot);
            xyzz_add<F>(acc, acc, p);
        }
        block_tree_sum<C>(acc, sh);
        if (threadIdx.x == 0)
            store_xyzz(buckets + g, acc);
        __syncthreads();
    }
}

// ---------------------------------------------------------------- bucket reduction
// Window w's sum is sum_b (b + 1) B_b over its B buckets (reference running-sum reduction,
// helios_msm_vartime.cpp:334-361).  Split the buckets into segments of G: segment s holds buckets
// [s G, s G + G), and with R_s = sum_j B_(sG+j) and T_s = sum_j (j + 1) B_(sG+j) (a running sum,
// one thread per segment)
//     window sum = sum_s T_s + G sum_s s R_s.
// sum_s s R_s = sum_i 2^i Y_i, where Y_i sums the R_s whose index s has bit i set.  These subset
// sums come out of a binary tree over the segments that keeps, besides each node's total, the
// totals of its right children at every level; no scalar multiple of any partial sum is formed,
// and the only doublings left are those of the final sum_i 2^(i + log2 G) Y_i, one short chain
// per window.
//   k_reduce_seg: one thread per segment; a workgroup reduces tw = min(segments per window, 256)
//                 consecutive segments of one window (256 / tw windows per workgroup) and leaves
//                 2 + log2 tw points per block: S = sum R, T = sum T, Q_L = sum of the R_s with
//                 bit L of s (within the block) set.
//   k_reduce_win: one wavefront per window runs the same tree over the window's blocks (the
//                 blocks' own Q_L add up; the block index's bits add subset sums of S), then
//                 forms T + sum_i 2^(i + log2 G) Y_i with every field element spread over a DPP
//                 row (rowcurve.h) and stores the window sum in Jacobian form.
constexpr uint32_t SEG_SLOTS = 10;        // S, T and Q_0..Q_7 of a block of up to 256 segments
constexpr uint32_t MAX_SEGS_PER_WIN = 8192; // k_reduce_win keeps (2 + 8) x 32 blocks in LDS

__host__ __device__ inline uint32_t ilog2u(uint32_t x)
{
    uint32_t r = 0;
    while (x >>= 1)
        r++;
    return r;
}

// segments per k_reduce_seg block
__host__ __device__ inline uint32_t seg_block(uint32_t segs_per_win) { return segs_per_win < 256u ? segs_per_win : 256u; }

// xyzz entries of k_reduce_seg's output for nwin windows
__host__ __device__ inline size_t seg_out_entries(size_t nwin, uint32_t segs_per_win)
{
    return nwin * (segs_per_win / seg_block(segs_per_win)) * SEG_SLOTS;
}

// LDS of one window's wavefront in k_reduce_win
__host__ __device__ inline size_t reduce_win_lds1(uint32_t segs_per_win)
{
    const uint32_t tw = seg_block(segs_per_win), np = segs_per_win / tw, lt = ilog2u(tw), lp = ilog2u(np);
    return (size_t)(2 + lt) * np * sizeof(xyzz) + (size_t)(lt + lp + 1) * sizeof(jac);
}

// dynamic LDS of k_reduce_win with wg windows per workgroup: each wavefront's own area, then one
// Jacobian exchange slot per wavefront
__host__ __device__ inline size_t reduce_win_lds(uint32_t segs_per_win, uint32_t wg = 1)
{
    return ((reduce_win_lds1(segs_per_win) + 15) & ~(size_t)15) * wg + (size_t)wg * sizeof(jac);
}

// Windows per group of the window combine: a workgroup of wg wavefronts reduces wg consecutive
// windows at once and adds them up as sum_j 2^(c j) S_(wg k + j).  The serial Horner over the
// groups then has ceil(W / wg) - 1 additions in place of W - 1 (a Jacobian addition costs about
// two doublings); its doublings, c (W - wg), plus the c (wg - 1) of the longest scaling beside it
// are the c (W - 1) of the plain Horner, and the group's own log2 wg additions run on the side.
// Four windows per group up to c = 13 (W >= 20) where the workgroup's LDS fits in 64 KiB, else
// two, else one.  At c = 16 (W = 16) two windows per group measured no faster on MI355X: the
// Horner's 7 fewer additions (-20 us Helios, -37 us Selene) went into the longer window
// reduction (+19 / +34 us).
__host__ __device__ inline uint32_t win_group(int c, uint32_t segs_per_win)
{
    uint32_t wg = c <= 13 ? 4u : 1u;
    while (wg > 1 && reduce_win_lds(segs_per_win, wg) > 64u * 1024u)
        wg >>= 1;
    return wg;
}

template <class F> __device__ __forceinline__ xyzz xyzz_sel(bool c, const xyzz &a, const xyzz &b)
{
    xyzz r;
#pragma unroll
    for (int i = 0; i < NL; i++)
    {
        r.X.v[i] = c ? a.X.v[i] : b.X.v[i];
        r.Y.v[i] = c ? a.Y.v[i] : b.Y.v[i];
        r.ZZ.v[i] = c ? a.ZZ.v[i] : b.ZZ.v[i];
        r.ZZZ.v[i] = c ? a.ZZZ.v[i] : b.ZZZ.v[i];
    }
    return r;
}

__device__ __forceinline__ xyzz xyzz_shfl_xor1(const xyzz &a)
{
    xyzz r;
#pragma unroll
    for (int i = 0; i < NL; i++)
    {
        r.X.v[i] = __shfl_xor(a.X.v[i], 1, 64);
        r.Y.v[i] = __shfl_xor(a.Y.v[i], 1, 64);
        r.ZZ.v[i] = __shfl_xor(a.ZZ.v[i], 1, 64);
        r.ZZZ.v[i] = __shfl_xor(a.ZZZ.v[i], 1, 64);
    }
    return r;
}

// Thread t of the launch handles global segment w_lo * (B / G) + t (nseg segments in all).
// Tree step s >= 1 combines, for each list, the entries at P and P + 2^s of one array (P = k 2^(s+1)
// plus the list's offset): list 0 is S (offset 0), list 1 is T (stored at half the index, its
// step 0 being done across neighbouring lanes), list 2 + L (L < s) the Q_L entries of S, which sit
// at offset 2^L: the right children of level L, left untouched by the levels above.  Within a step
// every write lands at an offset below 2^s (mod 2^(s+1)) and every partner read at one above, so
// one barrier per step suffices; each thread does at most one addition per step.
template <class C>
__global__ void __launch_bounds__(256) k_reduce_seg(const xyzz *__restrict__ buckets, const uint32_t *__restrict__ offsets,
                                                    uint32_t B, uint32_t G, uint32_t w_lo, uint32_t nseg,
                                                    xyzz *__restrict__ seg_out)
{
    using F = typename C::F;
    __shared__ xyzz shS[256];
    __shared__ xyzz shT[128];
    const uint32_t segs_per_win = B / G, tw = seg_block(segs_per_win), lt = ilog2u(tw);
    const uint32_t t0 = w_lo * segs_per_win, i = threadIdx.x;
    const uint32_t t = t0 + blockIdx.x * 256 + i;
    const bool live = t < t0 + nseg;
    xyzz R, T;
    xyzz_identity<F>(R);
    xyzz_identity<F>(T);
    if (live)
    {
        const uint32_t w = t / segs_per_win, sg = t % segs_per_win;
        const size_t g0 = (size_t)w * B + (size_t)sg * G;
        for (int j = (int)G - 1; j >= 0; j--)
        {
            if (offsets[g0 + j] != offsets[g0 + j + 1]) // empty buckets hold the identity
            {
                xyzz b;
                load_xyzz(b, buckets + g0 + j);
                xyzz_add<F>(R, R, b);
            }
            xyzz_add<F>(T, T, R);
        }
    }
    if (lt == 0)
    {
        // one segment per window: the block is the segment itself
        if (live)
        {
            store_xyzz(seg_out + (size_t)t * SEG_SLOTS, R);
            store_xyzz(seg_out + (size_t)t * SEG_SLOTS + 1, T);
        }
        return;
    }
    // step 0 across lanes 2k, 2k + 1: the even lane forms S = R_2k + R_2k+1, the odd lane
    // T_2k + T_2k+1 and keeps R_2k+1 (the Q_0 entry) in place
    const bool odd = (i & 1u) != 0;
    const xyzz other = xyzz_shfl_xor1(xyzz_sel<F>(odd, R, T));
    xyzz sum = xyzz_sel<F>(odd, T, R);
    xyzz_add<F>(sum, sum, other);
    store_xyzz(shS + i, xyzz_sel<F>(odd, R, sum));
    if (odd)
        store_xyzz(shT + (i >> 1), sum);
    __syncthreads();
    for (uint32_t s = 1; s < lt; s++)
    {
        const uint32_t nk = 128u >> s;
        if (i < (2 + s) * nk)
        {
            const uint32_t lst = i / nk, P = (i % nk) << (s + 1);
            xyzz *dst;
            uint32_t h;
            if (lst == 1)
            {
                dst = shT + (P >> 1);
                h = 1u << (s - 1);
            }
            else
            {
                dst = shS + (lst == 0 ? P : P + (1u << (lst - 2)));
                h = 1u << s;
            }
            xyzz a, b;
            load_xyzz(a, dst);
            load_xyzz(b, dst + h);
            xyzz_add<F>(a, a, b);
            store_xyzz(dst, a);
        }
        __syncthreads();
    }
    const uint32_t nslot = 2 + lt, nblk = 256 / tw;
    for (uint32_t x = i; x < nblk * nslot; x += 256)
    {
        const uint32_t j = x / nslot, slot = x % nslot, base = blockIdx.x * 256 + j * tw;
        if (base >= nseg)
            continue;
        const xyzz *src = slot == 0 ? shS + j * tw : slot == 1 ? shT + (j * tw >> 1) : shS + j * tw + (1u << (slot - 2));
        xyzz v;
        load_xyzz(v, src);
        store_xyzz(seg_out + (size_t)((t0 + base) / tw) * SEG_SLOTS + slot, v);
    }
}

// Workgroups of wg wavefronts (64 threads each; batches have thousands of small windows, whose
// workgroups must not hold idle waves).  Workgroup b handles group gi = b % groups of chain
// ch = b / groups: wavefront j reduces window wc = gi wg + j of the chain (global window
// w_lo + ch Wc + wc; a wavefront past the chain's Wc windows only joins the barriers), scales its
// sum by 2^(c j) and the group's sums are added by an LDS tree; the result, in Jacobian form,
// goes to win_out[w_lo + ch groups + gi].  Dynamic LDS: reduce_win_lds(segs_per_win, wg).
template <class C>
__global__ void __launch_bounds__(256) k_reduce_win(const xyzz *__restrict__ seg_out, uint32_t segs_per_win, uint32_t G,
                                                     uint32_t w_lo, uint32_t Wc, uint32_t groups, int c,
                                                     jac *__restrict__ win_out)
{
    using F = typename C::F;
    extern __shared__ uint4 lds_win[];
    const uint32_t tw = seg_block(segs_per_win), lt = ilog2u(tw), np = segs_per_win / tw, lp = ilog2u(np);
    const uint32_t nslot = 2 + lt, nY = lt + lp, i = threadIdx.x & 63u, j = threadIdx.x >> 6, wg = blockDim.x >> 6;
    const uint32_t ch = blockIdx.x / groups, gi = blockIdx.x % groups, wc = gi * wg + j;
    const bool live = wc < Wc;
    const uint32_t w = w_lo + ch * Wc + wc;
    const size_t area = (reduce_win_lds1(segs_per_win) + 15) & ~(size_t)15;
    xyzz *A = reinterpret_cast<xyzz *>(reinterpret_cast<char *>(lds_win) + area * j); // slot-major: A[slot * np + block]
    jac *J = reinterpret_cast<jac *>(A + (size_t)nslot * np);
    jac *X = reinterpret_cast<jac *>(reinterpret_cast<char *>(lds_win) + area * wg); // exchange slots
    if (live)
        for (uint32_t x = i; x < nslot * np; x += 64)
        {
            const uint32_t k = x / nslot, slot = x % nslot;
            xyzz v;
            load_xyzz(v, seg_out + ((size_t)w * np + k) * SEG_SLOTS + slot);
            store_xyzz(A + (size_t)slot * np + k, v);
        }
    __syncthreads();
    // the tree over the window's blocks: every slot array adds up, and the bits of the block index
    // give new subset sums of S at offsets 2^L
    for (uint32_t s = 0; s < lp; s++)
    {
        const uint32_t nk = np >> (s + 1);
        if (live)
            for (uint32_t x = i; x < (nslot + s) * nk; x += 64)
            {
                const uint32_t lst = x / nk, P = (x % nk) << (s + 1);
                xyzz *dst = lst < nslot ? A + (size_t)lst * np + P : A + P + (1u << (lst - nslot));
                xyzz a, b;
                load_xyzz(a, dst);
                load_xyzz(b, dst + (1u << s));
                xyzz_add<F>(a, a, b);
                store_xyzz(dst, a);
            }
        __syncthreads();
    }
    // Y_i (i < lt): the blocks' Q_i summed; Y_(lt + L): S subset sums; J[nY] = T
    if (live && i <= nY)
    {
        const xyzz *src = i == nY ? A + np : i < lt ? A + (size_t)(2 + i) * np : A + (1u << (i - lt));
        xyzz v;
        load_xyzz(v, src);
        jac r;
        xyzz_to_jac<F>(r, v);
        J[i] = r;
    }
    __syncthreads();
    // window sum = T + 2^g sum_i 2^i Y_i (Horner on one wavefront, row-distributed), then 2^(c j)
    // times it
    row::RowK<F> k;
    row::rowk_init(k);
    row::rjac acc, tt;
    if (!live)
        row::rjac_identity<F>(k, acc);
    else
    {
        if (nY == 0)
            row::rjac_load<F>(k, acc, J);
        else
        {
            row::rhorner<F>(k, acc, J, 0, (int)nY, 1, true, nullptr);
            if (!row::ris_zero<F>(acc.Z))
                row::rjac_dbl_n<F>(k, acc, (int)ilog2u(G));
            row::rjac_load<F>(k, tt, J + nY);
            row::rjac_add<F>(k, acc, acc, tt);
        }
        if (j > 0 && !row::ris_zero<F>(acc.Z))
            row::rjac_dbl_n<F>(k, acc, c * (int)j);
    }
    // the group's sums added up by a tree over the wavefronts (every branch below is wave-uniform)
    jac r;
    for (uint32_t s = 1; s < wg; s <<= 1)
    {
        if ((j & (2 * s - 1)) == s)
        {
            row::rjac_get<F>(r, acc);
            if (i == 0)
                X[j] = r;
        }
        __syncthreads();
        if ((j & (2 * s - 1)) == 0 && j + s < wg)
        {
            row::rjac_load<F>(k, tt, X + j + s);
            row::rjac_add<F>(k, acc, acc, tt);
        }
    }
    if (j != 0)
        return;
    row::rjac_get<F>(r, acc);
    if (i == 0)
        win_out[w_lo + ch * groups + gi] = r;
}

// ---------------------------------------------------------------- window combine + encode
// Horner over the window sums in Jacobian coordinates: acc = 2^c acc + S_w, w = W-1 .. 0
// (reference helios_msm_vartime.cpp:282-289, 366-378).  One thread; writes the Jacobian total.
// Keeps a value in vector registers: a single-thread chain of uniform values would otherwise be
// computed on the scalar unit, whose 32-bit multiplies are far slower than v_mad_u64_u32.
__device__ __forceinline__ void force_vgpr_jac(jac &p)
{
#pragma unroll
    for (int i = 0; i < NL; i++)
    {
        asm volatile("" : "+v"(p.X.v[i]));
        asm volatile("" : "+v"(p.Y.v[i]));
        asm volatile("" : "+v"(p.Z.v[i]));
    }
}

__device__ __forceinline__ void force_vgpr(xyzz &p)
{
#pragma unroll
    for (int i = 0; i < NL; i++)
    {
        asm volatile("" : "+v"(p.X.v[i]));
        asm volatile("" : "+v"(p.Y.v[i]));
        asm volatile("" : "+v"(p.ZZ.v[i]));
        asm volatile("" : "+v"(p.ZZZ.v[i]));
    }
}

// Horner step over windows [w_lo, w_hi), highest first: acc = 2^c acc + S_w.  `first` starts
// from the identity, otherwise from the Jacobian value in *total; the result goes back to *total.
// One quad of lanes runs the chain (quad-parallel doublings and additions).
template <class C>
__global__ void __launch_bounds__(64) k_horner(const jac *__restrict__ win_out, int w_lo, int w_hi, int c, int first,
                                               jac *total)
{
    using F = typename C::F;
    if (threadIdx.x >= 4 || blockIdx.x != 0)
        return;
    jac acc, t;
    int w = w_hi - 1;
    if (first)
    {
        acc = win_out[w];
        w--;
    }
    else
        acc = *total;
    for (; w >= w_lo; w--)
    {
        // doublings of the identity are skipped: windows above every scalar's top digit are empty
        if (!fe_is_zero_norm<F>(acc.Z))
            jac_dbl_n_q<F>(acc, c);
        t = win_out[w];
        jac_add_q<F>(acc, acc, t);
    }
    if (threadIdx.x == 0)
        *total = acc;
}

// out32: compressed point (8 words); out_jac: reference Jacobian struct (15 uint64): (x : y : 1)
// canonical, or (1 : 1 : 0) for the identity (helios_ops.h:44-50).  Runs on one lane.
template <class C>
__device__ __forceinline__ void finalize_point(const jac &acc, uint32_t *__restrict__ out32, uint64_t *__restrict__ out_jac)
{
    using F = typename C::F;
    aff a;
    bool pt = jac_to_aff<F>(a, acc);
    uint32_t enc[8];
    encode_aff<F>(enc, a, pt);
    if (out32)
        for (int k = 0; k < 8; k++)
            out32[k] = enc[k];
    if (out_jac)
    {
        uint64_t l[5];
        if (pt)
        {
            fe_to_r51<F>(l, a.x);
            for (int k = 0; k < 5; k++)
                out_jac[k] = l[k];
            fe_to_r51<F>(l, a.y);
            for (int k = 0; k < 5; k++)
                out_jac[5 + k] = l[k];
            for (int k = 0; k < 5; k++)
                out_jac[10 + k] = k == 0 ? 1u : 0u;
        }
        else
        {
            for (int k = 0; k < 15; k++)
                out_jac[k] = (k == 0 || k == 5) ? 1u : 0u;
        }
    }
}

template <class C>
__global__ void __launch_bounds__(64) k_finalize(const jac *__restrict__ total, uint32_t *__restrict__ out32,
                                                 uint64_t *__restrict__ out_jac)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    jac acc = *total;
    finalize_point<C>(acc, out32, out_jac);
}

// The Horner step with every field element spread over a DPP row (rowcurve.h): one whole
// wavefront, four products per formula level.  With an output (out32 or out_jac) the kernel also
// finalizes (affine conversion and encoding, on lane 0) instead of storing the total.
template <class C>
__global__ void __launch_bounds__(64) k_horner_row(const jac *__restrict__ win_out, int w_lo, int w_hi, int c,
                                                   int first, jac *total, uint32_t *__restrict__ out32,
                                                   uint64_t *__restrict__ out_jac)
{
    using F = typename C::F;
    row::RowK<F> k;
    row::rowk_init(k);
    row::rjac acc;
    row::rhorner<F>(k, acc, win_out, w_lo, w_hi, c, first != 0, total);
    jac r;
    row::rjac_get<F>(r, acc);
    if (threadIdx.x != 0)
        return;
    if (!out32 && !out_jac)
    {
        *total = r;
        return;
    }
    force_vgpr_jac(r); // the finalization of a wave-uniform value would otherwise run on the scalar unit
    finalize_point<C>(r, out32, out_jac);
}

} // namespace hs

#include "encode.h"

namespace hs
{

template <class C> struct CurveK;
template <> struct CurveK<HeliosC>
{
    using K = HeliosK;
};
template <> struct CurveK<SeleneC>
{
    using K = SeleneK;
};

template <class F> __device__ __forceinline__ void store_r51_aff(uint64_t *o, const aff &a, bool is_point)
{
    uint64_t l[5];
    if (is_point)
    {
        fe_to_r51<F>(l, a.x);
        for (int k = 0; k < 5; k++)
            o[k] = l[k];
        fe_to_r51<F>(l, a.y);
        for (int k = 0; k < 5; k++)
            o[5 + k] = l[k];
        for (int k = 0; k < 5; k++)
            o[10 + k] = k == 0 ? 1u : 0u;
    }
    else
    {
        for (int k = 0; k < 15; k++)
            o[k] = (k == 0 || k == 5) ? 1u : 0u;
    }
}

// SSWU hash-to-curve of n 32-byte inputs -> reference Jacobian structs (Z = 1) and/or compressed
template <class C>
__global__ void __launch_bounds__(256) k_map_to_curve(const uint32_t *__restrict__ u, uint32_t n,
                                                      uint64_t *__restrict__ out_jac, uint32_t *__restrict__ out32)
{
    using F = typename C::F;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    uint32_t w[8];
    for (int k = 0; k < 8; k++)
        w[k] = u[8ull * i + k];
    aff a;
    sswu<F, typename CurveK<C>::K>(a, w);
    if (out_jac)
        store_r51_aff<F>(out_jac + 15ull * i, a, true);
    if (out32)
    {
        uint32_t e[8];
        encode_aff<F>(e, a, true);
        for (int k = 0; k < 8; k++)
            out32[8ull * i + k] = e[k];
    }
}

// compressed points -> reference Jacobian structs; status[i] = 0 ok / 1 identity / 0xff invalid
template <class C>
__global__ void __launch_bounds__(256) k_decompress(const uint32_t *__restrict__ in32, uint32_t n,
                                                    uint64_t *__restrict__ out_jac, uint8_t *__restrict__ status,
                                                    uint32_t *__restrict__ nbad)
{
    using F = typename C::F;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    uint32_t w[8];
    for (int k = 0; k < 8; k++)
        w[k] = in32[8ull * i + k];
    aff a;
    int rc = decompress<F, typename CurveK<C>::K>(a, w);
    store_r51_aff<F>(out_jac + 15ull * i, a, rc == 0);
    if (status)
        status[i] = rc < 0 ? 0xffu : (uint8_t)rc;
    if (rc < 0)
        atomicAdd(nbad, 1u);
}

// INT32 multiply throughput probe: 8 independent v_mad_u64_u32 accumulation chains per thread
// (one 32x32->64 partial product plus a 64-bit add per instruction, no other work in the loop).
__global__ void __launch_bounds__(256) k_probe_mad64(uint32_t *out, int iters, uint32_t seed)
{
    uint64_t acc[8];
    uint32_t a[8];
    for (int j = 0; j < 8; j++)
    {
        acc[j] = threadIdx.x + j;
        a[j] = seed * (j + 3) + threadIdx.x;
    }
    const uint32_t y = seed | 0x80000001u;
    for (int i = 0; i < iters; i++)
    {
#pragma unroll
        for (int j = 0; j < 8; j++)
            asm volatile("v_mad_u64_u32 %0, vcc, %1, %2, %0" : "+v"(acc[j]) : "v"(a[j]), "v"(y) : "vcc");
    }
    uint32_t r = 0;
    for (int j = 0; j < 8; j++)
        r ^= (uint32_t)(acc[j] >> 32) ^ (uint32_t)acc[j];
    out[blockIdx.x * blockDim.x + threadIdx.x] = r;
}

} // namespace hs

namespace hs
{
// Encodes an array of xyzz points to compressed bytes (inspection of intermediate state).
template <class C>
__global__ void __launch_bounds__(64) k_encode_xyzz(const xyzz *__restrict__ p, uint32_t n, uint32_t *__restrict__ out32)
{
    using F = typename C::F;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    xyzz q;
    load_xyzz(q, p + i);
    aff a;
    bool pt = xyzz_to_aff<F>(a, q);
    uint32_t e[8];
    encode_aff<F>(e, a, pt);
    for (int k = 0; k < 8; k++)
        out32[8ull * i + k] = e[k];
}
// Jacobian points -> 32-byte encodings (identity -> 0^32); for debugging dumps.
template <class C>
__global__ void __launch_bounds__(64) k_encode_jac(const jac *__restrict__ p, uint32_t n, uint32_t *__restrict__ out32)
{
    using F = typename C::F;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const jac q = p[i];
    aff a;
    bool pt = jac_to_aff<F>(a, q);
    uint32_t e[8];
    encode_aff<F>(e, a, pt);
    for (int k = 0; k < 8; k++)
        out32[8ull * i + k] = e[k];
}
} // namespace hs


// ================================================================ batched MSMs
// A group of M MSMs of one curve sharing the window width c.  Points of all MSMs are concatenated
// (MSM m owns global points [p_off[m], p_off[m] + n[m])); each MSM has W windows, and virtual
// window v = m*W + w owns the entry slots [p_off[m]*W + w*n[m], ... + n[m]) of the sorted array.
// Zero digits become ENTRY_SKIP slots at the end of the window's last bucket, so the layout, every
// buffer size and every grid depend only on the batch shape.
namespace hs
{

struct BatchDesc
{
    const uint32_t *scalars; // n x 8 words (device)
    const uint64_t *points;  // n x 15 uint64 (reference Jacobian struct, device)
    uint32_t *out32;         // 8 words (device), may be null
    uint64_t *outjac;        // 15 uint64 (device), may be null
    uint32_t n, p_off, w8_compat, pad;
};

__device__ __forceinline__ uint32_t find_msm(const BatchDesc *__restrict__ d, uint32_t M, uint32_t i)
{
    uint32_t lo = 0, hi = M; // largest m with p_off[m] <= i
    while (hi - lo > 1)
    {
        const uint32_t mid = (lo + hi) >> 1;
        if (d[mid].p_off <= i)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

template <class C>
__global__ void __launch_bounds__(256) k_ingest_batch(const BatchDesc *__restrict__ d, uint32_t M, uint32_t N,
                                                      uint32_t *__restrict__ out, uint8_t *__restrict__ is_id)
{
    using F = typename C::F;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    const uint32_t m = find_msm(d, M, i), j = i - d[m].p_off;
    const uint64_t *p = d[m].points + 15ull * j;
    uint64_t l[5];
    fe X, Y, Z;
    for (int k = 0; k < 5; k++)
        l[k] = p[k];
    fe_from_r51<F>(X, l);
    for (int k = 0; k < 5; k++)
        l[k] = p[5 + k];
    fe_from_r51<F>(Y, l);
    for (int k = 0; k < 5; k++)
        l[k] = p[10 + k];
    fe_from_r51<F>(Z, l);
    aff a;
    uint32_t Zc[8];
    fe_to_words_canon<F>(Zc, Z);
    uint32_t zor = 0, one = Zc[0] == 1u;
    for (int k = 0; k < 8; k++)
        zor |= Zc[k];
    for (int k = 1; k < 8; k++)
        one &= Zc[k] == 0u;
    if (zor == 0)
    {
        fe_zero(a.x);
        fe_zero(a.y);
        is_id[i] = 1;
    }
    else
    {
        if (one)
        {
            a.x = X;
            a.y = Y;
        }
        else
        {
            fe zi, zi2, zi3;
            fe_inv<F>(zi, Z);
            fe_sq<F>(zi2, zi);
            fe_mul<F>(zi3, zi2, zi);
            fe_mul<F>(a.x, X, zi2);
            fe_mul<F>(a.y, Y, zi3);
        }
        is_id[i] = 0;
    }
    store_base_record<F>(out, i, a);
}

template <class C>
__global__ void __launch_bounds__(256) k_recode_batch(const BatchDesc *__restrict__ d, uint32_t M, uint32_t N,
                                                      const uint8_t *__restrict__ is_id, int c, int W,
                                                      uint16_t *__restrict__ digits)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    const uint32_t m = find_msm(d, M, i), j = i - d[m].p_off;
    const uint4 *sp = reinterpret_cast<const uint4 *>(d[m].scalars + 8ull * j);
    uint4 a = sp[0], b = sp[1];
    const uint32_t s[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    uint32_t k[8];
    scalar_prepare<typename C::O>(k, s, d[m].w8_compat != 0);
    const bool skip = is_id[i] != 0;
    uint32_t carry = 0;
    for (int w = 0; w < W; w++)
    {
        uint32_t dg = recode_digit(k, c, w, carry);
        digits[(size_t)w * N + i] =
            (dg == 0 || skip) ? DIGIT_ZERO : (uint16_t)(((dg & 0x7fffffffu) - 1) | ((dg >> 31) << 15));
    }
}

// One block per virtual window: LDS counting sort of the window's n digits into B buckets.
constexpr int WSORT_MAX_B = 4096;

__global__ void __launch_bounds__(256) k_window_sort(const BatchDesc *__restrict__ d, uint32_t N, int W, uint32_t B,
                                                     const uint16_t *__restrict__ digits, uint32_t *__restrict__ sorted,
                                                     uint32_t *__restrict__ offsets, uint32_t V)
{
    __shared__ uint32_t cnt[WSORT_MAX_B];
    __shared__ uint32_t part[256];
    const uint32_t v = blockIdx.x, m = v / W, w = v % W;
    const uint32_t n = d[m].n, p_off = d[m].p_off;
    const uint32_t base = p_off * (uint32_t)W + w * n;
    const uint16_t *dg = digits + (size_t)w * N + p_off;
    for (uint32_t b = threadIdx.x; b < B; b += blockDim.x)
        cnt[b] = 0;
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < n; j += blockDim.x)
    {
        const uint16_t code = dg[j];
        if (code != DIGIT_ZERO)
            atomicAdd(&cnt[code & 0x7fffu], 1u);
    }
    __syncthreads();
    // exclusive scan of B counters: each thread scans a contiguous slice, then the slice totals
    const uint32_t per = (B + blockDim.x - 1) / blockDim.x, b0 = threadIdx.x * per;
    uint32_t tot = 0;
    for (uint32_t b = b0; b < b0 + per && b < B; b++)
        tot += cnt[b];
    part[threadIdx.x] = tot;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        uint32_t run = 0;
        for (uint32_t t = 0; t < blockDim.x; t++)
        {
            const uint32_t x = part[t];
            part[t] = run;
            run += x;
        }
    }
    __syncthreads();
    uint32_t run = base + part[threadIdx.x];
    for (uint32_t b = b0; b < b0 + per && b < B; b++)
    {
        const uint32_t x = cnt[b];
        cnt[b] = run; // becomes the scatter cursor
        offsets[(size_t)v * B + b] = run;
        run += x;
    }
    if (v == V - 1 && threadIdx.x == 0)
        offsets[(size_t)V * B] = base + n;
    __syncthreads();
    // nonzero digits into their buckets; the remaining slots of the window are padding
    for (uint32_t j = threadIdx.x; j < n; j += blockDim.x)
    {
        const uint16_t code = dg[j];
        if (code != DIGIT_ZERO)
        {
            const uint32_t pos = atomicAdd(&cnt[code & 0x7fffu], 1u);
            sorted[pos] = (p_off + j) | ((uint32_t)(code >> 15) << 31);
        }
    }
    __syncthreads();
    // after the scatter cnt[B-1] points one past the last real entry of the window
    for (uint32_t j = cnt[B - 1] + threadIdx.x; j < base + n; j += blockDim.x)
        sorted[j] = ENTRY_SKIP;
}

// Horner per MSM (one quad of lanes each) over its W window sums.
template <class C>
__global__ void __launch_bounds__(64) k_horner_batch(const jac *__restrict__ win_out, uint32_t M, int W, int c,
                                                     jac *__restrict__ total)
{
    using F = typename C::F;
    const uint32_t m = (blockIdx.x * blockDim.x + threadIdx.x) >> 2;
    if (m >= M)
        return;
    const jac *ws = win_out + (size_t)m * W;
    jac acc, t;
    acc = ws[W - 1];
    for (int w = W - 2; w >= 0; w--)
    {
        if (!fe_is_zero_norm<F>(acc.Z)) // the identity doubles to itself
            jac_dbl_n_q<F>(acc, c);
        t = ws[w];
        jac_add_q<F>(acc, acc, t);
    }
    if ((threadIdx.x & 3) == 0)
        total[m] = acc;
}

// Horner per MSM with the row-distributed arithmetic, one wavefront (workgroup) per MSM, then the
// MSM's affine conversion and encoding on lane 0 (k_finalize_batch's work).
template <class C>
__global__ void __launch_bounds__(64) k_horner_batch_row(const jac *__restrict__ win_out, uint32_t M, int W, int c,
                                                         const BatchDesc *__restrict__ d)
{
    using F = typename C::F;
    const uint32_t m = blockIdx.x;
    if (m >= M)
        return;
    row::RowK<F> k;
    row::rowk_init(k);
    row::rjac acc;
    row::rhorner<F>(k, acc, win_out + (size_t)m * W, 0, W, c, true, nullptr);
    jac r;
    row::rjac_get<F>(r, acc);
    if (threadIdx.x != 0)
        return;
    force_vgpr_jac(r);
    aff a;
    const bool pt = jac_to_aff<F>(a, r);
    uint32_t enc[8];
    encode_aff<F>(enc, a, pt);
    if (d[m].out32)
        for (int j = 0; j < 8; j++)
            d[m].out32[j] = enc[j];
    if (d[m].outjac)
        store_r51_aff<F>(d[m].outjac, a, pt);
}

template <class C>
__global__ void __launch_bounds__(64) k_finalize_batch(const jac *__restrict__ total, const BatchDesc *__restrict__ d,
                                                       uint32_t M)
{
    using F = typename C::F;
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M)
        return;
    jac acc = total[m];
    aff a;
    const bool pt = jac_to_aff<F>(a, acc);
    uint32_t enc[8];
    encode_aff<F>(enc, a, pt);
    if (d[m].out32)
        for (int k = 0; k < 8; k++)
            d[m].out32[k] = enc[k];
    if (d[m].outjac)
        store_r51_aff<F>(d[m].outjac, a, pt);
}

} // namespace hs

namespace hs
{
// Sum of k partial MSM results (reference structs, affine with Z = 1 or the identity 1:1:0) by one
// quad of lanes, then the encoding: the total of an MSM computed in chunks.
template <class C>
__global__ void __launch_bounds__(64) k_sum_parts(const uint64_t *__restrict__ parts, uint32_t k,
                                                  uint32_t *__restrict__ out32, uint64_t *__restrict__ outjac)
{
    using F = typename C::F;
    if (threadIdx.x >= 4)
        return;
    jac acc, t;
    jac_identity<F>(acc);
    for (uint32_t j = 0; j < k; j++)
    {
        uint64_t l[5];
        for (int i = 0; i < 5; i++)
            l[i] = parts[15ull * j + i];
        fe_from_r51<F>(t.X, l);
        for (int i = 0; i < 5; i++)
            l[i] = parts[15ull * j + 5 + i];
        fe_from_r51<F>(t.Y, l);
        for (int i = 0; i < 5; i++)
            l[i] = parts[15ull * j + 10 + i];
        fe_from_r51<F>(t.Z, l);
        jac_add_q<F>(acc, acc, t);
    }
    if (threadIdx.x != 0)
        return;
    aff a;
    const bool pt = jac_to_aff<F>(a, acc);
    uint32_t enc[8];
    encode_aff<F>(enc, a, pt);
    if (out32)
        for (int i = 0; i < 8; i++)
            out32[i] = enc[i];
    if (outjac)
        store_r51_aff<F>(outjac, a, pt);
}

// identity outputs for the empty MSMs (n = 0) of a batch: 0^32 and the reference's identity
// Jacobian (X = 1, Y = 1, Z = 0 in radix 2^51)
__global__ void k_identity_out(const BatchDesc *__restrict__ d, uint32_t M)
{
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M)
        return;
    if (d[m].out32)
        for (int k = 0; k < 8; k++)
            d[m].out32[k] = 0;
    if (d[m].outjac)
        for (int k = 0; k < 15; k++)
            d[m].outjac[k] = (k == 0 || k == 5) ? 1u : 0u;
}

// graph-safe fill (captured graphs get no memset nodes)
__global__ void k_fill_u32(uint32_t *p, uint32_t n, uint32_t v)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}
} // namespace hs
