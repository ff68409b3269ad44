// Host side of the MI355X MSM library: per-device context, scratch management, launch sequence of
// the plain Pippenger path, and the C ABI declared in include/hs_gpu_msm.h.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <string>
#include <vector>

#include "../../include/hs_gpu_msm.h"
#include "msm_kernels.h"
#include "fixed_kernels.h"
#include "affine_kernels.h"

namespace
{

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HS_CHECK(expr)                                                                                        \
    do                                                                                                        \
    {                                                                                                         \
        hipError_t e_ = (expr);                                                                               \
        if (e_ != hipSuccess)                                                                                 \
            return fail(HS_GPU_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));                   \
    } while (0)

// Device bytes the library holds, and an optional cap on them (hs_gpu_set_device_limit, or the
// HS_GPU_DEVICE_LIMIT environment variable in bytes).  An allocation that would pass the cap
// fails with HS_GPU_ENOMEM, as a refused hipMalloc does.
size_t env_size(const char *name)
{
    const char *v = std::getenv(name);
    return v ? (size_t)std::strtoull(v, nullptr, 10) : 0;
}
std::atomic<size_t> g_dev_held{0};
std::atomic<size_t> g_dev_limit{env_size("HS_GPU_DEVICE_LIMIT")};

int dev_malloc(void **p, size_t bytes)
{
    const size_t lim = g_dev_limit.load();
    if (lim && g_dev_held.load() + bytes > lim)
    {
        *p = nullptr;
        return fail(HS_GPU_ENOMEM, "device allocation of " + std::to_string(bytes) + " bytes would pass the library's " +
                                       std::to_string(lim) + "-byte limit");
    }
    if (hipMalloc(p, bytes) != hipSuccess)
    {
        *p = nullptr;
        return fail(HS_GPU_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    g_dev_held += bytes;
    return 0;
}

void dev_free(void *p, size_t bytes)
{
    if (!p)
        return;
    hipFree(p);
    g_dev_held -= bytes;
}

// A growable device buffer (freed only by release()).
struct DevBuf
{
    void *p = nullptr;
    size_t cap = 0;
    void release()
    {
        dev_free(p, cap);
        p = nullptr;
        cap = 0;
    }
    int reserve(size_t bytes)
    {
        if (bytes <= cap)
            return 0;
        release();
        const size_t want = bytes + bytes / 4 + 4096;
        if (int r = dev_malloc(&p, want))
            return r;
        cap = want;
        return 0;
    }
    // exactly `bytes` (no growth slack), for buffers sized once such as fixed-base tables
    int reserve_exact(size_t bytes)
    {
        if (bytes <= cap)
            return 0;
        release();
        if (int r = dev_malloc(&p, bytes))
            return r;
        cap = bytes;
        return 0;
    }
    template <class T> T *as() const { return reinterpret_cast<T *>(p); }
};

// Overwrites host memory holding secrets; the volatile stores cannot be elided.
void secure_erase(void *p, size_t bytes)
{
    volatile unsigned char *q = (volatile unsigned char *)p;
    for (size_t i = 0; i < bytes; i++)
        q[i] = 0;
}

// Stream-ordered zeroing of `bytes` (rounded up to whole words) with the fill kernel, which
// stream capture records as an ordinary kernel node.
void zero_dev(void *p, size_t bytes, hipStream_t st)
{
    const size_t words = (bytes + 3) / 4;
    if (!p || !words)
        return;
    for (size_t off = 0; off < words; off += (size_t)1 << 30)
    {
        const uint32_t k = (uint32_t)std::min<size_t>(words - off, (size_t)1 << 30);
        hipLaunchKernelGGL(hs::k_fill_u32, dim3((k + 255) / 256), dim3(256), 0, st, (uint32_t *)p + off, k, 0u);
    }
}

struct Plan
{
    int c, W, Q;
    uint32_t B, G, nb, L;
    int affK = 0; // bucket method: 0 = xyzz mixed additions (k_bucket_flat), K = batched affine with K chains per lane
};

// Window width for an n-point MSM: balances W*n bucket additions against the W*2^c reduction
// work; c * W >= 256 always holds since W = ceil(256 / c).
int choose_c(size_t n)
{
    int lg = 0;
    while (((size_t)1 << (lg + 1)) <= n)
        lg++;
    int c = lg - 4;
    if (c < 4)
        c = 4;
    if (c > 16)
        c = 16;
    return c;
}

// Built-in run length (sorted entries per bucket-kernel thread) and segment size (buckets per
// reduction segment) of a plain MSM.  Below 2^15 points the MSM is a chain of latency-bound
// kernels: every bucket-kernel thread adds its L entries one after another and every reduction
// thread 2 G buckets, so short runs and segments finish first (measured on MI355X: 0.65 -> 0.56 ms
// per MSM over sizes 64 .. 2^14).  From 2^19 points, runs of 64 entries cut half as many buckets
// as runs of 32 (buckets average 32 entries at 2^20), which halves the merge: whole MSM at 2^20
// 1.61 / 2.27 ms against 1.66 / 2.33 ms (Helios / Selene).
uint32_t plain_default_L(size_t n)
{
    return n < ((size_t)1 << 15) ? 16u : n < ((size_t)1 << 19) ? 32u : 64u;
}
uint32_t plain_default_G(size_t n) { return n < ((size_t)1 << 15) ? 4u : 8u; }

// Tuning overrides for experiments: HS_GPU_C (window bits), HS_GPU_G (buckets per reduction
// segment, a power of two).
int env_int(const char *name, int dflt)
{
    const char *v = std::getenv(name);
    return v ? std::atoi(v) : dflt;
}

// Horner combines run with the row-distributed arithmetic (one wavefront per chain); HS_GPU_HORNER=quad
// selects the earlier quad-of-lanes kernels (same results) for comparisons.
const bool g_horner_quad = [] {
    const char *v = std::getenv("HS_GPU_HORNER");
    return v && std::strcmp(v, "quad") == 0;
}();

// ---------------------------------------------------------------- tuned plan parameters
// The autotuner's table: for (mode, curve, size class) the fastest (c, L, G) it measured.  mode 0 =
// plain MSM, 1 = a group of a batch; size class = floor(log2 n).  Lookups fall back to the nearest
// tuned class of the same mode and curve, then to the built-in defaults.
struct TuneEntry
{
    int c = 0, L = 32, G = 8;
    float ms = 0.f;
};

struct TuneTable
{
    std::mutex mu;
    std::map<int, TuneEntry> e; // key = mode * 1024 + curve * 64 + lg
    bool loaded_env = false;
};

TuneTable &tune()
{
    static TuneTable t;
    return t;
}

int size_class(size_t n)
{
    int lg = 0;
    while (((size_t)1 << (lg + 1)) <= n)
        lg++;
    return lg;
}

int tune_load_file(const char *path);

// Forced parameters while the autotuner measures a candidate (set on the tuning thread only).
thread_local const TuneEntry *g_force = nullptr;

bool tune_lookup(int mode, int curve, size_t n, TuneEntry &out)
{
    if (g_force)
    {
        out = *g_force;
        return true;
    }
    TuneTable &t = tune();
    std::lock_guard<std::mutex> lk(t.mu);
    if (!t.loaded_env)
    {
        t.loaded_env = true;
        if (const char *f = std::getenv("HS_GPU_TUNE_FILE"))
        {
            t.mu.unlock();
            tune_load_file(f);
            t.mu.lock();
        }
    }
    const int lg = size_class(n);
    int best = -1, bd = 1 << 30;
    for (const auto &kv : t.e)
    {
        const int m = kv.first / 1024, c = (kv.first / 64) % 16, l = kv.first % 64;
        if (m != mode || c != curve)
            continue;
        const int d = l > lg ? l - lg : lg - l;
        if (d < bd && d <= 1)
        {
            bd = d;
            best = kv.first;
        }
    }
    if (best < 0)
        return false;
    out = t.e[best];
    return true;
}

// Buckets per reduction segment: a power of two in [B / MAX_SEGS_PER_WIN, B] (the window
// reduction doubles its subset sums log2 G times and keeps at most MAX_SEGS_PER_WIN / 256 blocks of
// a window in LDS).
static uint32_t plan_G(uint32_t B, uint32_t G)
{
    uint32_t p = 1;
    while (2 * p <= G)
        p *= 2;
    p = std::min(p, B);
    return std::max(p, B / hs::MAX_SEGS_PER_WIN > 0 ? B / hs::MAX_SEGS_PER_WIN : 1u);
}

Plan make_plan(int curve, size_t n)
{
    Plan p;
    TuneEntry te;
    const bool tuned = tune_lookup(0, curve, n, te);
    p.c = env_int("HS_GPU_C", tuned ? te.c : choose_c(n));
    if (p.c < 4)
        p.c = 4;
    if (p.c > 16)
        p.c = 16;
    p.W = (256 + p.c - 1) / p.c;
    p.B = 1u << (p.c - 1);
    p.G = plan_G(p.B, (uint32_t)env_int("HS_GPU_G", tuned ? te.G : (int)plain_default_G(n)));
    p.nb = (uint32_t)p.W * p.B;
    p.L = (uint32_t)env_int("HS_GPU_L", tuned ? te.L : (int)plain_default_L(n));
    if (p.L < 1)
        p.L = 1;
    p.Q = env_int("HS_GPU_Q", 1);
    if (p.Q < 1)
        p.Q = 1;
    // HS_GPU_AFFINE_K = 1, 2 or 4 selects the batched-affine bucket kernel (measurements and tests)
    p.affK = env_int("HS_GPU_AFFINE_K", 0);
    if (p.affK != 1 && p.affK != 2 && p.affK != 4)
        p.affK = 0;
    return p;
}

int ref_window(size_t n)
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

// Scratch of one in-flight MSM; one workspace per stream, so MSMs issued on different streams
// may overlap on the device.
struct Workspace
{
    DevBuf aff, isid, digits, counts, offsets, block_sums, sorted, buckets, seg, win, big_list, big_count, entries,
        cnt_a, off_a, total, chunk_bucket, bnd, parts;
    // second stream for the window combine, which runs behind the bucket work of lower windows
    hipStream_t aux = nullptr;
    static constexpr int MAXQ = 8;
    hipEvent_t ev_group[MAXQ] = {}, ev_done = nullptr;

    void release_all()
    {
        for (DevBuf *d : {&aff, &isid, &digits, &counts, &offsets, &block_sums, &sorted, &buckets, &seg, &win, &big_list,
                          &big_count, &entries, &cnt_a, &off_a, &total, &chunk_bucket, &bnd, &parts})
            d->release();
    }

    int ensure_aux()
    {
        if (aux)
            return 0;
        HS_CHECK(hipStreamCreateWithFlags(&aux, hipStreamNonBlocking));
        for (int i = 0; i < MAXQ; i++)
            HS_CHECK(hipEventCreateWithFlags(&ev_group[i], hipEventDisableTiming));
        HS_CHECK(hipEventCreateWithFlags(&ev_done, hipEventDisableTiming));
        return 0;
    }
};

// Scratch of the no-allocation entry point (hs_gpu_msm_device_noalloc): one array per device, in
// the code object, so the entry point itself allocates nothing.  Host code reaches it only through
// the device address hipGetSymbolAddress returns at initialisation (Ctx::scratch_ptr); the array's
// name is a host-side stand-in.  Every call on a device shares it: calls that overlap on one device
// race.  64 MiB hold the plan of an MSM of up to 2^16 points.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
constexpr size_t NOALLOC_MAX_POINTS = (size_t)1 << 16;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

// Staging of one host-pointer MSM in flight: a stream of its own (and with it a workspace of its
// own) plus the upload and result buffers.  Concurrent host callers each take one from a pool.
// Host-to-device uploads from the caller's (pageable) memory.  HIP copies pageable memory through
// its own staging at about 26 GB/s on MI355X; here a few host threads copy each chunk into a pinned
// slot (66 GB/s with four threads) while the DMA engine moves the previous chunk from the other
// pinned slot (57 GB/s).
constexpr size_t UPLOAD_CHUNK = (size_t)16 << 20;   // bytes per pinned slot
constexpr size_t UPLOAD_DIRECT = (size_t)4 << 20;   // smaller uploads go straight through HIP

// memcpy split over a small pool of host threads (the caller takes one part).  The pool is
// created on first use and lives until the process ends; its workers sleep between jobs.
class CopyPool
{
  public:
    static CopyPool &get()
    {
        static CopyPool *p = new CopyPool(3);
        return *p;
    }
    void copy(void *dst, const void *src, size_t bytes)
    {
        std::lock_guard<std::mutex> one(call_mu_); // one job at a time
        const size_t parts = workers_.size() + 1, per = (bytes + parts - 1) / parts;
        {
            std::lock_guard<std::mutex> lk(mu_);
            dst_ = (char *)dst;
            src_ = (const char *)src;
            bytes_ = bytes;
            per_ = per;
            pending_ = workers_.size();
            gen_++;
        }
        cv_.notify_all();
        const size_t off = workers_.size() * per;
        if (off < bytes)
            std::memcpy((char *)dst + off, (const char *)src + off, bytes - off);
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [&] { return pending_ == 0; });
    }

  private:
    explicit CopyPool(size_t threads)
    {
        for (size_t i = 0; i < threads; i++)
            workers_.emplace_back([this, i] { run(i); });
        for (auto &w : workers_)
            w.detach();
    }
    void run(size_t i)
    {
        uint64_t seen = 0;
        std::unique_lock<std::mutex> lk(mu_);
        while (true)
        {
            cv_.wait(lk, [&] { return gen_ != seen; });
            seen = gen_;
            char *d = dst_;
            const char *s = src_;
            const size_t b = bytes_, off = i * per_, len = off < b ? std::min(per_, b - off) : 0;
            lk.unlock();
            if (len)
                std::memcpy(d + off, s + off, len);
            lk.lock();
            if (--pending_ == 0)
                done_.notify_all();
        }
    }
    std::vector<std::thread> workers_;
    std::mutex call_mu_, mu_;
    std::condition_variable cv_, done_;
    char *dst_ = nullptr;
    const char *src_ = nullptr;
    size_t bytes_ = 0, per_ = 0, pending_ = 0;
    uint64_t gen_ = 0;
};

struct HostStage
{
    hipStream_t st = nullptr;
    DevBuf scalars, points, out;
    unsigned char *pin[2] = {nullptr, nullptr}; // pinned upload slots (UPLOAD_CHUNK each)
    hipEvent_t pin_ev[2] = {nullptr, nullptr};  // the last DMA out of each slot
    bool pin_used[2] = {false, false};
    bool pin_dirty = false; // the slots hold data of an upload not yet erased

    // dev <- src (host, any memory), ordered on stream s (the slot's own stream by default);
    // returns 0 or an error code
    int upload(void *dev, const void *src, size_t bytes, hipStream_t s = nullptr)
    {
        if (!s)
            s = st;
        if (bytes < UPLOAD_DIRECT)
            return hipMemcpyAsync(dev, src, bytes, hipMemcpyHostToDevice, s) == hipSuccess
                       ? 0
                       : fail(HS_GPU_EDEVICE, "input upload failed");
        for (int k = 0; k < 2; k++)
            if (!pin[k])
            {
                if (hipHostMalloc((void **)&pin[k], UPLOAD_CHUNK, hipHostMallocDefault) != hipSuccess)
                {
                    pin[k] = nullptr;
                    return fail(HS_GPU_ENOMEM, "pinned upload buffer allocation failed");
                }
                if (hipEventCreateWithFlags(&pin_ev[k], hipEventDisableTiming) != hipSuccess)
                    return fail(HS_GPU_EDEVICE, "hipEventCreate failed");
            }
        size_t k = 0;
        for (size_t off = 0; off < bytes; off += UPLOAD_CHUNK, k ^= 1)
        {
            const size_t len = std::min(UPLOAD_CHUNK, bytes - off);
            if (pin_used[k] && hipEventSynchronize(pin_ev[k]) != hipSuccess) // the slot's last DMA is done
                return fail(HS_GPU_EDEVICE, "input upload failed");
            CopyPool::get().copy(pin[k], (const char *)src + off, len);
            if (hipMemcpyAsync((char *)dev + off, pin[k], len, hipMemcpyHostToDevice, s) != hipSuccess ||
                hipEventRecord(pin_ev[k], s) != hipSuccess)
                return fail(HS_GPU_EDEVICE, "input upload failed");
            pin_used[k] = true;
            pin_dirty = true;
        }
        return 0;
    }

    // zeroes the pinned slots once their DMAs are done, when an upload has passed through them
    // since the last erase (they may hold scalars)
    void erase_pinned()
    {
        if (!pin_dirty)
            return;
        for (int k = 0; k < 2; k++)
            if (pin[k])
            {
                if (pin_used[k])
                    hipEventSynchronize(pin_ev[k]);
                secure_erase(pin[k], UPLOAD_CHUNK);
            }
        pin_dirty = false;
    }
};

struct Ctx
{
    int device = -1;
    hipStream_t stream = nullptr;
    std::mutex mu;
    DevBuf scalars, points, out32, outjac, status, nbad, probe;
    std::vector<HostStage *> stages_free, stages_all;
    unsigned char *scratch_ptr = nullptr; // device address of `scratch` on this context's device
    Workspace *last_ws = nullptr;         // workspace of the last plain MSM (hs_gpu_debug_dump)
    Workspace noalloc_ws;                 // buffers carved from it (never released)
    std::vector<std::pair<hipStream_t, Workspace *>> wss;
    bool ready = false;

    Workspace &ws(hipStream_t st)
    {
        for (auto &p : wss)
            if (p.first == st)
                return *p.second;
        wss.emplace_back(st, new Workspace());
        return *wss.back().second;
    }
    // phase timing (hs_gpu_profile): HIP events recorded between the phases of every plain MSM
    // issued while profiling is on, on the stream the MSM runs on
    bool profile = false;
    static constexpr int NEV = 48;
    struct EvSet
    {
        hipEvent_t ev[NEV];
        int phase[NEV]; // phase of the interval ending at ev[i]
        int curve;
        int n;
    };
    std::vector<EvSet> evsets;
    size_t evused = 0;
    EvSet *cur = nullptr;

    int begin_profile(int curve)
    {
        cur = nullptr;
        if (!profile)
            return 0;
        if (evused == evsets.size())
        {
            EvSet e;
            for (int i = 0; i < NEV; i++)
                HS_CHECK(hipEventCreate(&e.ev[i]));
            evsets.push_back(e);
        }
        cur = &evsets[evused++];
        cur->curve = curve;
        cur->n = 0;
        return 0;
    }

    // records an event on st; the time since the previous mark is attributed to `phase`
    void mark(hipStream_t st, int phase)
    {
        if (cur && cur->n < NEV)
        {
            cur->phase[cur->n] = phase;
            hipEventRecord(cur->ev[cur->n++], st);
        }
    }

    int init()
    {
        if (ready)
            return 0;
        int dev = 0;
        HS_CHECK(hipGetDevice(&dev));
        hipDeviceProp_t prop;
        HS_CHECK(hipGetDeviceProperties(&prop, dev));
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
            return fail(HS_GPU_EDEVICE, std::string("device is ") + prop.gcnArchName + ", this library is built for gfx950");
        device = dev;
        HS_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        HS_CHECK(hipGetSymbolAddress((void **)&scratch_ptr, HIP_SYMBOL(scratch)));
        ready = true;
        return 0;
    }
};

Ctx &ctx()
{
    static Ctx c;
    return c;
}

// A free staging slot (cx.mu held); nullptr with the error set when none can be made.
HostStage *take_stage(Ctx &cx)
{
    if (!cx.stages_free.empty())
    {
        HostStage *h = cx.stages_free.back();
        cx.stages_free.pop_back();
        return h;
    }
    auto *h = new HostStage();
    if (hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking) != hipSuccess)
    {
        delete h;
        fail(HS_GPU_EDEVICE, "hipStreamCreate failed");
        return nullptr;
    }
    cx.stages_all.push_back(h);
    return h;
}

int scan_exclusive(Workspace &cx, const uint32_t *in, uint32_t *out, uint32_t len, hipStream_t st)
{
    const uint32_t per = hs::SCAN_BLOCK * hs::SCAN_ITEMS;
    const uint32_t nblk = (len + per - 1) / per;
    if (nblk > per)
        return fail(HS_GPU_EINVAL, "scan too long");
    if (int r = cx.block_sums.reserve(2 * (size_t)(nblk + 1) * 4 + 64))
        return r;
    uint32_t *bs = cx.block_sums.as<uint32_t>();
    uint32_t *bs_off = bs + nblk + 1;
    hipLaunchKernelGGL(hs::k_scan_local, dim3(nblk), dim3(hs::SCAN_BLOCK), 0, st, in, out, len, bs);
    uint32_t *dummy = bs_off + nblk + 1;
    hipLaunchKernelGGL(hs::k_scan_local, dim3(1), dim3(hs::SCAN_BLOCK), 0, st, bs, bs_off, nblk, dummy);
    hipLaunchKernelGGL(hs::k_scan_add, dim3(nblk), dim3(hs::SCAN_BLOCK), 0, st, out, len, bs_off);
    HS_CHECK(hipGetLastError());
    return 0;
}

// The plain MSM on device-resident inputs.  d_scalars: n x 8 words; d_points: n x 15 uint64.
// Point indices in the sort entries have 24 bits: larger MSMs run as chunks (msm_device_any).
constexpr size_t MSM_MAX_CHUNK = (size_t)1 << 23;

// Launch geometry of a plain MSM: two-level sort with nsb super-buckets of 2^shift buckets per
// window and coarse blocks of T points (about 256 blocks, one per CU, at least 256 points each).
struct MsmGeom
{
    int shift, nsb;
    uint32_t T, nblk, nh, scan_len;
    size_t lds, max_chunks;
};

MsmGeom msm_geom(const Plan &P, size_t n)
{
    MsmGeom g;
    const uint32_t nn = (uint32_t)n;
    g.shift = P.c - 1 > 8 ? P.c - 1 - 8 : 0;
    g.nsb = (int)(P.B >> g.shift);
    g.T = 256;
    while ((nn + g.T - 1) / g.T > 256)
        g.T *= 2;
    g.nblk = (nn + g.T - 1) / g.T;
    g.nh = (uint32_t)P.W * g.nsb;
    g.lds = (size_t)g.nh * 4;
    g.scan_len = g.nh * g.nblk + 1;
    g.max_chunks = (n * (size_t)P.W + P.L - 1) / P.L; // bound for any window group
    return g;
}

// Every device buffer one plain MSM of n points uses, with its size (one list for the workspace
// reservations and for carving the no-allocation scratch).
std::vector<std::pair<DevBuf *, size_t>> msm_buffers(Workspace &ws, const Plan &P, const MsmGeom &g, size_t n)
{
    const uint32_t per = hs::SCAN_BLOCK * hs::SCAN_ITEMS, sblk = (g.scan_len + per - 1) / per;
    return {{&ws.aff, n * hs::BASE_RECORD_BYTES},
            {&ws.isid, n},
            {&ws.digits, n * 2 * (size_t)P.W},
            {&ws.offsets, (size_t)(P.nb + 1) * 4},
            {&ws.entries, n * 4 * (size_t)P.W + 4},
            {&ws.cnt_a, (size_t)g.scan_len * 4},
            {&ws.off_a, (size_t)g.scan_len * 4},
            {&ws.sorted, n * 4 * (size_t)P.W + 4},
            {&ws.buckets, (size_t)P.nb * sizeof(hs::xyzz)},
            {&ws.seg, hs::seg_out_entries(P.W, P.B / P.G) * sizeof(hs::xyzz)},
            {&ws.win, (size_t)P.W * sizeof(hs::xyzz)},
            {&ws.big_list, (size_t)P.nb * 4},
            {&ws.big_count, 64},
            {&ws.total, sizeof(hs::jac)},
            {&ws.chunk_bucket, g.max_chunks * 4 + 4},
            {&ws.bnd, 2 * g.max_chunks * sizeof(hs::xyzz)},
            {&ws.block_sums, 2 * (size_t)(sblk + 1) * 4 + 64}};
}

// compat_n: the MSM size that decides the reference's window-8 behaviour (the whole MSM's size
// when this call is one chunk of it).  With `fixed_ws` (the no-allocation entry point) the plan is
// `*fixed_plan`, every buffer is already carved from the scratch array, and nothing that allocates
// runs: no profiling events, no auxiliary stream (one window group).
template <class C>
int msm_device(Ctx &cx, const uint32_t *d_scalars, const uint64_t *d_points, size_t n, unsigned flags, uint32_t *d_out32,
               uint64_t *d_outjac, hipStream_t st, size_t compat_n = 0, Workspace *fixed_ws = nullptr,
               const Plan *fixed_plan = nullptr)
{
    if (n >= (1ull << 24))
        return fail(HS_GPU_EINVAL, "n must be below 2^24 points per chunk");
    Workspace &ws = fixed_ws ? *fixed_ws : cx.ws(st);
    cx.last_ws = &ws;
    const Plan P = fixed_plan ? *fixed_plan : make_plan(C::id, n);
    const uint32_t nn = (uint32_t)n;
    const bool compat = !(flags & HS_GPU_EXACT) && ref_window(compat_n ? compat_n : n) == 8;
    const MsmGeom geo = msm_geom(P, n);
    const int shift = geo.shift, nsb = geo.nsb;
    const uint32_t T = geo.T, nblk = geo.nblk, nh = geo.nh;
    const size_t lds = geo.lds;
    int r;
    for (const auto &b : msm_buffers(ws, P, geo, n))
        if ((r = b.first->reserve(b.second)))
            return r;
    const uint32_t L = P.L;
    uint32_t *aff = ws.aff.as<uint32_t>();
    uint8_t *isid = ws.isid.as<uint8_t>();
    uint16_t *digits = ws.digits.as<uint16_t>();
    uint32_t *offsets = ws.offsets.as<uint32_t>(), *sorted = ws.sorted.as<uint32_t>(), *entries = ws.entries.as<uint32_t>(),
             *cnt_a = ws.cnt_a.as<uint32_t>(), *off_a = ws.off_a.as<uint32_t>();
    hs::xyzz *buckets = ws.buckets.as<hs::xyzz>(), *seg = ws.seg.as<hs::xyzz>();
    hs::jac *win = ws.win.as<hs::jac>(); // window sums, Jacobian
    uint32_t *big_list = ws.big_list.as<uint32_t>(), *big_count = ws.big_count.as<uint32_t>();

    const dim3 blk(256), grd((nn + 255) / 256);
    // Window groups (Q > 1) pipeline the combine on an auxiliary stream; with one group the whole
    // MSM stays on st (no cross-stream hop for the caller's stream to wait on).
    const int Q = fixed_ws ? 1 : P.Q < Workspace::MAXQ ? P.Q : Workspace::MAXQ;
    if (Q > 1 && (r = ws.ensure_aux()))
        return r;
    hipStream_t cst = Q > 1 ? ws.aux : st;
    if (fixed_ws)
        cx.cur = nullptr; // no profiling events on the no-allocation path
    else if ((r = cx.begin_profile(C::id)))
        return r;
    cx.mark(st, -1);
    hipLaunchKernelGGL(hs::k_ingest<C>, grd, blk, 0, st, d_points, nn, aff, isid);
    cx.mark(st, 0);
    hipLaunchKernelGGL(hs::k_recode_count<C>, dim3(nblk), dim3(hs::SORT_COARSE_THREADS), lds, st, d_scalars, isid, nn,
                       P.c, P.W, compat ? 1 : 0, nsb, shift, T, nblk, digits, cnt_a);
    cx.mark(st, 1);
    if ((r = scan_exclusive(ws, cnt_a, off_a, geo.scan_len, st)))
        return r;
    hipLaunchKernelGGL(hs::k_coarse_scatter, dim3(nblk), dim3(hs::SORT_COARSE_THREADS), lds, st, digits, nn, P.W, nsb, shift, T, nblk, off_a,
                       entries);
    hipLaunchKernelGGL(hs::k_fine_sort, dim3(nh), dim3(hs::SORT_FINE_THREADS), 0, st, entries, off_a, nblk, nsb, shift, P.B, nh, sorted,
                       offsets);
    cx.mark(st, 2);
    // Windows are processed from the top in Q groups; group q's combine (Horner over its windows)
    // runs on the combine stream while the next group's buckets are summed on st.
    const uint32_t segs_per_win = P.B / P.G;
    hs::jac *total = ws.total.as<hs::jac>();
    for (int q = 0; q < Q; q++)
    {
        const int w_hi = P.W - (P.W * q) / Q, w_lo = P.W - (P.W * (q + 1)) / Q;
        if (w_hi <= w_lo)
            continue;
        const uint32_t g_lo = (uint32_t)w_lo * P.B, g_hi = (uint32_t)w_hi * P.B, ng = g_hi - g_lo;
        const dim3 gnb((ng + 255) / 256);
        const size_t chunks = ((size_t)n * (w_hi - w_lo) + L - 1) / L;
        hipLaunchKernelGGL(hs::k_chunk_start, gnb, blk, 0, st, offsets, g_lo, g_hi, L, ws.chunk_bucket.as<uint32_t>(),
                           big_count);
        if (P.affK == 0)
            hipLaunchKernelGGL(hs::k_bucket_flat<C>, dim3((uint32_t)((chunks + 255) / 256)), blk, 0, st, offsets, sorted,
                               g_lo, g_hi, L, ws.chunk_bucket.as<uint32_t>(), aff, buckets, ws.bnd.as<hs::xyzz>());
        else
        {
            const uint32_t per = 64u * (uint32_t)P.affK, grid = (uint32_t)((chunks + per - 1) / per);
            if (P.affK == 1)
                hipLaunchKernelGGL((hs::k_bucket_affine<C, 1>), dim3(grid), dim3(64), 0, st, offsets, sorted, g_lo, g_hi, L,
                                   ws.chunk_bucket.as<uint32_t>(), aff, buckets, ws.bnd.as<hs::xyzz>());
            else if (P.affK == 2)
                hipLaunchKernelGGL((hs::k_bucket_affine<C, 2>), dim3(grid), dim3(64), 0, st, offsets, sorted, g_lo, g_hi, L,
                                   ws.chunk_bucket.as<uint32_t>(), aff, buckets, ws.bnd.as<hs::xyzz>());
            else
                hipLaunchKernelGGL((hs::k_bucket_affine<C, 4>), dim3(grid), dim3(64), 0, st, offsets, sorted, g_lo, g_hi, L,
                                   ws.chunk_bucket.as<uint32_t>(), aff, buckets, ws.bnd.as<hs::xyzz>());
        }
        cx.mark(st, 3);
        hipLaunchKernelGGL(hs::k_merge_span<C>, dim3((uint32_t)((chunks + 255) / 256)), blk, 0, st, offsets, g_lo, g_hi, L,
                           ws.chunk_bucket.as<uint32_t>(), ws.bnd.as<hs::xyzz>(), 8u, buckets, big_list, big_count);
        hipLaunchKernelGGL(hs::k_merge_long<C>, dim3(256), dim3(hs::BIG_THREADS), 0, st, offsets, g_lo, L,
                           ws.bnd.as<hs::xyzz>(), big_list, big_count, buckets);
        cx.mark(st, 6);
        const uint32_t nseg = (uint32_t)(w_hi - w_lo) * segs_per_win;
        hipLaunchKernelGGL(hs::k_reduce_seg<C>, dim3((nseg + 255) / 256), blk, 0, st, buckets, offsets, P.B, P.G,
                           (uint32_t)w_lo, nseg, seg);
        // with one window group the windows are combined in groups of wg (the Horner then steps by
        // c wg bits over ngw group sums); pipelined window groups keep one window per group
        const uint32_t nw = (uint32_t)(w_hi - w_lo), wg = Q > 1 ? 1u : hs::win_group(P.c, segs_per_win),
                       ngw = (nw + wg - 1) / wg;
        hipLaunchKernelGGL(hs::k_reduce_win<C>, dim3(ngw), dim3(64 * wg), hs::reduce_win_lds(segs_per_win, wg), st, seg,
                           segs_per_win, P.G, (uint32_t)w_lo, nw, ngw, P.c, win);
        cx.mark(st, 4);
        if (Q > 1)
        {
            HS_CHECK(hipEventRecord(ws.ev_group[q], st));
            HS_CHECK(hipStreamWaitEvent(cst, ws.ev_group[q], 0));
        }
        // group sums sit at win[w_lo .. w_lo + ngw) (wg == 1: the window sums themselves)
        const int h_hi = w_lo + (int)ngw, hc = P.c * (int)wg;
        if (g_horner_quad)
            hipLaunchKernelGGL(hs::k_horner<C>, dim3(1), dim3(64), 0, cst, win, w_lo, h_hi, hc, q == 0 ? 1 : 0, total);
        else
        {
            // the last group's kernel also finalizes (w_lo == 0 only there)
            const bool last = w_lo == 0;
            hipLaunchKernelGGL(hs::k_horner_row<C>, dim3(1), dim3(64), 0, cst, win, w_lo, h_hi, hc, q == 0 ? 1 : 0,
                               total, last ? d_out32 : nullptr, last ? d_outjac : nullptr);
        }
    }
    if (g_horner_quad)
        hipLaunchKernelGGL(hs::k_finalize<C>, dim3(1), dim3(64), 0, cst, total, d_out32, d_outjac);
    if (flags & HS_GPU_ERASE)
    {
        // the digits and the sorted entries (point index and digit sign) are functions of the scalars
        zero_dev(digits, n * 2 * (size_t)P.W, cst);
        zero_dev(entries, n * 4 * (size_t)P.W, cst);
        zero_dev(sorted, n * 4 * (size_t)P.W, cst);
    }
    if (Q > 1)
    {
        HS_CHECK(hipEventRecord(ws.ev_done, cst));
        HS_CHECK(hipStreamWaitEvent(st, ws.ev_done, 0));
    }
    cx.mark(st, 5);
    HS_CHECK(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------- batched MSMs
// Window width for MSMs of a batch (capped so one workgroup sorts a window in LDS).
int choose_c_batch(size_t n)
{
    int c = choose_c(n);
    return c > 13 ? 13 : c;
}

// Plan parameters (c, L, G) of a batch member.  A member of fewer than 2^15 points that is its
// curve's only MSM in the batch forms a group of its own, whose time is the latency of one MSM: it
// takes the plain MSM's parameters (tuning mode 0, latency of one MSM, measured on an idle device),
// with c capped at 13.  Other members take the batch parameters (mode 1, the time of many MSMs of
// that size in one batch).  Larger lone members keep the batch parameters: served alone next to
// other lanes' work, plain parameters made the server's largest requests slower (p99 over sizes up
// to 2^18: 3.03-3.10 ms with them, 2.05-2.20 ms without, same p50).
struct MemberPlan
{
    int c;
    uint32_t L, G;
};

MemberPlan member_plan(int curve, size_t n, bool alone)
{
    TuneEntry te;
    MemberPlan m;
    if (alone && n < ((size_t)1 << 15))
    {
        const bool tuned = tune_lookup(0, curve, n, te);
        m.c = tuned ? te.c : choose_c(n);
        m.L = tuned && te.L > 0 ? (uint32_t)te.L : plain_default_L(n);
        m.G = tuned ? (uint32_t)te.G : plain_default_G(n);
    }
    else
    {
        const bool tuned = tune_lookup(1, curve, n, te);
        m.c = tuned ? te.c : choose_c_batch(n);
        m.L = tuned && te.L > 0 ? (uint32_t)te.L : 32u;
        m.G = tuned ? (uint32_t)te.G : 8u;
    }
    m.c = m.c < 4 ? 4 : m.c > 13 ? 13 : m.c;
    return m;
}

// Bump allocator over one device block (the serving loop's per-lane memory).
struct Arena
{
    char *base = nullptr;
    size_t cap = 0, used = 0;
    void *take(size_t bytes)
    {
        const size_t a = (used + 255) & ~(size_t)255;
        if (a + bytes > cap)
            return nullptr;
        used = a + bytes;
        return base + a;
    }
};

struct BatchGroup
{
    int curve = 0, c = 0, W = 0;
    uint32_t B = 0, G = 0, M = 0, N = 0, V = 0, nb = 0, nseg = 0, L = 32;
    size_t chunks = 0;
    std::vector<hs::BatchDesc> hdesc;
    std::vector<size_t> index; // position of each MSM in the caller's descriptor array
    DevBuf desc, aff, isid, digits, sorted, offsets, buckets, seg, win, total, chunk_bucket, bnd, long_list, long_count;
};

} // namespace

struct hs_gpu_batch
{
    std::vector<BatchGroup *> groups;
    unsigned flags = 0;
    bool in_arena = false;   // device buffers carved from a caller's arena (not freed here)
    bool borrowed = false;   // worker streams and events belong to the caller
    std::vector<hipStream_t> workers;
    std::vector<hipEvent_t> ev;
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    size_t total_points = 0;
    uint32_t n_empty = 0; // MSMs with n = 0: identity outputs
    DevBuf empty;
    std::vector<hs::BatchDesc> hempty;
};

namespace
{

template <class C> int batch_group_enqueue(BatchGroup &g, hipStream_t st)
{
    const dim3 blk(256);
    const uint32_t N = g.N;
    hs::BatchDesc *d = g.desc.as<hs::BatchDesc>();
    uint32_t *aff = g.aff.as<uint32_t>(), *sorted = g.sorted.as<uint32_t>(), *offsets = g.offsets.as<uint32_t>();
    hs::xyzz *buckets = g.buckets.as<hs::xyzz>(), *seg = g.seg.as<hs::xyzz>();
    hs::jac *win = g.win.as<hs::jac>(); // window sums, Jacobian
    hipLaunchKernelGGL(hs::k_ingest_batch<C>, dim3((N + 255) / 256), blk, 0, st, d, g.M, N, aff, g.isid.as<uint8_t>());
    hipLaunchKernelGGL(hs::k_recode_batch<C>, dim3((N + 255) / 256), blk, 0, st, d, g.M, N, g.isid.as<uint8_t>(), g.c, g.W,
                       g.digits.as<uint16_t>());
    hipLaunchKernelGGL(hs::k_window_sort, dim3(g.V), blk, 0, st, d, N, g.W, g.B, g.digits.as<uint16_t>(), sorted, offsets,
                       g.V);
    const dim3 gnb((g.nb + 255) / 256);
    hipLaunchKernelGGL(hs::k_chunk_start, gnb, blk, 0, st, offsets, 0u, g.nb, g.L, g.chunk_bucket.as<uint32_t>(),
                       g.long_count.as<uint32_t>());
    hipLaunchKernelGGL(hs::k_bucket_flat<C>, dim3((uint32_t)((g.chunks + 255) / 256)), blk, 0, st, offsets, sorted, 0u,
                       g.nb, g.L, g.chunk_bucket.as<uint32_t>(), aff, buckets, g.bnd.as<hs::xyzz>());
    hipLaunchKernelGGL(hs::k_merge_span<C>, dim3((uint32_t)((g.chunks + 255) / 256)), blk, 0, st, offsets, 0u, g.nb, g.L,
                       g.chunk_bucket.as<uint32_t>(), g.bnd.as<hs::xyzz>(), 8u, buckets, g.long_list.as<uint32_t>(),
                       g.long_count.as<uint32_t>());
    hipLaunchKernelGGL(hs::k_merge_long<C>, dim3(128), dim3(hs::BIG_THREADS), 0, st, offsets, 0u, g.L,
                       g.bnd.as<hs::xyzz>(), g.long_list.as<uint32_t>(), g.long_count.as<uint32_t>(), buckets);
    hipLaunchKernelGGL(hs::k_reduce_seg<C>, dim3((g.nseg + 255) / 256), blk, 0, st, buckets, offsets, g.B, g.G, 0u, g.nseg,
                       seg);
    // each MSM's windows combined in groups of wg: ngw group sums per MSM, Horner steps of c wg bits
    const uint32_t spw = g.B / g.G, wg = hs::win_group(g.c, spw), ngw = ((uint32_t)g.W + wg - 1) / wg;
    hipLaunchKernelGGL(hs::k_reduce_win<C>, dim3(g.M * ngw), dim3(64 * wg), hs::reduce_win_lds(spw, wg), st, seg, spw,
                       g.G, 0u, (uint32_t)g.W, ngw, g.c, win);
    if (g_horner_quad)
    {
        hipLaunchKernelGGL(hs::k_horner_batch<C>, dim3((4 * g.M + 63) / 64), dim3(64), 0, st, win, g.M, (int)ngw,
                           g.c * (int)wg, g.total.as<hs::jac>());
        hipLaunchKernelGGL(hs::k_finalize_batch<C>, dim3((g.M + 63) / 64), dim3(64), 0, st, g.total.as<hs::jac>(), d,
                           g.M);
    }
    else
        hipLaunchKernelGGL(hs::k_horner_batch_row<C>, dim3(g.M), dim3(64), 0, st, win, g.M, (int)ngw, g.c * (int)wg, d);
    HS_CHECK(hipGetLastError());
    return 0;
}

// Enqueues every group of the batch: fork from st onto the worker streams, join back into st (a
// single group runs on st).  Works eagerly and under stream capture alike.
int batch_enqueue(hs_gpu_batch &b, hipStream_t st)
{
    const size_t K = b.workers.size();
    if (b.n_empty)
        hipLaunchKernelGGL(hs::k_identity_out, dim3((b.n_empty + 63) / 64), dim3(64), 0, st,
                           b.empty.as<hs::BatchDesc>(), b.n_empty);
    // one group has nothing to overlap with: it runs on st itself, without the fork and join
    if (b.groups.size() == 1)
    {
        BatchGroup &g = *b.groups[0];
        return g.curve == HS_GPU_HELIOS ? batch_group_enqueue<hs::HeliosC>(g, st) : batch_group_enqueue<hs::SeleneC>(g, st);
    }
    HS_CHECK(hipEventRecord(b.ev[0], st));
    for (size_t k = 0; k < K; k++)
        HS_CHECK(hipStreamWaitEvent(b.workers[k], b.ev[0], 0));
    for (size_t i = 0; i < b.groups.size(); i++)
    {
        BatchGroup &g = *b.groups[i];
        hipStream_t ws = b.workers[i % K];
        int r = g.curve == HS_GPU_HELIOS ? batch_group_enqueue<hs::HeliosC>(g, ws) : batch_group_enqueue<hs::SeleneC>(g, ws);
        if (r)
            return r;
    }
    for (size_t k = 0; k < K; k++)
    {
        HS_CHECK(hipEventRecord(b.ev[1 + k], b.workers[k]));
        HS_CHECK(hipStreamWaitEvent(st, b.ev[1 + k], 0));
    }
    return 0;
}

// Entry slots, bucket ids and point indices of one batch group stay below this (bit 31 of an entry
// is the digit sign, 0xffffffff marks padding).
constexpr size_t BATCH_GROUP_LIMIT = (size_t)1 << 31;

// Plans a batch.  With `arena`, every device buffer is carved from it and descriptors are
// uploaded asynchronously on `upload` (their host copies live in the plan); with `workers`
// (and `events`, one more than workers) the plan runs on the caller's streams.
int batch_create(hs_gpu_batch **out, const hs_gpu_msm_desc *descs, size_t count, unsigned flags, Arena *arena = nullptr,
                 hipStream_t upload = nullptr, const std::vector<hipStream_t> *workers = nullptr,
                 const std::vector<hipEvent_t> *events = nullptr)
{
    Ctx &cx = ctx();
    int r;
    {
        std::lock_guard<std::mutex> lk(cx.mu);
        if ((r = cx.init()))
            return r;
    }
    auto *b = new hs_gpu_batch();
    b->flags = flags;
    b->in_arena = arena != nullptr;
    // group by (curve, window width).  Every group is one chain of kernels whose latency (the
    // combine's 256 doublings above all) does not shrink with the group's size, so window widths
    // are merged into at most HS_GPU_BATCH_GROUPS (default 2) groups per curve: each width maps
    // to the largest width of its share of the curve's sorted distinct widths.  Batches planned
    // inside a serving arena keep their own widths (the admission footprint assumes them).
    size_t per_curve[2] = {0, 0};
    for (size_t i = 0; i < count; i++)
        if (descs[i].n != 0 && descs[i].n < (1u << 24) &&
            (descs[i].curve == HS_GPU_HELIOS || descs[i].curve == HS_GPU_SELENE))
            per_curve[descs[i].curve]++;
    std::vector<int> want_c(count, 0);
    std::vector<MemberPlan> mp(count);
    for (size_t i = 0; i < count; i++)
    {
        if (descs[i].n == 0 || descs[i].n >= (1u << 24) ||
            (descs[i].curve != HS_GPU_HELIOS && descs[i].curve != HS_GPU_SELENE))
            continue;
        mp[i] = member_plan(descs[i].curve, descs[i].n, per_curve[descs[i].curve] == 1);
        want_c[i] = mp[i].c;
    }
    const int max_groups = arena ? 64 : std::max(1, env_int("HS_GPU_BATCH_GROUPS", 2));
    int cmap[2][14] = {};
    for (int cv = 0; cv < 2; cv++)
    {
        std::vector<int> cs;
        for (size_t i = 0; i < count; i++)
            if (want_c[i] && descs[i].curve == cv && std::find(cs.begin(), cs.end(), want_c[i]) == cs.end())
                cs.push_back(want_c[i]);
        std::sort(cs.begin(), cs.end());
        const size_t k = std::min(cs.size(), (size_t)max_groups);
        for (size_t j = 0; j < cs.size(); j++)
        {
            const size_t part = j * k / cs.size();              // which of the k shares
            const size_t last = ((part + 1) * cs.size() + k - 1) / k - 1; // its largest member
            cmap[cv][cs[j]] = cs[std::min(last, cs.size() - 1)];
        }
    }
    std::vector<std::pair<int, int>> keys;
    std::vector<hs::BatchDesc> empties;
    // HS_GPU_BATCH_GROUP_LIMIT lowers the group limit (tests of the split at small sizes)
    const size_t lim_env = env_size("HS_GPU_BATCH_GROUP_LIMIT");
    const size_t group_limit = lim_env && lim_env < BATCH_GROUP_LIMIT ? lim_env : BATCH_GROUP_LIMIT;
    for (size_t i = 0; i < count; i++)
    {
        const hs_gpu_msm_desc &d = descs[i];
        if ((d.curve != HS_GPU_HELIOS && d.curve != HS_GPU_SELENE) || (d.n && (!d.scalars || !d.points)))
        {
            hs_gpu_batch_destroy(b);
            return fail(HS_GPU_EINVAL, "bad batch descriptor " + std::to_string(i));
        }
        if (d.n == 0)
        {
            hs::BatchDesc e{};
            e.out32 = (uint32_t *)d.out32;
            e.outjac = (uint64_t *)d.result;
            empties.push_back(e);
            continue;
        }
        if (d.n >= (1u << 24))
        {
            hs_gpu_batch_destroy(b);
            return fail(HS_GPU_EINVAL, "batch element " + std::to_string(i) + ": n >= 2^24 (the plain MSM takes larger sizes)");
        }
        const int c = cmap[d.curve][want_c[i]];
        const size_t Wc = (256 + c - 1) / c, Bc = (size_t)1 << (c - 1);
        // A group's entry slots (N * W), bucket ids (M * W * B) and point indices are 32-bit values
        // with bit 31 kept for the digit sign, so a group holds fewer than 2^31 of each: a member
        // that would pass that limit starts another group of the same (curve, width).
        BatchGroup *g = nullptr;
        for (size_t k = keys.size(); k-- > 0;)
            if (keys[k].first == d.curve && keys[k].second == c)
            {
                BatchGroup *h = b->groups[k];
                if (((size_t)h->N + d.n) * Wc < group_limit && ((size_t)h->hdesc.size() + 1) * Wc * Bc < group_limit)
                    g = h;
                break;
            }
        if (!g)
        {
            keys.emplace_back(d.curve, c);
            g = new BatchGroup();
            g->curve = d.curve;
            g->c = c;
            g->W = (int)Wc;
            g->B = (uint32_t)Bc;
            g->G = 0;
            g->L = 0;
            b->groups.push_back(g);
        }
        // the group runs with the largest run length and segment size of its members: its buffers
        // (chunk bounds, segment sums) are then no larger than the members' footprints together,
        // which is what the serving loop admits by
        g->G = std::max(g->G, mp[i].G);
        g->L = std::max(g->L, mp[i].L);
        hs::BatchDesc bd;
        bd.scalars = (const uint32_t *)d.scalars;
        bd.points = (const uint64_t *)d.points;
        bd.out32 = (uint32_t *)d.out32;
        bd.outjac = (uint64_t *)d.result;
        bd.n = (uint32_t)d.n;
        bd.p_off = g->N;
        bd.w8_compat = (!(flags & HS_GPU_EXACT) && ref_window(d.n) == 8) ? 1u : 0u;
        bd.pad = 0;
        g->hdesc.push_back(bd);
        g->index.push_back(i);
        g->N += (uint32_t)d.n;
        b->total_points += d.n;
    }
    for (BatchGroup *g : b->groups)
    {
        g->G = plan_G(g->B, g->G);
        g->L = std::max(1u, g->L);
        g->M = (uint32_t)g->hdesc.size();
        g->V = g->M * (uint32_t)g->W;
        g->nb = g->V * g->B;
        g->nseg = g->nb / g->G;
        g->chunks = ((size_t)g->N * g->W + g->L - 1) / g->L;
        const size_t E = (size_t)g->N * g->W;
        const std::pair<DevBuf *, size_t> bufs[] = {
            {&g->desc, g->M * sizeof(hs::BatchDesc)},
            {&g->aff, (size_t)g->N * hs::BASE_RECORD_BYTES},
            {&g->isid, g->N},
            {&g->digits, E * 2},
            {&g->sorted, E * 4 + 4},
            {&g->offsets, ((size_t)g->nb + 1) * 4},
            {&g->buckets, (size_t)g->nb * sizeof(hs::xyzz)},
            {&g->seg, hs::seg_out_entries(g->V, g->B / g->G) * sizeof(hs::xyzz)},
            {&g->win, (size_t)g->V * sizeof(hs::xyzz)},
            {&g->total, (size_t)g->M * sizeof(hs::jac)},
            {&g->chunk_bucket, g->chunks * 4 + 4},
            {&g->bnd, 2 * g->chunks * sizeof(hs::xyzz)},
            {&g->long_list, (size_t)g->nb * 4},
            {&g->long_count, 64}};
        for (const auto &bf : bufs)
        {
            if (arena)
            {
                bf.first->p = arena->take(bf.second);
                bf.first->cap = bf.second;
                r = bf.first->p ? 0 : fail(HS_GPU_ENOMEM, "batch does not fit its device arena");
            }
            else
                r = bf.first->reserve(bf.second);
            if (r)
            {
                hs_gpu_batch_destroy(b);
                return r;
            }
        }
        if (upload)
            HS_CHECK(hipMemcpyAsync(g->desc.p, g->hdesc.data(), g->M * sizeof(hs::BatchDesc), hipMemcpyHostToDevice, upload));
        else
            HS_CHECK(hipMemcpy(g->desc.p, g->hdesc.data(), g->M * sizeof(hs::BatchDesc), hipMemcpyHostToDevice));
    }
    if (!empties.empty())
    {
        b->n_empty = (uint32_t)empties.size();
        b->hempty = empties;
        const size_t eb = empties.size() * sizeof(hs::BatchDesc);
        if (arena)
        {
            b->empty.p = arena->take(eb);
            b->empty.cap = eb;
            r = b->empty.p ? 0 : fail(HS_GPU_ENOMEM, "batch does not fit its device arena");
        }
        else
            r = b->empty.reserve(eb);
        if (r)
        {
            hs_gpu_batch_destroy(b);
            return r;
        }
        if (upload)
            HS_CHECK(hipMemcpyAsync(b->empty.p, b->hempty.data(), eb, hipMemcpyHostToDevice, upload));
        else
            HS_CHECK(hipMemcpy(b->empty.p, b->hempty.data(), eb, hipMemcpyHostToDevice));
    }
    if (workers)
    {
        // the caller's streams and events (at least one worker, workers + 1 events)
        b->borrowed = true;
        b->workers = *workers;
        b->ev = *events;
        *out = b;
        return 0;
    }
    const size_t K = b->groups.size() < 4 ? (b->groups.empty() ? 1 : b->groups.size()) : 4;
    b->workers.resize(K);
    b->ev.resize(K + 1);
    for (size_t k = 0; k < K; k++)
        HS_CHECK(hipStreamCreateWithFlags(&b->workers[k], hipStreamNonBlocking));
    for (size_t k = 0; k < K + 1; k++)
        HS_CHECK(hipEventCreateWithFlags(&b->ev[k], hipEventDisableTiming));
    *out = b;
    return 0;
}

const char *const PHASE_NAMES[] = {"ingest", "recode", "sort", "bucket", "reduce", "combine", "merge"};

template <class C>
int map_to_curve_device(Ctx &cx, const uint32_t *d_u, size_t n, uint64_t *d_jac, uint32_t *d_out32, hipStream_t st)
{
    (void)cx;
    const uint32_t nn = (uint32_t)n;
    hipLaunchKernelGGL(hs::k_map_to_curve<C>, dim3((nn + 255) / 256), dim3(256), 0, st, d_u, nn, d_jac, d_out32);
    HS_CHECK(hipGetLastError());
    return 0;
}

template <class C>
int decompress_device(Ctx &cx, const uint32_t *d_in32, size_t n, uint64_t *d_jac, hipStream_t st, uint32_t *h_nbad)
{
    int r;
    if ((r = cx.nbad.reserve(64)))
        return r;
    const uint32_t nn = (uint32_t)n;
    HS_CHECK(hipMemsetAsync(cx.nbad.p, 0, 4, st));
    hipLaunchKernelGGL(hs::k_decompress<C>, dim3((nn + 255) / 256), dim3(256), 0, st, d_in32, nn, d_jac,
                       (uint8_t *)nullptr, cx.nbad.as<uint32_t>());
    HS_CHECK(hipGetLastError());
    HS_CHECK(hipMemcpyAsync(h_nbad, cx.nbad.p, 4, hipMemcpyDeviceToHost, st));
    return 0;
}

// MSMs of 2^24 points or more: chunks of 2^23 points, each leaving its sum (reference struct) in
// the stream's workspace, then one kernel adds the chunk sums and encodes the total.
template <class C>
int msm_device_chunked(Ctx &cx, const uint32_t *s, const uint64_t *p, size_t n, unsigned flags, uint32_t *o32,
                       uint64_t *oj, hipStream_t st)
{
    Workspace &ws = cx.ws(st);
    // HS_GPU_CHUNK (points per chunk, for tests of this path at small n) overrides the default
    const size_t chunk = std::min<size_t>(MSM_MAX_CHUNK, std::max<size_t>(1024, (size_t)env_int("HS_GPU_CHUNK", (int)MSM_MAX_CHUNK)));
    const size_t k = (n + chunk - 1) / chunk;
    int r;
    if ((r = ws.parts.reserve(k * (15 * 8 + 32))))
        return r;
    uint64_t *parts = ws.parts.as<uint64_t>();
    uint32_t *scratch32 = (uint32_t *)(parts + 15 * k);
    for (size_t j = 0; j < k; j++)
    {
        const size_t off = j * chunk, len = std::min(chunk, n - off);
        if ((r = msm_device<C>(cx, s + 8 * off, p + 15 * off, len, flags, scratch32 + 8 * j, parts + 15 * j, st, n)))
            return r;
    }
    hipLaunchKernelGGL(hs::k_sum_parts<C>, dim3(1), dim3(64), 0, st, parts, (uint32_t)k, o32, oj);
    HS_CHECK(hipGetLastError());
    return 0;
}

int msm_device_any(int curve, Ctx &cx, const void *s, const void *p, size_t n, unsigned flags, void *o32, void *oj,
                   hipStream_t st)
{
    const size_t chunk = (size_t)env_int("HS_GPU_CHUNK", 0);
    if (n >= ((size_t)1 << 24) || (chunk >= 1024 && n > chunk))
        return curve == HS_GPU_HELIOS
                   ? msm_device_chunked<hs::HeliosC>(cx, (const uint32_t *)s, (const uint64_t *)p, n, flags,
                                                     (uint32_t *)o32, (uint64_t *)oj, st)
                   : msm_device_chunked<hs::SeleneC>(cx, (const uint32_t *)s, (const uint64_t *)p, n, flags,
                                                     (uint32_t *)o32, (uint64_t *)oj, st);
    if (curve == HS_GPU_HELIOS)
        return msm_device<hs::HeliosC>(cx, (const uint32_t *)s, (const uint64_t *)p, n, flags, (uint32_t *)o32,
                                       (uint64_t *)oj, st);
    if (curve == HS_GPU_SELENE)
        return msm_device<hs::SeleneC>(cx, (const uint32_t *)s, (const uint64_t *)p, n, flags, (uint32_t *)o32,
                                       (uint64_t *)oj, st);
    return fail(HS_GPU_EINVAL, "unknown curve id");
}

void set_identity(hs_gpu_jacobian *r)
{
    std::memset(r, 0, sizeof *r);
    r->X[0] = 1;
    r->Y[0] = 1;
}

int msm_host(int curve, hs_gpu_jacobian *result, unsigned char *out32, const unsigned char *scalars,
             const hs_gpu_jacobian *points, size_t n, unsigned flags)
{
    if (curve != HS_GPU_HELIOS && curve != HS_GPU_SELENE)
        return fail(HS_GPU_EINVAL, "unknown curve id");
    if (n == 0)
    {
        if (result)
            set_identity(result);
        if (out32)
            std::memset(out32, 0, 32);
        return 0;
    }
    if (!scalars || !points)
        return fail(HS_GPU_EINVAL, "null input with n > 0");
    if (n > SIZE_MAX / sizeof(hs_gpu_jacobian))
        return fail(HS_GPU_EINVAL, "n too large");
    Ctx &cx = ctx();
    HostStage *hsg = nullptr;
    int r = 0;
    {
        std::lock_guard<std::mutex> lk(cx.mu);
        if ((r = cx.init()))
            return r;
        if (!(hsg = take_stage(cx)))
            return HS_GPU_EDEVICE;
        const size_t ob = 32 + sizeof(hs_gpu_jacobian);
        if (!(r = hsg->scalars.reserve(n * 32)) && !(r = hsg->points.reserve(n * sizeof(hs_gpu_jacobian))))
            r = hsg->out.reserve(ob);
        if (r)
        {
            cx.stages_free.push_back(hsg);
            return r;
        }
    }
    // upload outside the lock (the staging slot is this call's alone), so concurrent callers copy
    // in parallel; launch under the lock (workspace lookup, profiling state); wait outside it
    hipStream_t st = hsg->st;
    r = hsg->upload(hsg->scalars.p, scalars, n * 32);
    if (!r)
        r = hsg->upload(hsg->points.p, points, n * sizeof(hs_gpu_jacobian));
    if (flags & HS_GPU_ERASE)
        hsg->erase_pinned();
    {
        std::lock_guard<std::mutex> lk(cx.mu);
        unsigned char *o = hsg->out.as<unsigned char>();
        if (!r)
            r = msm_device_any(curve, cx, hsg->scalars.p, hsg->points.p, n, flags, o, o + 32, st);
        if (flags & HS_GPU_ERASE)
            zero_dev(hsg->scalars.p, n * 32, st);
        if (r)
        {
            hipStreamSynchronize(hsg->st);
            cx.stages_free.push_back(hsg);
            return r;
        }
    }
    unsigned char o[32 + sizeof(hs_gpu_jacobian)];
    if (hipMemcpyAsync(o, hsg->out.p, sizeof o, hipMemcpyDeviceToHost, hsg->st) != hipSuccess)
        r = fail(HS_GPU_EDEVICE, "result copy failed");
    const hipError_t e = hipStreamSynchronize(hsg->st);
    if (!r && e != hipSuccess)
        r = fail(HS_GPU_EDEVICE, std::string("MSM failed: ") + hipGetErrorString(e));
    {
        std::lock_guard<std::mutex> lk(cx.mu);
        cx.stages_free.push_back(hsg);
    }
    if (r)
        return r;
    if (out32)
        std::memcpy(out32, o, 32);
    if (result)
        std::memcpy(result, o + 32, sizeof(hs_gpu_jacobian));
    return 0;
}

// Registered CPU fallbacks of the void drop-ins, per curve.
std::atomic<hs_gpu_cpu_msm_fn> g_fb_msm[2];
std::atomic<hs_gpu_cpu_msm_fixed_fn> g_fb_fixed[2];
std::atomic<uint64_t> g_fb_calls{0};

[[noreturn]] void die(const char *what, int curve, int r)
{
    std::fprintf(stderr,
                 "hs_gpu %s_%s failed (%d): %s; no CPU fallback is registered (hs_gpu_set_cpu_fallback), so the "
                 "process stops here rather than return a wrong point\n",
                 curve == 0 ? "helios" : "selene", what, r, g_err.c_str());
    std::fflush(stderr);
    std::abort();
}

void slot_flags(int curve, hs_gpu_jacobian *result, const unsigned char *scalars, const hs_gpu_jacobian *points,
                size_t n, unsigned flags)
{
    const int r = msm_host(curve, result, nullptr, scalars, points, n, flags);
    if (r == 0)
        return;
    if (curve != HS_GPU_HELIOS && curve != HS_GPU_SELENE)
        die("msm_vartime", curve, r);
    if (hs_gpu_cpu_msm_fn fb = g_fb_msm[curve].load())
    {
        g_fb_calls++;
        fb(result, scalars, points, n);
        return;
    }
    die("msm_vartime", curve, r);
}

void slot(int curve, hs_gpu_jacobian *result, const unsigned char *scalars, const hs_gpu_jacobian *points, size_t n)
{
    slot_flags(curve, result, scalars, points, n, 0);
}

} // namespace

// ---------------------------------------------------------------- fixed-base sets
struct hs_gpu_fixed
{
    int curve = 0;
    uint32_t n = 0;
    DevBuf tables, points; // n x 16 canonical affine entries; n reference Jacobian structs
    int wb = 5;            // table layout: window width (n x 2^(wb-1) entries) or hs::FIXED_COMB16
    DevBuf scal, sums;     // per-call workspace: prepared scalars, window sums
    hs_gpu_batch *pip = nullptr; // cached Pippenger plan (method HS_GPU_FIXED_PIPPENGER)
    const void *pip_s = nullptr;
    void *pip_o = nullptr, *pip_j = nullptr;
    size_t pip_K = 0;
    unsigned pip_flags = 0;
    std::mutex mu;
};

namespace
{

// The n bases of a fixed set must fit the 32-bit table indexing and the batch plan limits.
constexpr size_t FIXED_MAX_BASES = (size_t)1 << 24;

int fixed_alloc(hs_gpu_fixed **out, int curve, size_t n)
{
    if (curve != HS_GPU_HELIOS && curve != HS_GPU_SELENE)
        return fail(HS_GPU_EINVAL, "unknown curve id");
    if (n == 0 || n >= FIXED_MAX_BASES)
        return fail(HS_GPU_EINVAL, "fixed-base set needs 0 < n < 2^24");
    Ctx &cx = ctx();
    int r;
    {
        std::lock_guard<std::mutex> lk(cx.mu);
        if ((r = cx.init()))
            return r;
    }
    auto *f = new hs_gpu_fixed();
    f->curve = curve;
    f->n = (uint32_t)n;
    if ((r = f->points.reserve_exact(n * sizeof(hs_gpu_jacobian))))
    {
        hs_gpu_fixed_destroy(f);
        return r;
    }
    *out = f;
    return 0;
}

// Tables of a set built from points: the 16-tooth comb (2^15 entries per base, width code
// hs::FIXED_COMB16) when its tables (2 MiB per base) fit FIXED_TABLE_BUDGET, else the widest window
// of 14, 13, 12, 10 and 8 bits whose 2^(w-1) entries of 64 B per base fit, else 5 bits.  Fewer
// windows mean fewer additions per base and MSM (16 for the comb instead of 52); the table
// gathers are random 64-byte reads that HBM serves well.  2^12 + 1 bases get the comb (8 GiB),
// 2^20 bases width 8.  Entry indices stay below 2^32.
constexpr size_t FIXED_TABLE_BUDGET = (size_t)9 << 30;

int fixed_width(size_t n)
{
    // HS_GPU_FIXED_WIDTH forces a layout that fits (tests of every width at small n): 16 = comb
    const int forced = env_int("HS_GPU_FIXED_WIDTH", 0);
    if (forced == 16 && n * ((size_t)64 << 15) <= FIXED_TABLE_BUDGET)
        return hs::FIXED_COMB16;
    for (int wb : {14, 13, 12, 10, 8, 5})
        if (forced == wb && n * ((size_t)64 << (wb - 1)) <= FIXED_TABLE_BUDGET)
            return wb;
    if (n * ((size_t)64 << 15) <= FIXED_TABLE_BUDGET)
        return hs::FIXED_COMB16;
    for (int wb : {14, 13, 12, 10, 8})
        if (n * ((size_t)64 << (wb - 1)) <= FIXED_TABLE_BUDGET)
            return wb;
    return 5;
}

// Comb tables: tooth bases 2^(16 j) P_i (staged in the per-call window-sum workspace), then every
// entry from them.
template <class C> int fixed_build_comb(hs_gpu_fixed &f, hipStream_t st)
{
    constexpr int E = hs::FixedW<hs::FIXED_COMB16>::entries;
    int r;
    // the per-call scalar workspace stages the affine bases, then (once the teeth are built from
    // them) the half tables, with the quarter tables behind them
    if ((r = f.tables.reserve_exact((size_t)f.n * E * 64)) ||
        (r = f.scal.reserve((size_t)f.n * (hs::COMB_HALF + hs::COMB_QUARTER) * 64)) ||
        (r = f.sums.reserve((size_t)f.n * 16 * 64)))
        return r;
    uint32_t *quarters = f.scal.as<uint32_t>() + (size_t)f.n * hs::COMB_HALF * 16;
    hipLaunchKernelGGL(hs::k_fixed_bases<C>, dim3((f.n + 255) / 256), dim3(256), 0, st, f.points.as<uint64_t>(), f.n,
                       f.scal.as<uint32_t>());
    hipLaunchKernelGGL(hs::k_comb_teeth<C>, dim3((16 * f.n + 255) / 256), dim3(256), 0, st, f.scal.as<uint32_t>(),
                       f.n, f.sums.as<uint32_t>());
    const uint32_t tq = f.n * hs::COMB_QUARTER, th = f.n * 24u, t = f.n * (128u * (256u / hs::COMB_RUN));
    hipLaunchKernelGGL(hs::k_comb_quarters<C>, dim3((tq + 255) / 256), dim3(256), 0, st, f.sums.as<uint32_t>(), f.n,
                       quarters);
    hipLaunchKernelGGL(hs::k_comb_halves<C>, dim3((th + 255) / 256), dim3(256), 0, st, quarters, f.n,
                       f.scal.as<uint32_t>());
    hipLaunchKernelGGL(hs::k_comb_tables_runs<C>, dim3((t + 255) / 256), dim3(256), 0, st, f.scal.as<uint32_t>(), f.n,
                       f.tables.as<uint32_t>());
    HS_CHECK(hipGetLastError());
    f.wb = hs::FIXED_COMB16;
    return 0;
}

template <class C, int WB> int fixed_build_tables_w(hs_gpu_fixed &f, uint64_t *d_ref_out, hipStream_t st)
{
    constexpr int E = hs::FixedW<WB>::entries;
    int r;
    // the affine bases are staged in the per-call scalar workspace
    if ((r = f.tables.reserve_exact((size_t)f.n * E * 64)) || (r = f.scal.reserve((size_t)f.n * 64)))
        return r;
    hipLaunchKernelGGL(hs::k_fixed_bases<C>, dim3((f.n + 255) / 256), dim3(256), 0, st, f.points.as<uint64_t>(), f.n,
                       f.scal.as<uint32_t>());
    const uint32_t t = f.n * (uint32_t)E;
    DevBuf halves; // wide tables: entries as sums over two half tables (freed after the build)
    bool use_halves = false;
    if (WB >= 8 && !d_ref_out)
    {
        const std::string err = g_err;
        use_halves = halves.reserve_exact((size_t)f.n * hs::win_half(E) * 64) == 0;
        if (!use_halves)
        {
            // no room for the half tables: the entries are built one by one instead, which is not
            // an error (clear what the refused allocation recorded)
            (void)hipGetLastError();
            g_err = err;
        }
    }
    if (use_halves)
    {
        const uint32_t th = f.n * hs::win_half(E), tr = f.n * (E / 32 + 1) * (32u / hs::COMB_RUN);
        hipLaunchKernelGGL((hs::k_win_halves<C, E>), dim3((th + 255) / 256), dim3(256), 0, st, f.scal.as<uint32_t>(),
                           f.n, halves.as<uint32_t>());
        hipLaunchKernelGGL((hs::k_win_tables_runs<C, E>), dim3((tr + 255) / 256), dim3(256), 0, st,
                           halves.as<uint32_t>(), f.n, f.tables.as<uint32_t>());
        const hipError_t e1 = hipGetLastError(), e2 = hipStreamSynchronize(st);
        halves.release();
        if (e1 != hipSuccess || e2 != hipSuccess)
            return fail(HS_GPU_EDEVICE, "fixed-base table build failed");
    }
    else
    {
        // one multiplication and one inversion per entry (the width-5 tables, or no room for the
        // half tables)
        hipLaunchKernelGGL((hs::k_fixed_tables<C, E>), dim3((t + 255) / 256), dim3(256), 0, st, f.scal.as<uint32_t>(),
                           f.n, f.tables.as<uint32_t>(), d_ref_out);
        HS_CHECK(hipGetLastError());
    }
    f.wb = WB;
    return 0;
}

// Tables of a set from its points.  With d_ref_out (the reference-format precompute) they are the
// width-5 tables; otherwise the width is fixed_width(n).
template <class C> int fixed_build_tables(hs_gpu_fixed &f, uint64_t *d_ref_out, hipStream_t st)
{
    switch (d_ref_out ? 5 : fixed_width(f.n))
    {
    case hs::FIXED_COMB16:
        return fixed_build_comb<C>(f, st);
    case 14:
        return fixed_build_tables_w<C, 14>(f, nullptr, st);
    case 13:
        return fixed_build_tables_w<C, 13>(f, nullptr, st);
    case 12:
        return fixed_build_tables_w<C, 12>(f, nullptr, st);
    case 10:
        return fixed_build_tables_w<C, 10>(f, nullptr, st);
    case 8:
        return fixed_build_tables_w<C, 8>(f, nullptr, st);
    default:
        return fixed_build_tables_w<C, 5>(f, d_ref_out, st);
    }
}

// windows (or comb columns) summed per wavefront: 2 and 4 measured alike, both 2.5% ahead of 1
// (a 64-lane tree per window costs ~10% of the additions)
constexpr int FIXED_CPW = 2;

// count MSMs over the set; comb method (width-WB tables, shared doublings)
template <class C, int WB>
int fixed_comb_enqueue(hs_gpu_fixed &f, const void *d_scalars, size_t count, unsigned flags, void *d_out32,
                       void *d_outjac, hipStream_t st)
{
    constexpr int W = hs::FixedW<WB>::windows;
    int r;
    const size_t rows = count * f.n;
    if ((r = f.scal.reserve(rows * 32)) || (r = f.sums.reserve(count * W * sizeof(hs::jac))))
        return r;
    const uint32_t compat = (!(flags & HS_GPU_EXACT) && ref_window(f.n) == 8) ? 1u : 0u;
    hipLaunchKernelGGL(hs::k_fixed_prep<C>, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, st,
                       (const uint32_t *)d_scalars, rows, f.scal.as<uint32_t>(), compat,
                       WB == hs::FIXED_COMB16 ? 1u : 0u);
    hipLaunchKernelGGL((hs::k_fixed_window_sums<C, WB, FIXED_CPW>), dim3((unsigned)(count * ((W + FIXED_CPW - 1) / FIXED_CPW))), dim3(64), 0, st,
                       f.tables.as<uint32_t>(), f.scal.as<uint32_t>(), f.n, f.sums.as<hs::jac>());
    hipLaunchKernelGGL((hs::k_fixed_horner<C, WB>), dim3((unsigned)((4 * count + 63) / 64)), dim3(64), 0, st,
                       f.sums.as<hs::jac>(), (uint32_t)count, (uint32_t *)d_out32, (uint64_t *)d_outjac);
    HS_CHECK(hipGetLastError());
    return 0;
}

template <class C>
int fixed_comb_dispatch(hs_gpu_fixed &f, const void *d_scalars, size_t count, unsigned flags, void *d_out32,
                        void *d_outjac, hipStream_t st)
{
    switch (f.wb)
    {
    case hs::FIXED_COMB16:
        return fixed_comb_enqueue<C, hs::FIXED_COMB16>(f, d_scalars, count, flags, d_out32, d_outjac, st);
    case 14:
        return fixed_comb_enqueue<C, 14>(f, d_scalars, count, flags, d_out32, d_outjac, st);
    case 13:
        return fixed_comb_enqueue<C, 13>(f, d_scalars, count, flags, d_out32, d_outjac, st);
    case 12:
        return fixed_comb_enqueue<C, 12>(f, d_scalars, count, flags, d_out32, d_outjac, st);
    case 10:
        return fixed_comb_enqueue<C, 10>(f, d_scalars, count, flags, d_out32, d_outjac, st);
    case 8:
        return fixed_comb_enqueue<C, 8>(f, d_scalars, count, flags, d_out32, d_outjac, st);
    default:
        return fixed_comb_enqueue<C, 5>(f, d_scalars, count, flags, d_out32, d_outjac, st);
    }
}

// Pippenger method: one batched plan whose MSMs all read the resident bases
int fixed_pippenger_enqueue(hs_gpu_fixed &f, const void *d_scalars, size_t count, unsigned flags, void *d_out32,
                            void *d_outjac, hipStream_t st)
{
    if (!f.pip || f.pip_s != d_scalars || f.pip_K != count || f.pip_o != d_out32 || f.pip_j != d_outjac ||
        f.pip_flags != flags)
    {
        if (f.pip)
        {
            HS_CHECK(hipStreamSynchronize(st));
            hs_gpu_batch_destroy(f.pip);
            f.pip = nullptr;
        }
        std::vector<hs_gpu_msm_desc> d(count);
        for (size_t k = 0; k < count; k++)
        {
            d[k].curve = f.curve;
            d[k].n = f.n;
            d[k].scalars = (const unsigned char *)d_scalars + k * f.n * 32;
            d[k].points = f.points.as<hs_gpu_jacobian>();
            d[k].out32 = d_out32 ? (unsigned char *)d_out32 + 32 * k : nullptr;
            d[k].result = d_outjac ? (hs_gpu_jacobian *)d_outjac + k : nullptr;
        }
        int r;
        if ((r = batch_create(&f.pip, d.data(), count, flags)))
            return r;
        f.pip_s = d_scalars;
        f.pip_K = count;
        f.pip_o = d_out32;
        f.pip_j = d_outjac;
        f.pip_flags = flags;
    }
    return batch_enqueue(*f.pip, st);
}

int fixed_enqueue(hs_gpu_fixed &f, const void *d_scalars, size_t count, unsigned flags, int method, void *d_out32,
                  void *d_outjac, hipStream_t st)
{
    if (method == HS_GPU_FIXED_AUTO)
        method = HS_GPU_FIXED_COMB;
    if (method == HS_GPU_FIXED_PIPPENGER)
        return fixed_pippenger_enqueue(f, d_scalars, count, flags, d_out32, d_outjac, st);
    if (method != HS_GPU_FIXED_COMB)
        return fail(HS_GPU_EINVAL, "unknown fixed-base method");
    if (count * hs::FIXED_WINDOWS >= ((size_t)1 << 31))
        return fail(HS_GPU_EINVAL, "too many MSMs in one fixed-base call");
    return f.curve == HS_GPU_HELIOS ? fixed_comb_dispatch<hs::HeliosC>(f, d_scalars, count, flags, d_out32, d_outjac, st)
                                    : fixed_comb_dispatch<hs::SeleneC>(f, d_scalars, count, flags, d_out32, d_outjac, st);
}

// Host-buffer round trip of `count` MSMs over f (synchronous on the library stream).
int fixed_msm_host(hs_gpu_fixed &f, const unsigned char *scalars, size_t count, unsigned flags, int method,
                   unsigned char *out32, hs_gpu_jacobian *results)
{
    if (count == 0)
        return 0;
    if (!scalars)
        return fail(HS_GPU_EINVAL, "null scalars");
    Ctx &cx = ctx();
    hipStream_t st = cx.stream;
    DevBuf ds, dout;
    int r;
    if ((r = ds.reserve(count * f.n * 32)) || (r = dout.reserve(count * (32 + sizeof(hs_gpu_jacobian)))))
    {
        ds.release();
        dout.release();
        return r;
    }
    unsigned char *o32 = dout.as<unsigned char>();
    hs_gpu_jacobian *oj = (hs_gpu_jacobian *)(o32 + count * 32);
    // the scalars go up through a staging slot's pinned buffers (taken for this call only)
    HostStage *hsg = nullptr;
    {
        std::lock_guard<std::mutex> lk(cx.mu);
        if ((r = cx.init()) == 0 && !(hsg = take_stage(cx)))
            r = HS_GPU_EDEVICE;
    }
    if (!r)
        r = hsg->upload(ds.p, scalars, count * f.n * 32, st);
    if (hsg && (flags & HS_GPU_ERASE))
        hsg->erase_pinned();
    if (!r)
        r = fixed_enqueue(f, ds.p, count, flags, method, o32, oj, st);
    if (!r && out32 && hipMemcpyAsync(out32, o32, count * 32, hipMemcpyDeviceToHost, st) != hipSuccess)
        r = fail(HS_GPU_EDEVICE, "result copy failed");
    if (!r && results && hipMemcpyAsync(results, oj, count * sizeof(hs_gpu_jacobian), hipMemcpyDeviceToHost, st) != hipSuccess)
        r = fail(HS_GPU_EDEVICE, "result copy failed");
    if (flags & HS_GPU_ERASE)
    {
        zero_dev(ds.p, count * f.n * 32, st);
        zero_dev(f.scal.p, f.scal.cap, st); // prepared scalars / column patterns
        if (f.pip)
            for (BatchGroup *g : f.pip->groups)
            {
                zero_dev(g->digits.p, g->digits.cap, st);
                zero_dev(g->sorted.p, g->sorted.cap, st);
            }
    }
    if (hipStreamSynchronize(st) != hipSuccess && !r)
        r = fail(HS_GPU_EDEVICE, "fixed-base MSM failed");
    if (hsg)
    {
        std::lock_guard<std::mutex> lk(cx.mu);
        cx.stages_free.push_back(hsg);
    }
    if (f.pip && method == HS_GPU_FIXED_PIPPENGER)
    {
        hs_gpu_batch_destroy(f.pip); // its descriptors point at the buffers freed below
        f.pip = nullptr;
    }
    ds.release();
    dout.release();
    return r;
}

void slot_msm_fixed(int curve, hs_gpu_jacobian *result, const unsigned char *scalars, const hs_gpu_affine *const *tables,
                    size_t n)
{
    if (n == 0)
    {
        if (result)
            set_identity(result);
        return;
    }
    int r = 0;
    hs_gpu_fixed *f = nullptr;
    if (!scalars || !tables)
        r = fail(HS_GPU_EINVAL, "null input with n > 0");
    if (!r)
        r = fixed_alloc(&f, curve, n);
    if (!r)
        r = f->tables.reserve_exact(n * hs::FIXED_ENTRIES * 64);
    if (!r)
    {
        // stage the caller's tables (n pointers to 16 reference affine entries) in one upload
        std::vector<hs_gpu_affine> h(n * hs::FIXED_ENTRIES);
        for (size_t i = 0; i < n && !r; i++)
        {
            if (!tables[i])
                r = fail(HS_GPU_EINVAL, "null table pointer");
            else
                std::memcpy(&h[i * hs::FIXED_ENTRIES], tables[i], hs::FIXED_ENTRIES * sizeof(hs_gpu_affine));
        }
        // the tables are public (multiples of the bases); the scalars are erased in fixed_msm_host
        DevBuf staged;
        hipStream_t st = ctx().stream;
        if (!r && !(r = staged.reserve(h.size() * sizeof(hs_gpu_affine))))
        {
            if (hipMemcpyAsync(staged.p, h.data(), h.size() * sizeof(hs_gpu_affine), hipMemcpyHostToDevice, st) !=
                hipSuccess)
                r = fail(HS_GPU_EDEVICE, "table upload failed");
            const uint32_t entries = (uint32_t)h.size();
            if (!r)
            {
                if (curve == HS_GPU_HELIOS)
                    hipLaunchKernelGGL(hs::k_fixed_tables_from_ref<hs::HeliosC>, dim3((entries + 255) / 256), dim3(256),
                                       0, st, staged.as<uint64_t>(), entries, f->tables.as<uint32_t>());
                else
                    hipLaunchKernelGGL(hs::k_fixed_tables_from_ref<hs::SeleneC>, dim3((entries + 255) / 256), dim3(256),
                                       0, st, staged.as<uint64_t>(), entries, f->tables.as<uint32_t>());
                r = fixed_msm_host(*f, scalars, 1, HS_GPU_EXACT | HS_GPU_ERASE, HS_GPU_FIXED_COMB, nullptr, result);
            }
            staged.release();
        }
    }
    if (f)
        hs_gpu_fixed_destroy(f);
    if (r)
    {
        if (hs_gpu_cpu_msm_fixed_fn fb = (curve == HS_GPU_HELIOS || curve == HS_GPU_SELENE) ? g_fb_fixed[curve].load() : nullptr)
        {
            g_fb_calls++;
            fb(result, scalars, (const struct hs_gpu_affine *const *)tables, n);
            return;
        }
        die("msm_fixed", curve, r);
    }
}

void slot_pedersen(int curve, hs_gpu_jacobian *result, const unsigned char *blinding, const hs_gpu_jacobian *H,
                   const unsigned char *values, const hs_gpu_jacobian *generators, size_t n)
{
    // guard as helios_pedersen.h:63-68, then one MSM of n + 1 terms (:70-84)
    if (n > SIZE_MAX / 32 - 1 || !blinding || !H || (n && (!values || !generators)))
    {
        if (result)
            set_identity(result);
        return;
    }
    std::vector<unsigned char> s(32 * (n + 1));
    std::vector<hs_gpu_jacobian> p(n + 1);
    std::memcpy(s.data(), blinding, 32);
    p[0] = *H;
    if (n)
    {
        std::memcpy(s.data() + 32, values, 32 * n);
        std::memcpy(p.data() + 1, generators, n * sizeof(hs_gpu_jacobian));
    }
    hs_gpu_jacobian r;
    slot_flags(curve, &r, s.data(), p.data(), n + 1, HS_GPU_ERASE);
    secure_erase(s.data(), s.size()); // blinding and values (helios_pedersen.h:85)
    if (result)
        *result = r;
}

} // namespace

// ---------------------------------------------------------------- serving loop
// A scheduler thread owns `lanes` device lanes.  Each lane has a stream, worker streams, and one
// device arena of device_bytes / lanes.  Submitted requests wait in a FIFO.  Whenever a lane is
// idle, the scheduler admits the longest FIFO prefix whose summed footprint fits the lane's arena
// (up to max_batch requests), stages host inputs into the arena, and plans the admitted requests
// as one batch in that arena.  It then launches the batch and copies the outputs back to pinned
// memory.  Requests that arrive while others run are admitted as soon as a lane frees up, and
// lanes run concurrently on the device (continuous batching).
struct SrvReq
{
    hs_gpu_msm_desc d{};
    bool device_inputs = false;
    size_t foot = 0;
    int status = 1; // 1 = pending
    std::chrono::steady_clock::time_point t_submit, t_done;
};

struct SrvLane
{
    hipStream_t st = nullptr;
    std::vector<hipStream_t> workers;
    std::vector<hipEvent_t> ev;
    hipEvent_t done = nullptr;
    Arena arena;
    unsigned char *pin_out = nullptr;
    std::vector<SrvReq *> reqs;
    hs_gpu_batch *plan = nullptr;
    bool busy = false;
};

// Request latencies of a server: totals over every completion since the last reset, and the most
// recent SERVER_LAT_WINDOW values for the percentiles (a ring, so a long-lived server keeps a
// bounded log).
constexpr size_t SERVER_LAT_WINDOW = (size_t)1 << 20;

struct LatencyLog
{
    std::vector<double> ring;
    size_t next = 0;
    uint64_t count = 0;
    double sum = 0, max = 0;
    void add(double us)
    {
        if (ring.size() < SERVER_LAT_WINDOW)
            ring.push_back(us);
        else
        {
            ring[next] = us;
            next = (next + 1) % SERVER_LAT_WINDOW;
        }
        count++;
        sum += us;
        max = std::max(max, us);
    }
    void clear()
    {
        ring.clear();
        next = 0;
        count = 0;
        sum = max = 0;
    }
};

struct hs_gpu_server
{
    size_t lane_bytes = 0;
    int max_batch = 256;
    unsigned flags = 0;
    std::vector<SrvLane> lanes;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<SrvReq *> pending;
    std::unordered_map<uint64_t, SrvReq *> reqs;
    uint64_t next_id = 1;
    bool stop = false;
    std::thread th;
    LatencyLog lat;
    uint64_t points_done = 0, batches = 0;
    std::chrono::steady_clock::time_point t0;
};

namespace
{

// Device bytes one request adds to a batch: staged inputs, outputs and its share of every batch
// buffer (all of which grow linearly with the member MSMs), plus alignment slack.
size_t request_footprint(int curve, size_t n, bool device_inputs)
{
    if (n == 0)
        return 1024;
    // the request's batch may hold it alone (plain parameters) or with others (batch parameters):
    // admit by the larger footprint
    size_t b = 0;
    for (const bool alone : {false, true})
    {
        const MemberPlan m = member_plan(curve, n, alone);
        const size_t W = (256 + m.c - 1) / m.c, B = (size_t)1 << (m.c - 1);
        const size_t G = plan_G((uint32_t)B, m.G), L = std::max<size_t>(1, m.L);
        const size_t E = n * W, nb = W * B, chunks = (E + L - 1) / L + 1;
        b = std::max(b, sizeof(hs::BatchDesc) + n * hs::BASE_RECORD_BYTES + n + E * 2 + E * 4 + 4 + (nb + 1) * 4 +
                            nb * sizeof(hs::xyzz) + hs::seg_out_entries(W, (uint32_t)(B / G)) * sizeof(hs::xyzz) +
                            W * sizeof(hs::xyzz) + sizeof(hs::jac) + chunks * 4 + 4 + 2 * chunks * sizeof(hs::xyzz) + nb * 4 +
                            64);
    }
    if (!device_inputs)
        b += n * (32 + sizeof(hs_gpu_jacobian));
    return b + 32 + sizeof(hs_gpu_jacobian) + 16 * 256;
}

void srv_finish(hs_gpu_server &s, SrvLane &ln, int status)
{
    // called with s.mu held
    const auto now = std::chrono::steady_clock::now();
    const size_t k = ln.reqs.size();
    for (size_t i = 0; i < k; i++)
    {
        SrvReq *q = ln.reqs[i];
        if (status == 0)
        {
            if (q->d.out32)
                std::memcpy(q->d.out32, ln.pin_out + 32 * i, 32);
            if (q->d.result)
                std::memcpy(q->d.result, ln.pin_out + 32 * k + sizeof(hs_gpu_jacobian) * i, sizeof(hs_gpu_jacobian));
            s.lat.add(std::chrono::duration<double, std::micro>(now - q->t_submit).count());
            s.points_done += q->d.n;
        }
        q->t_done = now;
        q->status = status;
    }
    ln.reqs.clear();
    if (ln.plan)
    {
        hs_gpu_batch_destroy(ln.plan);
        ln.plan = nullptr;
    }
    ln.busy = false;
    s.cv_done.notify_all();
}

// Launches ln.reqs as one batch on the lane.  Called without s.mu held.
int srv_launch(hs_gpu_server &s, SrvLane &ln)
{
    const size_t k = ln.reqs.size();
    ln.arena.used = 0;
    std::vector<hs_gpu_msm_desc> d(k);
    int r = 0;
    for (size_t i = 0; i < k && !r; i++)
    {
        const SrvReq &q = *ln.reqs[i];
        d[i] = q.d;
        if (q.d.n && !q.device_inputs)
        {
            void *ds = ln.arena.take(q.d.n * 32), *dp = ln.arena.take(q.d.n * sizeof(hs_gpu_jacobian));
            if (!ds || !dp)
                return fail(HS_GPU_ENOMEM, "request inputs do not fit the lane arena");
            HS_CHECK(hipMemcpyAsync(ds, q.d.scalars, q.d.n * 32, hipMemcpyHostToDevice, ln.st));
            HS_CHECK(hipMemcpyAsync(dp, q.d.points, q.d.n * sizeof(hs_gpu_jacobian), hipMemcpyHostToDevice, ln.st));
            d[i].scalars = (const unsigned char *)ds;
            d[i].points = (const hs_gpu_jacobian *)dp;
        }
    }
    unsigned char *dout = (unsigned char *)ln.arena.take(k * (32 + sizeof(hs_gpu_jacobian)));
    if (!dout)
        return fail(HS_GPU_ENOMEM, "request outputs do not fit the lane arena");
    for (size_t i = 0; i < k; i++)
    {
        d[i].out32 = dout + 32 * i;
        d[i].result = (hs_gpu_jacobian *)(dout + 32 * k + sizeof(hs_gpu_jacobian) * i);
    }
    if ((r = batch_create(&ln.plan, d.data(), k, s.flags, &ln.arena, ln.st, &ln.workers, &ln.ev)))
        return r;
    if ((r = batch_enqueue(*ln.plan, ln.st)))
        return r;
    HS_CHECK(hipMemcpyAsync(ln.pin_out, dout, k * (32 + sizeof(hs_gpu_jacobian)), hipMemcpyDeviceToHost, ln.st));
    HS_CHECK(hipEventRecord(ln.done, ln.st));
    return 0;
}

void srv_loop(hs_gpu_server *sp)
{
    hs_gpu_server &s = *sp;
    hipSetDevice(ctx().device);
    std::unique_lock<std::mutex> lk(s.mu);
    while (true)
    {
        bool any_busy = false;
        // completions
        for (SrvLane &ln : s.lanes)
        {
            if (!ln.busy)
                continue;
            const hipError_t e = hipEventQuery(ln.done);
            if (e == hipSuccess)
                srv_finish(s, ln, 0);
            else if (e != hipErrorNotReady)
                srv_finish(s, ln, fail(HS_GPU_EDEVICE, std::string("serving batch failed: ") + hipGetErrorString(e)));
            else
                any_busy = true;
        }
        if (s.stop && s.pending.empty() && !any_busy)
            break;
        // admissions
        for (SrvLane &ln : s.lanes)
        {
            if (ln.busy || s.pending.empty())
                continue;
            size_t used = 0;
            while (!s.pending.empty() && (int)ln.reqs.size() < s.max_batch)
            {
                SrvReq *q = s.pending.front();
                if (q->foot > s.lane_bytes)
                {
                    // can never fit: reject rather than block the queue
                    s.pending.pop_front();
                    q->status = fail(HS_GPU_ENOMEM, "request footprint exceeds the per-lane device budget");
                    q->t_done = std::chrono::steady_clock::now();
                    s.cv_done.notify_all();
                    continue;
                }
                if (used + q->foot > s.lane_bytes)
                    break;
                used += q->foot;
                ln.reqs.push_back(q);
                s.pending.pop_front();
            }
            if (ln.reqs.empty())
                continue;
            ln.busy = true;
            s.batches++;
            lk.unlock();
            const int r = srv_launch(s, ln);
            lk.lock();
            if (r)
            {
                hipStreamSynchronize(ln.st);
                srv_finish(s, ln, r);
            }
            else
                any_busy = true;
        }
        if (any_busy)
            s.cv_work.wait_for(lk, std::chrono::microseconds(20));
        else if (s.pending.empty() && !s.stop)
            s.cv_work.wait(lk);
    }
}


// ---------------------------------------------------------------- autotuner
int tune_load_file(const char *path)
{
    FILE *f = std::fopen(path, "r");
    if (!f)
        return fail(HS_GPU_EINVAL, std::string("cannot open tuning file ") + path);
    int mode, curve, lg, c, L, G, k = 0;
    float ms;
    TuneTable &t = tune();
    std::lock_guard<std::mutex> lk(t.mu);
    while (std::fscanf(f, "%d %d %d %d %d %d %f", &mode, &curve, &lg, &c, &L, &G, &ms) == 7)
    {
        if (mode < 0 || mode > 1 || curve < 0 || curve > 1 || lg < 0 || lg > 40 || c < 4 || c > 16 || L < 1 || G < 1 ||
            (G & (G - 1)))
            continue;
        TuneEntry e;
        e.c = c;
        e.L = L;
        e.G = G;
        e.ms = ms;
        t.e[mode * 1024 + curve * 64 + lg] = e;
        k++;
    }
    std::fclose(f);
    return k;
}

// deterministic synthetic inputs on the device: SSWU points of a counter-based byte stream and
// 255-bit scalars
struct TuneInputs
{
    DevBuf u, pts, sc, out;
    int make(int curve, size_t n, hipStream_t st)
    {
        int r;
        if ((r = u.reserve(n * 32)) || (r = pts.reserve(n * sizeof(hs_gpu_jacobian))) || (r = sc.reserve(n * 32)) ||
            (r = out.reserve(4096 * 32)))
            return r;
        std::vector<uint32_t> h(n * 8);
        uint64_t x = 0x9e3779b97f4a7c15ull ^ (uint64_t)n;
        for (auto &w : h)
        {
            x ^= x << 13;
            x ^= x >> 7;
            x ^= x << 17;
            w = (uint32_t)(x >> 11);
        }
        HS_CHECK(hipMemcpy(u.p, h.data(), n * 32, hipMemcpyHostToDevice));
        for (size_t i = 0; i < n; i++)
            h[8 * i + 7] &= 0x7fffffffu;
        HS_CHECK(hipMemcpy(sc.p, h.data(), n * 32, hipMemcpyHostToDevice));
        Ctx &cx = ctx();
        if ((r = curve == HS_GPU_HELIOS ? map_to_curve_device<hs::HeliosC>(cx, u.as<uint32_t>(), n, pts.as<uint64_t>(), nullptr, st)
                                        : map_to_curve_device<hs::SeleneC>(cx, u.as<uint32_t>(), n, pts.as<uint64_t>(), nullptr, st)))
            return r;
        HS_CHECK(hipStreamSynchronize(st));
        return 0;
    }
    ~TuneInputs()
    {
        for (DevBuf *d : {&u, &pts, &sc, &out})
            d->release();
    }
};

// Time (ms, best of `reps`) of one configuration: a plain MSM of n points (mode 0), or a batch of
// k MSMs of n points each (mode 1).
int tune_time(int mode, int curve, size_t n, size_t k, size_t ninp, TuneInputs &in, const TuneEntry &cand, int reps,
              float &best)
{
    Ctx &cx = ctx();
    hipStream_t st = cx.stream;
    g_force = &cand;
    int r = 0;
    hs_gpu_batch *b = nullptr;
    if (mode == 1)
    {
        std::vector<hs_gpu_msm_desc> d(k);
        for (size_t i = 0; i < k; i++)
        {
            d[i].curve = curve;
            d[i].n = n;
            const size_t off = (i * n) % (ninp - n + 1); // distinct slices of the inputs
            d[i].scalars = in.sc.as<unsigned char>() + 32 * off;
            d[i].points = in.pts.as<hs_gpu_jacobian>() + off;
            d[i].out32 = in.out.as<unsigned char>() + 32 * (i % 4096);
            d[i].result = nullptr;
        }
        r = batch_create(&b, d.data(), k, 0);
    }
    g_force = nullptr;
    if (r)
        return r;
    hipEvent_t e0, e1;
    HS_CHECK(hipEventCreate(&e0));
    HS_CHECK(hipEventCreate(&e1));
    best = 1e30f;
    for (int i = 0; i <= reps && !r; i++)
    {
        HS_CHECK(hipEventRecord(e0, st));
        if (mode == 0)
        {
            g_force = &cand;
            std::lock_guard<std::mutex> lk(cx.mu);
            r = msm_device_any(curve, cx, in.sc.p, in.pts.p, n, 0, in.out.p, nullptr, st);
            g_force = nullptr;
        }
        else
            r = batch_enqueue(*b, st);
        HS_CHECK(hipEventRecord(e1, st));
        HS_CHECK(hipEventSynchronize(e1));
        float ms;
        HS_CHECK(hipEventElapsedTime(&ms, e0, e1));
        if (i > 0 && ms < best) // the first run warms the workspace
            best = ms;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (b)
        hs_gpu_batch_destroy(b);
    return r;
}

} // namespace

extern "C"
{

int hs_gpu_server_create(const hs_gpu_server_config *cfg, hs_gpu_server **out)
{
    if (!out)
        return fail(HS_GPU_EINVAL, "null argument");
    Ctx &cx = ctx();
    int r;
    {
        std::lock_guard<std::mutex> lk(cx.mu);
        if ((r = cx.init()))
            return r;
    }
    const int lanes = cfg && cfg->lanes > 0 ? cfg->lanes : 4;
    const size_t bytes = cfg && cfg->device_bytes ? cfg->device_bytes : ((size_t)8 << 30);
    auto *s = new hs_gpu_server();
    s->max_batch = cfg && cfg->max_batch > 0 ? cfg->max_batch : 256;
    s->flags = cfg ? cfg->flags : 0;
    s->lane_bytes = bytes / (size_t)lanes;
    s->lanes.resize(lanes);
    s->t0 = std::chrono::steady_clock::now();
    for (SrvLane &ln : s->lanes)
    {
        bool ok = hipStreamCreateWithFlags(&ln.st, hipStreamNonBlocking) == hipSuccess &&
                  hipEventCreateWithFlags(&ln.done, hipEventDisableTiming) == hipSuccess &&
                  dev_malloc((void **)&ln.arena.base, s->lane_bytes) == 0 &&
                  hipHostMalloc((void **)&ln.pin_out, (size_t)s->max_batch * (32 + sizeof(hs_gpu_jacobian)), 0) ==
                      hipSuccess;
        ln.arena.cap = ok ? s->lane_bytes : 0;
        ln.workers.resize(4);
        ln.ev.resize(5);
        for (auto &w : ln.workers)
            ok = ok && hipStreamCreateWithFlags(&w, hipStreamNonBlocking) == hipSuccess;
        for (auto &e : ln.ev)
            ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
        if (!ok)
        {
            hs_gpu_server_destroy(s);
            return fail(HS_GPU_ENOMEM, "serving lane allocation failed (device_bytes too large?)");
        }
    }
    s->th = std::thread(srv_loop, s);
    *out = s;
    return 0;
}

int hs_gpu_server_submit(hs_gpu_server *s, const hs_gpu_msm_desc *req, int device_inputs, uint64_t *ticket)
{
    if (!s || !req || !ticket)
        return fail(HS_GPU_EINVAL, "null argument");
    if ((req->curve != HS_GPU_HELIOS && req->curve != HS_GPU_SELENE) || (req->n && (!req->scalars || !req->points)))
        return fail(HS_GPU_EINVAL, "bad request");
    if (req->n >= ((size_t)1 << 24))
        return fail(HS_GPU_EINVAL, "request n >= 2^24 (the plain MSM takes larger sizes)");
    auto *q = new SrvReq();
    q->d = *req;
    q->device_inputs = device_inputs != 0;
    q->foot = request_footprint(req->curve, req->n, q->device_inputs);
    q->t_submit = std::chrono::steady_clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop)
        {
            delete q;
            return fail(HS_GPU_EINVAL, "server is stopping");
        }
        *ticket = s->next_id++;
        s->reqs[*ticket] = q;
        s->pending.push_back(q);
    }
    s->cv_work.notify_one();
    return 0;
}

int hs_gpu_server_wait(hs_gpu_server *s, uint64_t ticket, int timeout_ms)
{
    if (!s)
        return fail(HS_GPU_EINVAL, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end())
        return fail(HS_GPU_EINVAL, "unknown ticket");
    SrvReq *q = it->second;
    auto ready = [q] { return q->status != 1; };
    if (timeout_ms < 0)
        s->cv_done.wait(lk, ready);
    else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready))
        return 1;
    const int st = q->status;
    s->reqs.erase(it);
    delete q;
    return st;
}

int hs_gpu_server_get_stats(hs_gpu_server *s, hs_gpu_server_stats *out)
{
    if (!s || !out)
        return fail(HS_GPU_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    std::vector<double> v = s->lat.ring;
    std::sort(v.begin(), v.end());
    auto pct = [&v](double p) { return v.empty() ? 0.0 : v[std::min(v.size() - 1, (size_t)(p * (v.size() - 1) + 0.5))]; };
    out->completed = s->lat.count;
    out->batches = s->batches;
    out->points = s->points_done;
    out->p50_us = pct(0.50);
    out->p99_us = pct(0.99);
    out->max_us = s->lat.max;
    out->mean_us = s->lat.count ? s->lat.sum / (double)s->lat.count : 0.0;
    out->elapsed_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - s->t0).count();
    return 0;
}

int hs_gpu_server_reset_stats(hs_gpu_server *s)
{
    if (!s)
        return fail(HS_GPU_EINVAL, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    s->lat.clear();
    s->points_done = 0;
    s->batches = 0;
    s->t0 = std::chrono::steady_clock::now();
    return 0;
}

void hs_gpu_server_destroy(hs_gpu_server *s)
{
    if (!s)
        return;
    if (s->th.joinable())
    {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
        }
        s->cv_work.notify_all();
        s->th.join();
    }
    for (SrvLane &ln : s->lanes)
    {
        if (ln.plan)
            hs_gpu_batch_destroy(ln.plan);
        if (ln.st)
            hipStreamSynchronize(ln.st);
        for (auto w : ln.workers)
            if (w)
                hipStreamDestroy(w);
        for (auto e : ln.ev)
            if (e)
                hipEventDestroy(e);
        if (ln.done)
            hipEventDestroy(ln.done);
        if (ln.st)
            hipStreamDestroy(ln.st);
        if (ln.arena.base)
            dev_free(ln.arena.base, s->lane_bytes);
        if (ln.pin_out)
            hipHostFree(ln.pin_out);
    }
    for (auto &kv : s->reqs)
        delete kv.second;
    delete s;
}

} // extern "C"

namespace
{
} // namespace

extern "C"
{

void hs_gpu_helios_msm_vartime(hs_gpu_jacobian *result, const unsigned char *scalars, const hs_gpu_jacobian *points,
                               size_t n)
{
    slot(HS_GPU_HELIOS, result, scalars, points, n);
}

void hs_gpu_selene_msm_vartime(hs_gpu_jacobian *result, const unsigned char *scalars, const hs_gpu_jacobian *points,
                               size_t n)
{
    slot(HS_GPU_SELENE, result, scalars, points, n);
}

int hs_gpu_msm_ex(int curve, hs_gpu_jacobian *result, unsigned char out32[32], const unsigned char *scalars,
                  const hs_gpu_jacobian *points, size_t n, unsigned flags)
{
    return msm_host(curve, result, out32, scalars, points, n, flags);
}

int hs_gpu_msm_device(int curve, const void *d_scalars, const void *d_points, size_t n, unsigned flags, void *d_out32,
                      void *d_out_jac, void *stream)
{
    if (n == 0 || !d_scalars || !d_points || !d_out32)
        return fail(HS_GPU_EINVAL, "hs_gpu_msm_device needs n > 0 and device buffers");
    Ctx &cx = ctx();
    std::lock_guard<std::mutex> lk(cx.mu);
    int r;
    if ((r = cx.init()))
        return r;
    hipStream_t st = stream ? (hipStream_t)stream : cx.stream;
    return msm_device_any(curve, cx, d_scalars, d_points, n, flags, d_out32, d_out_jac, st);
}

// Plan of the no-allocation entry point: the built-in parameters (window bits choose_c(n), run
// length 32, segments of 8 buckets, one window group), independent of the tuning table and the
// environment, so the scratch it needs is a function of n alone.
static Plan noalloc_plan(size_t n)
{
    Plan p;
    p.c = choose_c(n);
    p.W = (256 + p.c - 1) / p.c;
    p.B = 1u << (p.c - 1);
    p.G = p.B < 8u ? p.B : 8u;
    p.nb = (uint32_t)p.W * p.B;
    p.L = 32;
    p.Q = 1;
    return p;
}

int hs_gpu_msm_device_noalloc(int curve, const void *d_scalars, const void *d_points, size_t n, unsigned flags,
                              void *d_out32, void *d_out_jac, void *stream)
{
    if (curve != HS_GPU_HELIOS && curve != HS_GPU_SELENE)
        return fail(HS_GPU_EINVAL, "unknown curve id");
    if (n == 0 || !d_scalars || !d_points || !d_out32)
        return fail(HS_GPU_EINVAL, "hs_gpu_msm_device_noalloc needs n > 0 and device buffers");
    if (n > NOALLOC_MAX_POINTS)
        return fail(HS_GPU_EINVAL, "hs_gpu_msm_device_noalloc takes at most " + std::to_string(NOALLOC_MAX_POINTS) +
                                       " points (its scratch holds " + std::to_string(NOALLOC_SCRATCH_BYTES) +
                                       " bytes); n = " + std::to_string(n));
    Ctx &cx = ctx();
    std::lock_guard<std::mutex> lk(cx.mu);
    int r;
    if ((r = cx.init())) // first call only; hs_gpu_init beforehand keeps every call allocation-free
        return r;
    const Plan P = noalloc_plan(n);
    const MsmGeom geo = msm_geom(P, n);
    // carve every buffer of the plan from the scratch array (256-byte aligned)
    size_t off = 0;
    for (const auto &b : msm_buffers(cx.noalloc_ws, P, geo, n))
    {
        off = (off + 255) & ~(size_t)255;
        if (off + b.second > NOALLOC_SCRATCH_BYTES)
            return fail(HS_GPU_EINVAL, "hs_gpu_msm_device_noalloc: the plan of n = " + std::to_string(n) +
                                           " does not fit the scratch array");
        b.first->p = cx.scratch_ptr + off;
        b.first->cap = b.second;
        off += b.second;
    }
    hipStream_t st = stream ? (hipStream_t)stream : cx.stream;
    return curve == HS_GPU_HELIOS
               ? msm_device<hs::HeliosC>(cx, (const uint32_t *)d_scalars, (const uint64_t *)d_points, n, flags,
                                         (uint32_t *)d_out32, (uint64_t *)d_out_jac, st, 0, &cx.noalloc_ws, &P)
               : msm_device<hs::SeleneC>(cx, (const uint32_t *)d_scalars, (const uint64_t *)d_points, n, flags,
                                         (uint32_t *)d_out32, (uint64_t *)d_out_jac, st, 0, &cx.noalloc_ws, &P);
}

int hs_gpu_synchronize(void)
{
    Ctx &cx = ctx();
    if (!cx.ready)
        return 0;
    HS_CHECK(hipStreamSynchronize(cx.stream));
    return 0;
}

int hs_gpu_init(void)
{
    Ctx &cx = ctx();
    std::lock_guard<std::mutex> lk(cx.mu);
    return cx.init();
}

int hs_gpu_set_cpu_fallback(int curve, hs_gpu_cpu_msm_fn msm, hs_gpu_cpu_msm_fixed_fn msm_fixed)
{
    if (curve != HS_GPU_HELIOS && curve != HS_GPU_SELENE)
        return fail(HS_GPU_EINVAL, "unknown curve id");
    g_fb_msm[curve].store(msm);
    g_fb_fixed[curve].store(msm_fixed);
    return 0;
}

uint64_t hs_gpu_fallback_count(void) { return g_fb_calls.load(); }

int hs_gpu_set_device_limit(size_t bytes)
{
    g_dev_limit.store(bytes);
    return 0;
}

size_t hs_gpu_device_bytes(void) { return g_dev_held.load(); }

int hs_gpu_trim(void)
{
    Ctx &cx = ctx();
    std::lock_guard<std::mutex> lk(cx.mu);
    if (!cx.ready)
        return 0;
    // every stream the cached workspaces serve has finished before they are freed
    HS_CHECK(hipDeviceSynchronize());
    for (auto &kv : cx.wss)
        kv.second->release_all();
    for (HostStage *h : cx.stages_all)
        for (DevBuf *d : {&h->scalars, &h->points, &h->out})
            d->release();
    for (DevBuf *d : {&cx.scalars, &cx.points, &cx.out32, &cx.outjac, &cx.status, &cx.nbad, &cx.probe})
        d->release();
    return 0;
}

const char *hs_gpu_last_error(void) { return g_err.c_str(); }

int hs_gpu_batch_create(const hs_gpu_msm_desc *descs, size_t count, unsigned flags, hs_gpu_batch **out)
{
    if (!out || (count && !descs))
        return fail(HS_GPU_EINVAL, "null argument");
    return batch_create(out, descs, count, flags);
}

int hs_gpu_batch_launch(hs_gpu_batch *b, void *stream)
{
    if (!b)
        return fail(HS_GPU_EINVAL, "null batch");
    hipStream_t st = stream ? (hipStream_t)stream : ctx().stream;
    if (b->exec)
    {
        HS_CHECK(hipGraphLaunch(b->exec, st));
        return 0;
    }
    return batch_enqueue(*b, st);
}

int hs_gpu_batch_capture(hs_gpu_batch *b)
{
    if (!b)
        return fail(HS_GPU_EINVAL, "null batch");
    if (b->exec)
        return 0;
    if (!b->cap_stream)
        HS_CHECK(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
    HS_CHECK(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
    int r = batch_enqueue(*b, b->cap_stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
    if (r)
        return r;
    if (e != hipSuccess)
        return fail(HS_GPU_EDEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    b->graph = g;
    HS_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return 0;
}

size_t hs_gpu_batch_points(const hs_gpu_batch *b) { return b ? b->total_points : 0; }

void hs_gpu_batch_destroy(hs_gpu_batch *b)
{
    if (!b)
        return;
    if (b->exec)
        hipGraphExecDestroy(b->exec);
    if (b->graph)
        hipGraphDestroy(b->graph);
    for (BatchGroup *g : b->groups)
    {
        if (!b->in_arena)
            for (DevBuf *d : {&g->desc, &g->aff, &g->isid, &g->digits, &g->sorted, &g->offsets, &g->buckets, &g->seg,
                              &g->win, &g->total, &g->chunk_bucket, &g->bnd, &g->long_list, &g->long_count})
                d->release();
        delete g;
    }
    if (!b->borrowed)
    {
        for (hipStream_t s : b->workers)
            hipStreamDestroy(s);
        for (hipEvent_t e : b->ev)
            hipEventDestroy(e);
    }
    if (b->cap_stream)
        hipStreamDestroy(b->cap_stream);
    if (!b->in_arena)
        b->empty.release();
    delete b;
}

int hs_gpu_msm_batch(const hs_gpu_msm_desc *descs, size_t count, unsigned flags)
{
    if (count && !descs)
        return fail(HS_GPU_EINVAL, "null descriptors");
    // validate everything before any allocation
    size_t n_tot = 0;
    for (size_t i = 0; i < count; i++)
    {
        const hs_gpu_msm_desc &d = descs[i];
        if (d.curve != HS_GPU_HELIOS && d.curve != HS_GPU_SELENE)
            return fail(HS_GPU_EINVAL, "bad curve in batch descriptor " + std::to_string(i));
        if (d.n && (!d.scalars || !d.points))
            return fail(HS_GPU_EINVAL, "null input in batch descriptor " + std::to_string(i));
        if (d.n >= (1u << 24))
            return fail(HS_GPU_EINVAL, "batch element " + std::to_string(i) + ": n >= 2^24 (the plain MSM takes larger sizes)");
        n_tot += d.n;
    }
    Ctx &cx = ctx();
    int r;
    // the inputs go up through a staging slot's pinned buffers (taken for this call only)
    HostStage *hsg = nullptr;
    {
        std::lock_guard<std::mutex> lk(cx.mu);
        if ((r = cx.init()))
            return r;
        if (!(hsg = take_stage(cx)))
            return HS_GPU_EDEVICE;
    }
    // stage every input in one device buffer, run the batch, copy the outputs back; every exit
    // below goes through the cleanup at the end
    DevBuf in, outb;
    hs_gpu_batch *b = nullptr;
    hipStream_t st = cx.stream;
    const size_t in_bytes = n_tot * (32 + sizeof(hs_gpu_jacobian));
    std::vector<hs_gpu_msm_desc> dd(descs, descs + count);
    if (!(r = in.reserve(in_bytes + 64)) && !(r = outb.reserve(count * (32 + sizeof(hs_gpu_jacobian)) + 64)))
    {
        unsigned char *sc_base = in.as<unsigned char>(), *pt_base = sc_base + n_tot * 32;
        unsigned char *o32_base = outb.as<unsigned char>(), *oj_base = o32_base + count * 32;
        size_t off = 0;
        for (size_t i = 0; i < count && !r; i++)
        {
            const size_t n = descs[i].n;
            if (n && !(r = hsg->upload(sc_base + off * 32, descs[i].scalars, n * 32, st)))
                r = hsg->upload(pt_base + off * sizeof(hs_gpu_jacobian), descs[i].points, n * sizeof(hs_gpu_jacobian), st);
            dd[i].scalars = sc_base + off * 32;
            dd[i].points = (const hs_gpu_jacobian *)(pt_base + off * sizeof(hs_gpu_jacobian));
            dd[i].out32 = o32_base + i * 32;
            dd[i].result = (hs_gpu_jacobian *)(oj_base + i * sizeof(hs_gpu_jacobian));
            off += n;
        }
        if (!r && hipStreamSynchronize(st) != hipSuccess)
            r = fail(HS_GPU_EDEVICE, "batch input upload failed");
        if (flags & HS_GPU_ERASE)
            hsg->erase_pinned(); // also after a failed upload
        if (!r)
            r = batch_create(&b, dd.data(), count, flags);
        if (!r)
            r = batch_enqueue(*b, st);
        if (!r)
        {
            std::vector<unsigned char> o32(count * 32);
            std::vector<hs_gpu_jacobian> oj(count);
            if (flags & HS_GPU_ERASE)
            {
                zero_dev(sc_base, n_tot * 32, st);
                for (BatchGroup *g : b->groups)
                {
                    zero_dev(g->digits.p, g->digits.cap, st);
                    zero_dev(g->sorted.p, g->sorted.cap, st);
                }
            }
            hipError_t e1 = hipMemcpyAsync(o32.data(), o32_base, count * 32, hipMemcpyDeviceToHost, st);
            hipError_t e2 = hipMemcpyAsync(oj.data(), oj_base, count * sizeof(hs_gpu_jacobian), hipMemcpyDeviceToHost, st);
            hipError_t e3 = hipStreamSynchronize(st);
            if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess)
                r = fail(HS_GPU_EDEVICE, "batch result copy failed");
            for (size_t i = 0; i < count && !r; i++)
            {
                if (descs[i].n == 0)
                {
                    if (descs[i].out32)
                        std::memset(descs[i].out32, 0, 32);
                    if (descs[i].result)
                        set_identity(descs[i].result);
                    continue;
                }
                if (descs[i].out32)
                    std::memcpy(descs[i].out32, &o32[i * 32], 32);
                if (descs[i].result)
                    descs[i].result[0] = oj[i];
            }
        }
    }
    hipStreamSynchronize(st);
    {
        std::lock_guard<std::mutex> lk(cx.mu);
        cx.stages_free.push_back(hsg);
    }
    hs_gpu_batch_destroy(b);
    in.release();
    outb.release();
    return r;
}

int hs_gpu_profile(int on)
{
    Ctx &cx = ctx();
    std::lock_guard<std::mutex> lk(cx.mu);
    int r;
    if ((r = cx.init()))
        return r;
    cx.profile = on != 0;
    cx.evused = 0;
    cx.cur = nullptr;
    return 0;
}

int hs_gpu_phase_summary(int curve, float *ms, int cap, int *calls)
{
    Ctx &cx = ctx();
    std::lock_guard<std::mutex> lk(cx.mu);
    int k = 0, nc = 0;
    for (int i = 0; i < cap; i++)
        ms[i] = 0.f;
    for (size_t s = 0; s < cx.evused; s++)
    {
        Ctx::EvSet &e = cx.evsets[s];
        if ((curve >= 0 && e.curve != curve) || e.n < 2)
            continue;
        HS_CHECK(hipEventSynchronize(e.ev[e.n - 1]));
        nc++;
        for (int i = 1; i < e.n; i++)
        {
            const int ph = e.phase[i];
            if (ph < 0 || ph >= cap)
                continue;
            float t;
            HS_CHECK(hipEventElapsedTime(&t, e.ev[i - 1], e.ev[i]));
            ms[ph] += t;
            if (ph + 1 > k)
                k = ph + 1;
        }
    }
    if (calls)
        *calls = nc;
    return k;
}

int hs_gpu_phase_times(float *ms, int cap)
{
    Ctx &cx = ctx();
    std::lock_guard<std::mutex> lk(cx.mu);
    for (int i = 0; i < cap; i++)
        ms[i] = 0.f;
    if (cx.evused == 0)
        return 0;
    Ctx::EvSet &e = cx.evsets[cx.evused - 1];
    if (e.n < 2)
        return 0;
    HS_CHECK(hipEventSynchronize(e.ev[e.n - 1]));
    int k = 0;
    for (int i = 1; i < e.n; i++)
    {
        const int ph = e.phase[i];
        if (ph < 0 || ph >= cap)
            continue;
        float t;
        HS_CHECK(hipEventElapsedTime(&t, e.ev[i - 1], e.ev[i]));
        ms[ph] += t;
        if (ph + 1 > k)
            k = ph + 1;
    }
    return k;
}

const char *hs_gpu_phase_name(int i)
{
    return (i >= 0 && i < (int)(sizeof PHASE_NAMES / sizeof PHASE_NAMES[0])) ? PHASE_NAMES[i] : "";
}

int hs_gpu_debug_dump(int curve, int which, void *out, size_t cap_bytes, size_t *count)
{
    Ctx &cx = ctx();
    std::lock_guard<std::mutex> lk(cx.mu);
    if (!cx.ready)
        return fail(HS_GPU_EINVAL, "no MSM has run");
    *count = 0;
    if (!cx.last_ws)
        return fail(HS_GPU_EINVAL, "no MSM has run");
    Workspace &ws = *cx.last_ws;
    if (which == 3 || which == 4)
    {
        // 3: sorted entries, 4: bucket offsets (raw uint32)
        DevBuf &b = which == 3 ? ws.sorted : ws.offsets;
        size_t nb = std::min(cap_bytes, b.cap);
        HS_CHECK(hipMemcpy(out, b.p, nb, hipMemcpyDeviceToHost));
        *count = nb / 4;
        return 0;
    }
    DevBuf &src = which == 0 ? ws.win : which == 1 ? ws.seg : which == 5 ? ws.bnd : ws.buckets;
    // window sums are Jacobian, segment and bucket sums xyzz
    size_t n = std::min(cap_bytes / 32, src.cap / (which == 0 ? sizeof(hs::jac) : sizeof(hs::xyzz)));
    if (n == 0)
        return 0;
    HS_CHECK(hipDeviceSynchronize()); // the last MSM may have run on any stream
    int r;
    if ((r = cx.probe.reserve(n * 32 + 64)))
        return r;
    const uint32_t nn = (uint32_t)n;
    const dim3 grid((nn + 63) / 64);
    if (which == 0)
    {
        if (curve == HS_GPU_HELIOS)
            hipLaunchKernelGGL(hs::k_encode_jac<hs::HeliosC>, grid, dim3(64), 0, cx.stream, src.as<hs::jac>(), nn,
                               cx.probe.as<uint32_t>());
        else
            hipLaunchKernelGGL(hs::k_encode_jac<hs::SeleneC>, grid, dim3(64), 0, cx.stream, src.as<hs::jac>(), nn,
                               cx.probe.as<uint32_t>());
    }
    else if (curve == HS_GPU_HELIOS)
        hipLaunchKernelGGL(hs::k_encode_xyzz<hs::HeliosC>, grid, dim3(64), 0, cx.stream, src.as<hs::xyzz>(), nn,
                           cx.probe.as<uint32_t>());
    else
        hipLaunchKernelGGL(hs::k_encode_xyzz<hs::SeleneC>, grid, dim3(64), 0, cx.stream, src.as<hs::xyzz>(), nn,
                           cx.probe.as<uint32_t>());
    HS_CHECK(hipGetLastError());
    HS_CHECK(hipStreamSynchronize(cx.stream));
    HS_CHECK(hipMemcpy(out, cx.probe.p, n * 32, hipMemcpyDeviceToHost));
    *count = n;
    return 0;
}

int hs_gpu_probe_int_mul_peak(double *products_per_second)
{
    Ctx &cx = ctx();
    std::lock_guard<std::mutex> lk(cx.mu);
    int r;
    if ((r = cx.init()))
        return r;
    hipDeviceProp_t prop;
    HS_CHECK(hipGetDeviceProperties(&prop, cx.device));
    const int blocks = prop.multiProcessorCount * 16, threads = 256, iters = 4096;
    if ((r = cx.probe.reserve((size_t)blocks * threads * 4)))
        return r;
    hipEvent_t e0, e1;
    HS_CHECK(hipEventCreate(&e0));
    HS_CHECK(hipEventCreate(&e1));
    hipLaunchKernelGGL(hs::k_probe_mad64, dim3(blocks), dim3(threads), 0, cx.stream, cx.probe.as<uint32_t>(), 64, 7u);
    float best = 1e30f;
    for (int rep = 0; rep < 3; rep++)
    {
        HS_CHECK(hipEventRecord(e0, cx.stream));
        hipLaunchKernelGGL(hs::k_probe_mad64, dim3(blocks), dim3(threads), 0, cx.stream, cx.probe.as<uint32_t>(),
                           iters, 7u + rep);
        HS_CHECK(hipEventRecord(e1, cx.stream));
        HS_CHECK(hipEventSynchronize(e1));
        float ms;
        HS_CHECK(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best)
            best = ms;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    *products_per_second = (double)blocks * threads * iters * 8 / (best * 1e-3);
    return 0;
}

int hs_gpu_device_info(int *compute_units, int *max_clock_khz, char *name, size_t name_cap)
{
    Ctx &cx = ctx();
    std::lock_guard<std::mutex> lk(cx.mu);
    int r;
    if ((r = cx.init()))
        return r;
    hipDeviceProp_t prop;
    HS_CHECK(hipGetDeviceProperties(&prop, cx.device));
    if (compute_units)
        *compute_units = prop.multiProcessorCount;
    if (max_clock_khz)
        *max_clock_khz = prop.clockRate;
    if (name && name_cap)
    {
        std::snprintf(name, name_cap, "%s (%s)", prop.name, prop.gcnArchName);
    }
    return 0;
}

int hs_gpu_map_to_curve(int curve, const unsigned char *u, size_t n, hs_gpu_jacobian *out_points, unsigned char *out32)
{
    if (curve != HS_GPU_HELIOS && curve != HS_GPU_SELENE)
        return fail(HS_GPU_EINVAL, "unknown curve id");
    if (n == 0)
        return 0;
    if (!u || (!out_points && !out32))
        return fail(HS_GPU_EINVAL, "null argument");
    Ctx &cx = ctx();
    std::lock_guard<std::mutex> lk(cx.mu);
    int r;
    if ((r = cx.init()) || (r = cx.scalars.reserve(n * 32)) || (r = cx.points.reserve(n * sizeof(hs_gpu_jacobian))) ||
        (r = cx.status.reserve(n * 32)))
        return r;
    hipStream_t st = cx.stream;
    HS_CHECK(hipMemcpyAsync(cx.scalars.p, u, n * 32, hipMemcpyHostToDevice, st));
    uint64_t *dj = out_points ? cx.points.as<uint64_t>() : nullptr;
    uint32_t *d32 = out32 ? cx.status.as<uint32_t>() : nullptr;
    r = curve == HS_GPU_HELIOS ? map_to_curve_device<hs::HeliosC>(cx, cx.scalars.as<uint32_t>(), n, dj, d32, st)
                               : map_to_curve_device<hs::SeleneC>(cx, cx.scalars.as<uint32_t>(), n, dj, d32, st);
    if (r)
        return r;
    if (out_points)
        HS_CHECK(hipMemcpyAsync(out_points, dj, n * sizeof(hs_gpu_jacobian), hipMemcpyDeviceToHost, st));
    if (out32)
        HS_CHECK(hipMemcpyAsync(out32, d32, n * 32, hipMemcpyDeviceToHost, st));
    HS_CHECK(hipStreamSynchronize(st));
    return 0;
}

int hs_gpu_map_to_curve_device(int curve, const void *d_u, size_t n, void *d_points, void *d_out32, void *stream)
{
    if (curve != HS_GPU_HELIOS && curve != HS_GPU_SELENE)
        return fail(HS_GPU_EINVAL, "unknown curve id");
    if (n == 0)
        return 0;
    Ctx &cx = ctx();
    std::lock_guard<std::mutex> lk(cx.mu);
    int r;
    if ((r = cx.init()))
        return r;
    hipStream_t st = stream ? (hipStream_t)stream : cx.stream;
    return curve == HS_GPU_HELIOS
               ? map_to_curve_device<hs::HeliosC>(cx, (const uint32_t *)d_u, n, (uint64_t *)d_points, (uint32_t *)d_out32, st)
               : map_to_curve_device<hs::SeleneC>(cx, (const uint32_t *)d_u, n, (uint64_t *)d_points, (uint32_t *)d_out32, st);
}

int hs_gpu_msm_bytes(int curve, unsigned char out32[32], const unsigned char *scalars, const unsigned char *points32,
                     size_t n, unsigned flags)
{
    if (curve != HS_GPU_HELIOS && curve != HS_GPU_SELENE)
        return fail(HS_GPU_EINVAL, "unknown curve id");
    if (!out32)
        return fail(HS_GPU_EINVAL, "null output");
    if (n == 0)
    {
        std::memset(out32, 0, 32);
        return 0;
    }
    if (!scalars || !points32)
        return fail(HS_GPU_EINVAL, "null input with n > 0");
    Ctx &cx = ctx();
    std::lock_guard<std::mutex> lk(cx.mu);
    int r;
    if ((r = cx.init()) || (r = cx.scalars.reserve(n * 32)) || (r = cx.points.reserve(n * sizeof(hs_gpu_jacobian))) ||
        (r = cx.status.reserve(n * 32)) || (r = cx.out32.reserve(64)))
        return r;
    hipStream_t st = cx.stream;
    // the inputs go up through a staging slot's pinned buffers; the slot is returned on every exit
    // (cx.mu is held throughout)
    HostStage *hsg = take_stage(cx);
    if (!hsg)
        return HS_GPU_EDEVICE;
    struct StageBack
    {
        Ctx &cx;
        HostStage *h;
        ~StageBack() { cx.stages_free.push_back(h); }
    } back{cx, hsg};
    r = hsg->upload(cx.scalars.p, scalars, n * 32, st);
    if (!r)
        r = hsg->upload(cx.status.p, points32, n * 32, st);
    if (flags & HS_GPU_ERASE)
        hsg->erase_pinned();
    if (r)
        return r;
    uint32_t nbad = 0;
    r = curve == HS_GPU_HELIOS
            ? decompress_device<hs::HeliosC>(cx, cx.status.as<uint32_t>(), n, cx.points.as<uint64_t>(), st, &nbad)
            : decompress_device<hs::SeleneC>(cx, cx.status.as<uint32_t>(), n, cx.points.as<uint64_t>(), st, &nbad);
    if (r)
        return r;
    HS_CHECK(hipStreamSynchronize(st));
    if (nbad)
        return fail(HS_GPU_EDECODE, std::to_string(nbad) + " point(s) failed to decode");
    r = msm_device_any(curve, cx, cx.scalars.p, cx.points.p, n, flags, cx.out32.p, nullptr, st);
    if (flags & HS_GPU_ERASE)
        zero_dev(cx.scalars.p, n * 32, st); // the device copy of the scalars
    if (r)
    {
        hipStreamSynchronize(st);
        return r;
    }
    HS_CHECK(hipMemcpyAsync(out32, cx.out32.p, 32, hipMemcpyDeviceToHost, st));
    HS_CHECK(hipStreamSynchronize(st));
    return 0;
}

const char *hs_gpu_version(void) { return "hs_gpu 0.1 gfx950"; }

int hs_gpu_window_bits(int curve, size_t n) { return make_plan(curve, n).c; }

int hs_gpu_reference_window_bits(size_t n) { return ref_window(n); }

int hs_gpu_fixed_create(int curve, const hs_gpu_jacobian *points, size_t n, hs_gpu_fixed **out)
{
    if (!out || !points)
        return fail(HS_GPU_EINVAL, "null argument");
    hs_gpu_fixed *f = nullptr;
    int r;
    if ((r = fixed_alloc(&f, curve, n)))
        return r;
    hipStream_t st = ctx().stream;
    if (hipMemcpyAsync(f->points.p, points, n * sizeof(hs_gpu_jacobian), hipMemcpyHostToDevice, st) != hipSuccess)
        r = fail(HS_GPU_EDEVICE, "base upload failed");
    if (!r)
        r = curve == HS_GPU_HELIOS ? fixed_build_tables<hs::HeliosC>(*f, nullptr, st)
                                   : fixed_build_tables<hs::SeleneC>(*f, nullptr, st);
    if (!r && hipStreamSynchronize(st) != hipSuccess)
        r = fail(HS_GPU_EDEVICE, "table build failed");
    if (r)
    {
        hs_gpu_fixed_destroy(f);
        return r;
    }
    *out = f;
    return 0;
}

int hs_gpu_fixed_create_device(int curve, const void *d_points, size_t n, void *stream, hs_gpu_fixed **out)
{
    if (!out || !d_points)
        return fail(HS_GPU_EINVAL, "null argument");
    hs_gpu_fixed *f = nullptr;
    int r;
    if ((r = fixed_alloc(&f, curve, n)))
        return r;
    hipStream_t st = stream ? (hipStream_t)stream : ctx().stream;
    if (hipMemcpyAsync(f->points.p, d_points, n * sizeof(hs_gpu_jacobian), hipMemcpyDeviceToDevice, st) != hipSuccess)
        r = fail(HS_GPU_EDEVICE, "base copy failed");
    if (!r)
        r = curve == HS_GPU_HELIOS ? fixed_build_tables<hs::HeliosC>(*f, nullptr, st)
                                   : fixed_build_tables<hs::SeleneC>(*f, nullptr, st);
    if (!r && hipStreamSynchronize(st) != hipSuccess)
        r = fail(HS_GPU_EDEVICE, "table build failed");
    if (r)
    {
        hs_gpu_fixed_destroy(f);
        return r;
    }
    *out = f;
    return 0;
}

size_t hs_gpu_fixed_size(const hs_gpu_fixed *f) { return f ? f->n : 0; }

int hs_gpu_fixed_msm_device(hs_gpu_fixed *f, const void *d_scalars, size_t count, unsigned flags, int method,
                            void *d_out32, void *d_out_jac, void *stream)
{
    if (!f || (count && !d_scalars))
        return fail(HS_GPU_EINVAL, "null argument");
    if (count == 0)
        return 0;
    std::lock_guard<std::mutex> lk(f->mu);
    return fixed_enqueue(*f, d_scalars, count, flags, method, d_out32, d_out_jac,
                         stream ? (hipStream_t)stream : ctx().stream);
}

int hs_gpu_fixed_msm(hs_gpu_fixed *f, const unsigned char *scalars, size_t count, unsigned flags, int method,
                     unsigned char *out32, hs_gpu_jacobian *results)
{
    if (!f)
        return fail(HS_GPU_EINVAL, "null fixed-base set");
    std::lock_guard<std::mutex> lk(f->mu);
    return fixed_msm_host(*f, scalars, count, flags, method, out32, results);
}

int hs_gpu_pedersen_commit_batch(hs_gpu_fixed *f, const unsigned char *blindings, const unsigned char *values,
                                 size_t count, unsigned flags, int method, unsigned char *out32, hs_gpu_jacobian *results)
{
    if (!f || (count && (!blindings || (f->n > 1 && !values))))
        return fail(HS_GPU_EINVAL, "null argument");
    const size_t nv = f->n - 1;
    std::vector<unsigned char> rows(count * f->n * 32);
    for (size_t k = 0; k < count; k++)
    {
        std::memcpy(&rows[k * f->n * 32], blindings + 32 * k, 32);
        if (nv)
            std::memcpy(&rows[k * f->n * 32 + 32], values + k * nv * 32, nv * 32);
    }
    std::lock_guard<std::mutex> lk(f->mu);
    const int r = fixed_msm_host(*f, rows.data(), count, flags | HS_GPU_ERASE, method, out32, results);
    secure_erase(rows.data(), rows.size());
    return r;
}

void hs_gpu_fixed_destroy(hs_gpu_fixed *f)
{
    if (!f)
        return;
    if (f->pip)
        hs_gpu_batch_destroy(f->pip);
    for (DevBuf *d : {&f->tables, &f->points, &f->scal, &f->sums})
        d->release();
    delete f;
}

int hs_gpu_fixed_precompute(int curve, const hs_gpu_jacobian *points, size_t n, hs_gpu_affine *tables)
{
    if (n == 0)
        return 0;
    if (!points || !tables)
        return fail(HS_GPU_EINVAL, "null argument");
    hs_gpu_fixed *f = nullptr;
    int r;
    if ((r = fixed_alloc(&f, curve, n)))
        return r;
    DevBuf ref;
    hipStream_t st = ctx().stream;
    const size_t bytes = n * hs::FIXED_ENTRIES * sizeof(hs_gpu_affine);
    if (!(r = ref.reserve(bytes)))
    {
        if (hipMemcpyAsync(f->points.p, points, n * sizeof(hs_gpu_jacobian), hipMemcpyHostToDevice, st) != hipSuccess)
            r = fail(HS_GPU_EDEVICE, "base upload failed");
        if (!r)
            r = curve == HS_GPU_HELIOS ? fixed_build_tables<hs::HeliosC>(*f, ref.as<uint64_t>(), st)
                                       : fixed_build_tables<hs::SeleneC>(*f, ref.as<uint64_t>(), st);
        if (!r && hipMemcpyAsync(tables, ref.p, bytes, hipMemcpyDeviceToHost, st) != hipSuccess)
            r = fail(HS_GPU_EDEVICE, "table copy failed");
        if (!r && hipStreamSynchronize(st) != hipSuccess)
            r = fail(HS_GPU_EDEVICE, "table build failed");
        ref.release();
    }
    hs_gpu_fixed_destroy(f);
    return r;
}

void hs_gpu_helios_msm_fixed(hs_gpu_jacobian *result, const unsigned char *scalars, const hs_gpu_affine *const *tables,
                             size_t n)
{
    slot_msm_fixed(HS_GPU_HELIOS, result, scalars, tables, n);
}

void hs_gpu_selene_msm_fixed(hs_gpu_jacobian *result, const unsigned char *scalars, const hs_gpu_affine *const *tables,
                             size_t n)
{
    slot_msm_fixed(HS_GPU_SELENE, result, scalars, tables, n);
}

void hs_gpu_helios_pedersen_commit(hs_gpu_jacobian *result, const unsigned char *blinding, const hs_gpu_jacobian *H,
                                   const unsigned char *values, const hs_gpu_jacobian *generators, size_t n)
{
    slot_pedersen(HS_GPU_HELIOS, result, blinding, H, values, generators, n);
}

void hs_gpu_selene_pedersen_commit(hs_gpu_jacobian *result, const unsigned char *blinding, const hs_gpu_jacobian *H,
                                   const unsigned char *values, const hs_gpu_jacobian *generators, size_t n)
{
    slot_pedersen(HS_GPU_SELENE, result, blinding, H, values, generators, n);
}


int hs_gpu_autotune(int curve, int mode, int lg_lo, int lg_hi, const char *path)
{
    if ((curve != HS_GPU_HELIOS && curve != HS_GPU_SELENE) || mode < 0 || mode > 1 || lg_lo < 6 || lg_hi > 22 ||
        lg_lo > lg_hi)
        return fail(HS_GPU_EINVAL, "autotune: bad curve, mode or size range (2^6 .. 2^22)");
    Ctx &cx = ctx();
    int r;
    {
        std::lock_guard<std::mutex> lk(cx.mu);
        if ((r = cx.init()))
            return r;
    }
    const int cmax = mode == 0 ? 16 : 13;
    TuneInputs in;
    const size_t nmax = (size_t)1 << lg_hi;
    // a batch of MSMs of 2^lg points reads the same 2^18-point inputs (or all of them, if larger)
    const size_t ninp = mode == 0 ? nmax : std::max(nmax, (size_t)1 << 18);
    if ((r = in.make(curve, ninp, cx.stream)))
        return r;
    for (int lg = lg_lo; lg <= lg_hi; lg++)
    {
        const size_t n = (size_t)1 << lg;
        const size_t k = mode == 0 ? 1 : std::max<size_t>(4, std::min<size_t>(256, ((size_t)1 << 19) >> lg));
        TuneEntry best;
        best.c = mode == 0 ? choose_c(n) : choose_c_batch(n);
        best.L = mode == 0 ? (int)plain_default_L(n) : 32;
        best.G = mode == 0 ? (int)plain_default_G(n) : 8;
        if ((r = tune_time(mode, curve, n, k, ninp, in, best, 3, best.ms)))
            return r;
        // coordinate descent: window bits, then run length, then segment size
        // a candidate replaces the current choice only when it is faster by more than the timing
        // noise (3%), so the defaults stay unless something is measurably better
        auto try_cand = [&](TuneEntry cand) -> int {
            float ms;
            int rr = tune_time(mode, curve, n, k, ninp, in, cand, 3, ms);
            if (!rr && ms < 0.97f * best.ms)
            {
                best = cand;
                best.ms = ms;
            }
            return rr;
        };
        const int c0 = best.c;
        for (int dc : {-2, -1, 1})
        {
            TuneEntry cand = best;
            cand.c = c0 + dc;
            if (cand.c >= 4 && cand.c <= cmax && (r = try_cand(cand)))
                return r;
        }
        const int L0 = best.L;
        for (int L : {8, 16, 32, 64})
        {
            TuneEntry cand = best;
            cand.L = L;
            if (L != L0 && (r = try_cand(cand)))
                return r;
        }
        const int G0 = best.G;
        for (int G : {2, 4, 8, 16})
        {
            TuneEntry cand = best;
            cand.G = G;
            if (G != G0 && (r = try_cand(cand)))
                return r;
        }
        {
            TuneTable &t = tune();
            std::lock_guard<std::mutex> lk(t.mu);
            t.loaded_env = true;
            t.e[mode * 1024 + curve * 64 + lg] = best;
        }
        if (path)
        {
            FILE *f = std::fopen(path, "a");
            if (!f)
                return fail(HS_GPU_EINVAL, std::string("cannot write tuning file ") + path);
            std::fprintf(f, "%d %d %d %d %d %d %.4f\n", mode, curve, lg, best.c, best.L, best.G, best.ms);
            std::fclose(f);
        }
    }
    return 0;
}

int hs_gpu_tune_load(const char *path)
{
    if (!path)
        return fail(HS_GPU_EINVAL, "null path");
    const int k = tune_load_file(path);
    {
        TuneTable &t = tune();
        std::lock_guard<std::mutex> lk(t.mu);
        t.loaded_env = true;
    }
    return k < 0 ? k : 0;
}

int hs_gpu_tune_get(int curve, int mode, size_t n, int *c, int *L, int *G)
{
    if (!c || !L || !G)
        return fail(HS_GPU_EINVAL, "null argument");
    TuneEntry te;
    if (mode == 0)
    {
        const Plan p = make_plan(curve, n);
        *c = p.c;
        *L = (int)p.L;
        *G = (int)p.G;
        return 0;
    }
    const bool tuned = tune_lookup(1, curve, n, te);
    *c = tuned ? te.c : choose_c_batch(n);
    *L = tuned ? te.L : 32;
    *G = tuned ? te.G : 8;
    return 0;
}

int hs_gpu_tune_clear(void)
{
    TuneTable &t = tune();
    std::lock_guard<std::mutex> lk(t.mu);
    t.e.clear();
    t.loaded_env = true;
    return 0;
}

} // extern "C"

