// Test infrastructure: plain C entry points over the reference library, which oracle/Makefile
// compiles from its own sources into oracle/_ref/libhsref.so.  Used only by tests/ (as the
// checker) and bench.py (as the CPU baseline); never by the product library.
//
// Point arguments are the reference's own structs (helios_jacobian / selene_jacobian, 120 bytes,
// helios/include/helios.h:41-46) or 32-byte compressed encodings (helios_tobytes.cpp:42-66).
// The bytes forms map 0^32 to the identity on input, like the reference test loader
// (src/tests.cpp:7479-7496).

#include "helioselene_dispatch.h"
#include "helioselene_cpuid.h"
#include "helios.h"
#include "helios_add.h"
#include "helios_constants.h"
#include "helios_dbl.h"
#include "helios_frombytes.h"
#include "helios_map_to_curve.h"
#include "helios_msm_fixed.h"
#include "helios_msm_vartime.h"
#include "helios_ops.h"
#include "helios_pedersen.h"
#include "helios_scalar.h"
#include "helios_scalarmult_fixed.h"
#include "helios_scalarmult_vartime.h"
#include "helios_tobytes.h"
#include "selene.h"
#include "selene_add.h"
#include "selene_constants.h"
#include "selene_dbl.h"
#include "selene_frombytes.h"
#include "selene_map_to_curve.h"
#include "selene_msm_fixed.h"
#include "selene_msm_vartime.h"
#include "selene_ops.h"
#include "selene_pedersen.h"
#include "selene_scalar.h"
#include "selene_scalarmult_fixed.h"
#include "selene_scalarmult_vartime.h"
#include "selene_tobytes.h"

#include <chrono>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace
{
    bool is_zero32(const unsigned char *b)
    {
        unsigned char acc = 0;
        for (int i = 0; i < 32; i++)
            acc |= b[i];
        return acc == 0;
    }

    // One template body per curve; the two curves differ only in the functions bound here.
    struct HeliosOps
    {
        using jac = helios_jacobian;
        using aff = helios_affine;
        static void identity(jac *r) { helios_identity(r); }
        static int frombytes(jac *r, const unsigned char *s) { return helios_frombytes(r, s); }
        static void tobytes(unsigned char *s, const jac *p) { helios_tobytes(s, p); }
        static void msm(jac *r, const unsigned char *sc, const jac *p, size_t n) { helios_msm_vartime(r, sc, p, n); }
        static void smul(jac *r, const unsigned char *s, const jac *p) { helios_scalarmult_vartime(r, s, p); }
        static void add(jac *r, const jac *a, const jac *b) { helios_add(r, a, b); }
        static void dbl(jac *r, const jac *a) { helios_dbl(r, a); }
        static void map(jac *r, const unsigned char *u) { helios_map_to_curve(r, u); }
        static void pedersen(jac *r, const unsigned char *bl, const jac *H, const unsigned char *v, const jac *g, size_t n)
        {
            helios_pedersen_commit(r, bl, H, v, g, n);
        }
        static void precompute(aff *t, const jac *p) { helios_scalarmult_fixed_precompute(t, p); }
        static void msm_fixed(jac *r, const unsigned char *s, const aff *const *t, size_t n) { helios_msm_fixed(r, s, t, n); }
        static void generator(jac *r)
        {
            fp_copy(r->X, HELIOS_GX);
            fp_copy(r->Y, HELIOS_GY);
            fp_1(r->Z);
        }
        static void reduce_wide(unsigned char *out, const unsigned char *w)
        {
            fq_fe t;
            helios_scalar_reduce_wide(t, w);
            helios_scalar_to_bytes(out, t);
        }
    };

    struct SeleneOps
    {
        using jac = selene_jacobian;
        using aff = selene_affine;
        static void identity(jac *r) { selene_identity(r); }
        static int frombytes(jac *r, const unsigned char *s) { return selene_frombytes(r, s); }
        static void tobytes(unsigned char *s, const jac *p) { selene_tobytes(s, p); }
        static void msm(jac *r, const unsigned char *sc, const jac *p, size_t n) { selene_msm_vartime(r, sc, p, n); }
        static void smul(jac *r, const unsigned char *s, const jac *p) { selene_scalarmult_vartime(r, s, p); }
        static void add(jac *r, const jac *a, const jac *b) { selene_add(r, a, b); }
        static void dbl(jac *r, const jac *a) { selene_dbl(r, a); }
        static void map(jac *r, const unsigned char *u) { selene_map_to_curve(r, u); }
        static void pedersen(jac *r, const unsigned char *bl, const jac *H, const unsigned char *v, const jac *g, size_t n)
        {
            selene_pedersen_commit(r, bl, H, v, g, n);
        }
        static void precompute(aff *t, const jac *p) { selene_scalarmult_fixed_precompute(t, p); }
        static void msm_fixed(jac *r, const unsigned char *s, const aff *const *t, size_t n) { selene_msm_fixed(r, s, t, n); }
        static void generator(jac *r)
        {
            fq_copy(r->X, SELENE_GX);
            fq_copy(r->Y, SELENE_GY);
            fq_1(r->Z);
        }
        static void reduce_wide(unsigned char *out, const unsigned char *w)
        {
            fp_fe t;
            selene_scalar_reduce_wide(t, w);
            selene_scalar_to_bytes(out, t);
        }
    };

    template <class O> int decode_points(std::vector<typename O::jac> &out, const unsigned char *pts, size_t n)
    {
        out.resize(n);
        for (size_t i = 0; i < n; i++)
        {
            if (is_zero32(pts + 32 * i))
                O::identity(&out[i]);
            else if (O::frombytes(&out[i], pts + 32 * i) != 0)
                return -1;
        }
        return 0;
    }

    template <class O> int msm_bytes(unsigned char *out, const unsigned char *scalars, const unsigned char *pts, size_t n)
    {
        std::vector<typename O::jac> p;
        if (decode_points<O>(p, pts, n) != 0)
            return -1;
        typename O::jac r;
        O::msm(&r, scalars, p.data(), n);
        O::tobytes(out, &r);
        return 0;
    }

    template <class O> int naive_bytes(unsigned char *out, const unsigned char *scalars, const unsigned char *pts, size_t n)
    {
        std::vector<typename O::jac> p;
        if (decode_points<O>(p, pts, n) != 0)
            return -1;
        typename O::jac acc, t, s;
        O::identity(&acc);
        for (size_t i = 0; i < n; i++)
        {
            O::smul(&t, scalars + 32 * i, &p[i]);
            O::add(&s, &acc, &t);
            acc = s;
        }
        O::tobytes(out, &acc);
        return 0;
    }

    template <class O> int pedersen_bytes(unsigned char *out, const unsigned char *blinding, const unsigned char *H,
                                          const unsigned char *values, const unsigned char *gens, size_t n)
    {
        std::vector<typename O::jac> g, h;
        if (decode_points<O>(h, H, 1) != 0 || decode_points<O>(g, gens, n) != 0)
            return -1;
        typename O::jac r;
        O::pedersen(&r, blinding, h.data(), values, g.data(), n);
        O::tobytes(out, &r);
        return 0;
    }

    template <class O> int msm_fixed_bytes(unsigned char *out, const unsigned char *scalars, const unsigned char *pts, size_t n)
    {
        std::vector<typename O::jac> p;
        if (decode_points<O>(p, pts, n) != 0)
            return -1;
        std::vector<typename O::aff> tables(16 * n);
        std::vector<const typename O::aff *> ptrs(n);
        for (size_t i = 0; i < n; i++)
        {
            O::precompute(&tables[16 * i], &p[i]);
            ptrs[i] = &tables[16 * i];
        }
        typename O::jac r;
        O::msm_fixed(&r, scalars, ptrs.data(), n);
        O::tobytes(out, &r);
        return 0;
    }

    template <class O> void map_bytes(unsigned char *out, const unsigned char *u, size_t n)
    {
        for (size_t i = 0; i < n; i++)
        {
            typename O::jac r;
            O::map(&r, u + 32 * i);
            O::tobytes(out + 32 * i, &r);
        }
    }

    template <class O> void map_jac(typename O::jac *out, const unsigned char *u, size_t n)
    {
        for (size_t i = 0; i < n; i++)
            O::map(&out[i], u + 32 * i);
    }

    // The FCMP++ benchmark's bases: G, 2G, 4G, ... by successive doublings (src/benchmark_fcmpp.cpp:51-69)
    template <class O> void dbl_chain(typename O::jac *out, size_t n)
    {
        if (n == 0)
            return;
        O::generator(&out[0]);
        for (size_t i = 1; i < n; i++)
            O::dbl(&out[i], &out[i - 1]);
    }

    template <class O> void msm_jac_bytes(unsigned char *o, const unsigned char *s, const typename O::jac *p, size_t n)
    {
        typename O::jac r;
        O::msm(&r, s, p, n);
        O::tobytes(o, &r);
    }

    template <class O> int to_jac(typename O::jac *out, const unsigned char *pts, size_t n)
    {
        std::vector<typename O::jac> p;
        if (decode_points<O>(p, pts, n) != 0)
            return -1;
        std::memcpy(out, p.data(), n * sizeof(typename O::jac));
        return 0;
    }

    // Times `reps` whole MSMs over `threads` host threads (each thread runs complete, independent
    // reference MSMs; the reference function is reentrant).  Returns wall seconds.
    template <class O> double time_msm(const unsigned char *scalars, const typename O::jac *pts, size_t n, int reps, int threads)
    {
        auto t0 = std::chrono::steady_clock::now();
        std::vector<std::thread> ws;
        for (int t = 0; t < threads; t++)
            ws.emplace_back([=]() {
                typename O::jac r;
                for (int k = t; k < reps; k += threads)
                    O::msm(&r, scalars, pts, n);
            });
        for (auto &w : ws)
            w.join();
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
} // namespace

extern "C"
{
    unsigned hsref_init(void)
    {
        helioselene_init();
        return helioselene_cpu_features();
    }
    size_t hsref_sizeof_jacobian(void) { return sizeof(helios_jacobian); }

    int hsref_helios_msm_bytes(unsigned char *o, const unsigned char *s, const unsigned char *p, size_t n) { return msm_bytes<HeliosOps>(o, s, p, n); }
    int hsref_selene_msm_bytes(unsigned char *o, const unsigned char *s, const unsigned char *p, size_t n) { return msm_bytes<SeleneOps>(o, s, p, n); }
    int hsref_helios_naive_bytes(unsigned char *o, const unsigned char *s, const unsigned char *p, size_t n) { return naive_bytes<HeliosOps>(o, s, p, n); }
    int hsref_selene_naive_bytes(unsigned char *o, const unsigned char *s, const unsigned char *p, size_t n) { return naive_bytes<SeleneOps>(o, s, p, n); }

    void hsref_helios_msm(helios_jacobian *r, const unsigned char *s, const helios_jacobian *p, size_t n) { helios_msm_vartime(r, s, p, n); }
    void hsref_selene_msm(selene_jacobian *r, const unsigned char *s, const selene_jacobian *p, size_t n) { selene_msm_vartime(r, s, p, n); }
    void hsref_helios_tobytes(unsigned char *o, const helios_jacobian *p) { helios_tobytes(o, p); }
    void hsref_selene_tobytes(unsigned char *o, const selene_jacobian *p) { selene_tobytes(o, p); }
    int hsref_helios_to_jac(helios_jacobian *o, const unsigned char *p, size_t n) { return to_jac<HeliosOps>(o, p, n); }
    int hsref_selene_to_jac(selene_jacobian *o, const unsigned char *p, size_t n) { return to_jac<SeleneOps>(o, p, n); }

    int hsref_helios_pedersen_bytes(unsigned char *o, const unsigned char *b, const unsigned char *H, const unsigned char *v, const unsigned char *g, size_t n)
    {
        return pedersen_bytes<HeliosOps>(o, b, H, v, g, n);
    }
    int hsref_selene_pedersen_bytes(unsigned char *o, const unsigned char *b, const unsigned char *H, const unsigned char *v, const unsigned char *g, size_t n)
    {
        return pedersen_bytes<SeleneOps>(o, b, H, v, g, n);
    }
    int hsref_helios_msm_fixed_bytes(unsigned char *o, const unsigned char *s, const unsigned char *p, size_t n) { return msm_fixed_bytes<HeliosOps>(o, s, p, n); }
    int hsref_selene_msm_fixed_bytes(unsigned char *o, const unsigned char *s, const unsigned char *p, size_t n) { return msm_fixed_bytes<SeleneOps>(o, s, p, n); }

    void hsref_helios_map_to_curve_bytes(unsigned char *o, const unsigned char *u, size_t n) { map_bytes<HeliosOps>(o, u, n); }
    void hsref_selene_map_to_curve_bytes(unsigned char *o, const unsigned char *u, size_t n) { map_bytes<SeleneOps>(o, u, n); }
    void hsref_helios_map_to_curve_jac(helios_jacobian *o, const unsigned char *u, size_t n) { map_jac<HeliosOps>(o, u, n); }
    void hsref_selene_map_to_curve_jac(selene_jacobian *o, const unsigned char *u, size_t n) { map_jac<SeleneOps>(o, u, n); }

    void hsref_helios_dbl_chain(helios_jacobian *o, size_t n) { dbl_chain<HeliosOps>(o, n); }
    void hsref_selene_dbl_chain(selene_jacobian *o, size_t n) { dbl_chain<SeleneOps>(o, n); }
    void hsref_helios_msm_jac_bytes(unsigned char *o, const unsigned char *s, const helios_jacobian *p, size_t n)
    {
        msm_jac_bytes<HeliosOps>(o, s, p, n);
    }
    void hsref_selene_msm_jac_bytes(unsigned char *o, const unsigned char *s, const selene_jacobian *p, size_t n)
    {
        msm_jac_bytes<SeleneOps>(o, s, p, n);
    }

    void hsref_helios_reduce_wide(unsigned char *o, const unsigned char *w) { HeliosOps::reduce_wide(o, w); }
    void hsref_selene_reduce_wide(unsigned char *o, const unsigned char *w) { SeleneOps::reduce_wide(o, w); }

    void hsref_helios_generator(helios_jacobian *r) { HeliosOps::generator(r); }
    void hsref_selene_generator(selene_jacobian *r) { SeleneOps::generator(r); }

    double hsref_helios_time_msm(const unsigned char *s, const helios_jacobian *p, size_t n, int reps, int threads)
    {
        return time_msm<HeliosOps>(s, p, n, reps, threads);
    }
    double hsref_selene_time_msm(const unsigned char *s, const selene_jacobian *p, size_t n, int reps, int threads)
    {
        return time_msm<SeleneOps>(s, p, n, reps, threads);
    }
}
