// Test infrastructure: a lane-by-lane host model of one 64-lane wavefront, supplying the lane
// primitives of helioselene_amd/csrc/rowfield.h (DPP row moves, permlane swaps, 64-bit pieces) so
// the row-distributed field and curve arithmetic runs on the CPU unchanged.  Every addition,
// subtraction, shift and multiply-add checks that it does not wrap; g_emu_wrap counts violations.
// DPP semantics as measured on gfx950: row_shr:k / row_shl:k move within a row of 16 lanes with
// zero fill (bound_ctrl), row_newbcast:k copies lane k of each row; permlane16_swap(a, b) gives
// rows (a0, b0, a2, b2) and (a1, b1, a3, b3), permlane32_swap gives (a0, a1, b0, b1) and
// (a2, a3, b2, b3).
#pragma once
#include <stdint.h>

#define HS_ROW_EMU 1
#define HS_RW inline

namespace hs
{
namespace row
{

extern long g_emu_wrap;

inline void chk(bool ok)
{
    if (!ok)
        g_emu_wrap++;
}

struct bv
{
    bool l[64];
};

struct u32v
{
    uint32_t l[64];
    u32v()
    {
        for (auto &x : l)
            x = 0;
    }
    u32v(uint32_t x)
    {
        for (auto &y : l)
            y = x;
    }
};

struct u64v
{
    uint64_t l[64];
    u64v()
    {
        for (auto &x : l)
            x = 0;
    }
    u64v(uint64_t x)
    {
        for (auto &y : l)
            y = x;
    }
};

#define HS_EMU_LANES for (int i = 0; i < 64; i++)

inline u32v operator+(const u32v &a, const u32v &b)
{
    u32v r;
    HS_EMU_LANES
    {
        r.l[i] = a.l[i] + b.l[i];
        chk(r.l[i] >= a.l[i]);
    }
    return r;
}
inline u32v operator-(const u32v &a, const u32v &b)
{
    u32v r;
    HS_EMU_LANES
    {
        chk(a.l[i] >= b.l[i]);
        r.l[i] = a.l[i] - b.l[i];
    }
    return r;
}
inline u32v operator*(const u32v &a, const u32v &b)
{
    u32v r;
    HS_EMU_LANES
    {
        const uint64_t p = (uint64_t)a.l[i] * b.l[i];
        chk(p >> 32 == 0);
        r.l[i] = (uint32_t)p;
    }
    return r;
}
inline u32v operator&(const u32v &a, const u32v &b)
{
    u32v r;
    HS_EMU_LANES r.l[i] = a.l[i] & b.l[i];
    return r;
}
inline u32v operator>>(const u32v &a, int s)
{
    u32v r;
    HS_EMU_LANES r.l[i] = a.l[i] >> s;
    return r;
}
inline u32v operator<<(const u32v &a, int s)
{
    u32v r;
    HS_EMU_LANES
    {
        r.l[i] = a.l[i] << s;
        chk((r.l[i] >> s) == a.l[i]);
    }
    return r;
}
inline u64v operator|(const u64v &a, const u64v &b)
{
    u64v r;
    HS_EMU_LANES r.l[i] = a.l[i] | b.l[i];
    return r;
}
inline u64v operator<<(const u64v &a, int s)
{
    u64v r;
    HS_EMU_LANES
    {
        r.l[i] = a.l[i] << s;
        chk((r.l[i] >> s) == a.l[i]);
    }
    return r;
}

inline u32v lane_id()
{
    u32v r;
    HS_EMU_LANES r.l[i] = (uint32_t)i;
    return r;
}
template <int K> inline u32v shr(const u32v &x)
{
    u32v r;
    HS_EMU_LANES r.l[i] = (i & 15) >= K ? x.l[i - K] : 0u;
    return r;
}
template <int K> inline u32v shl(const u32v &x)
{
    u32v r;
    HS_EMU_LANES r.l[i] = (i & 15) + K < 16 ? x.l[i + K] : 0u;
    return r;
}
template <int K> inline u32v bcast(const u32v &x)
{
    u32v r;
    HS_EMU_LANES r.l[i] = x.l[(i & ~15) + K];
    return r;
}
inline u32v lo(const u64v &x)
{
    u32v r;
    HS_EMU_LANES r.l[i] = (uint32_t)x.l[i];
    return r;
}
inline u32v hi(const u64v &x)
{
    u32v r;
    HS_EMU_LANES r.l[i] = (uint32_t)(x.l[i] >> 32);
    return r;
}
inline u64v wide(const u32v &x)
{
    u64v r;
    HS_EMU_LANES r.l[i] = x.l[i];
    return r;
}
inline u64v mad64(const u32v &a, const u32v &b, const u64v &c)
{
    u64v r;
    HS_EMU_LANES
    {
        const unsigned __int128 t = (unsigned __int128)a.l[i] * b.l[i] + c.l[i];
        chk((t >> 64) == 0);
        r.l[i] = (uint64_t)t;
    }
    return r;
}
inline u32v mad24(const u32v &a, const u32v &b, const u32v &c)
{
    u32v r;
    HS_EMU_LANES
    {
        chk(a.l[i] < (1u << 24) && b.l[i] < (1u << 24));
        const uint64_t t = (uint64_t)a.l[i] * b.l[i] + c.l[i];
        chk((t >> 32) == 0);
        r.l[i] = (uint32_t)t;
    }
    return r;
}
inline u64v add64(const u64v &a, const u64v &b)
{
    u64v r;
    HS_EMU_LANES
    {
        r.l[i] = a.l[i] + b.l[i];
        chk(r.l[i] >= a.l[i]);
    }
    return r;
}
inline u64v sh29w(const u64v &x)
{
    u64v r;
    HS_EMU_LANES r.l[i] = x.l[i] >> 29;
    return r;
}
inline u32v sh29(const u64v &x)
{
    u32v r;
    HS_EMU_LANES
    {
        chk((x.l[i] >> 61) == 0);
        r.l[i] = (uint32_t)(x.l[i] >> 29);
    }
    return r;
}
inline bv eq(const u32v &a, uint32_t b)
{
    bv r;
    HS_EMU_LANES r.l[i] = a.l[i] == b;
    return r;
}
inline bv lt(const u32v &a, uint32_t b)
{
    bv r;
    HS_EMU_LANES r.l[i] = a.l[i] < b;
    return r;
}
inline bv nz(const u32v &a)
{
    bv r;
    HS_EMU_LANES r.l[i] = a.l[i] != 0;
    return r;
}
inline u32v sel(const bv &c, const u32v &a, const u32v &b)
{
    u32v r;
    HS_EMU_LANES r.l[i] = c.l[i] ? a.l[i] : b.l[i];
    return r;
}
inline u64v sel(const bv &c, const u64v &a, const u64v &b)
{
    u64v r;
    HS_EMU_LANES r.l[i] = c.l[i] ? a.l[i] : b.l[i];
    return r;
}
inline void swap16(u32v &a, u32v &b)
{
    u32v x, y;
    HS_EMU_LANES
    {
        const int row = i >> 4, l = i & 15;
        x.l[i] = (row & 1) ? b.l[((row - 1) << 4) + l] : a.l[i];
        y.l[i] = (row & 1) ? b.l[i] : a.l[((row + 1) << 4) + l];
    }
    a = x;
    b = y;
}
inline void swap32(u32v &a, u32v &b)
{
    u32v x, y;
    HS_EMU_LANES
    {
        x.l[i] = i < 32 ? a.l[i] : b.l[i - 32];
        y.l[i] = i < 32 ? a.l[i + 32] : b.l[i];
    }
    a = x;
    b = y;
}
inline uint32_t readlane(const u32v &x, int l) { return x.l[l]; }
inline u32v load_if(const bv &c, const uint32_t *p, const u32v &idx)
{
    u32v r;
    HS_EMU_LANES r.l[i] = c.l[i] ? p[idx.l[i]] : 0u;
    return r;
}

#undef HS_EMU_LANES

} // namespace row
} // namespace hs
