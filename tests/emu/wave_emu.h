// Host lock-step emulator of one 64-lane wavefront, for running the tier-1
// cores (roger_amd/csrc/t1_core.h) on the CPU.
//
// Every lane is a fiber on its own stack.  A cross-lane call (shuffle, any,
// ballot, readlane, barrier) deposits the lane's value and yields; once all
// 64 lanes have arrived the scheduler resumes them in order and each reads the
// deposited values.  Deposits alternate between two buffers, so a lane that
// runs ahead to its next cross-lane call never overwrites values that slower
// lanes still have to read.
//
// Checks made on every cross-lane call (any failure aborts with a message):
//   * all live lanes reach the same call site (wave-uniform control flow),
//   * no lane leaves the kernel while others wait at a cross-lane call,
//   * a value passed to wv_uniform is the same in every lane.
#pragma once
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define RG_DEV inline
#define RG_CONST static const

struct i32x4 {
    int x, y, z, w;
};
struct alignas(16) u32x4 {
    uint32_t x, y, z, w;
};
static inline i32x4 mk_i32x4(int x, int y, int z, int w) { return i32x4{x, y, z, w}; }
static inline u32x4 mk_u32x4(uint32_t x, uint32_t y, uint32_t z, uint32_t w) { return u32x4{x, y, z, w}; }
static inline u32x4 wv_load16(const uint8_t *p)
{
    if (((uintptr_t)p & 15) != 0) { fprintf(stderr, "wave emulator: misaligned 16-byte load\n"); exit(3); }
    u32x4 v;
    memcpy(&v, p, 16);
    return v;
}
// Byte stores into [base, base + size); an offset outside is dropped (the
// kernels' raw buffer range check).
struct ByteSink {
    uint8_t *base;
    int size;
};
constexpr int SINK_DROP = 0x40000000;
static inline ByteSink wv_byte_sink(uint8_t *base, int size) { return ByteSink{base, size}; }
static inline void wv_sink_store(const ByteSink &s, int off, int k, uint32_t v)
{
    if (off < 0 || off >= s.size) return;
    if (off + k >= s.size) { fprintf(stderr, "wave emulator: sink store past its range\n"); exit(3); }
    s.base[off + k] = (uint8_t)v;
}
static inline void wv_store16(uint8_t *p, uint32_t a, uint32_t b, uint32_t c, uint32_t d)
{
    if (((uintptr_t)p & 15) != 0) { fprintf(stderr, "wave emulator: misaligned 16-byte store\n"); exit(3); }
    const u32x4 v{a, b, c, d};
    memcpy(p, &v, 16);
}

namespace emu {
int lane();
// deposits v for the current lane, waits for the wave; returns the 64 deposits
const uint64_t *exchange(uint64_t v, int check_uniform);
}  // namespace emu

static inline int wv_lane() { return emu::lane(); }
static inline void wv_sync() { emu::exchange(0, 0); }
// one emulated wavefront: the workgroup is that wavefront
static inline int wv_wave() { return 0; }
static inline void wg_sync() { emu::exchange(0, 0); }
static inline int wv_clz(uint32_t x) { return x ? __builtin_clz(x) : 32; }
static inline bool wv_any(bool p)
{
    const uint64_t *v = emu::exchange(p, 0);
    for (int l = 0; l < 64; ++l)
        if (v[l]) return true;
    return false;
}
static inline uint64_t wv_ballot(bool p)
{
    const uint64_t *v = emu::exchange(p, 0);
    uint64_t m = 0;
    for (int l = 0; l < 64; ++l) m |= (uint64_t)(v[l] != 0) << l;
    return m;
}
static inline int wv_readlane(int x, int j) { return (int)(uint32_t)emu::exchange((uint32_t)x, 0)[j]; }
static inline int wv_writelane(int v, int j, int x) { return emu::lane() == j ? x : v; }
static inline int wv_uniform(int x) { return (int)(uint32_t)emu::exchange((uint32_t)x, 1)[0]; }
template <class T> static inline T wv_shfl_up(T x, int d)
{
    const uint64_t *v = emu::exchange((uint64_t)x, 0);
    const int s = emu::lane() - d;
    return s >= 0 ? (T)v[s] : x;
}
template <class T> static inline T wv_shfl_down(T x, int d)
{
    const uint64_t *v = emu::exchange((uint64_t)x, 0);
    const int s = emu::lane() + d;
    return s < 64 ? (T)v[s] : x;
}
template <class T> static inline T wv_shfl_xor(T x, int m)
{
    const uint64_t *v = emu::exchange((uint64_t)x, 0);
    return (T)v[emu::lane() ^ m];
}
static inline uint32_t wv_from_left(uint32_t x)
{
    const uint64_t *v = emu::exchange(x, 0);
    return emu::lane() == 0 ? 0u : (uint32_t)v[emu::lane() - 1];
}
static inline uint32_t wv_from_right(uint32_t x)
{
    const uint64_t *v = emu::exchange(x, 0);
    return emu::lane() == 63 ? 0u : (uint32_t)v[emu::lane() + 1];
}
static inline void wv_sched_fence() {}
static inline uint32_t wv_sload_u8(const uint8_t *p) { return *p; }
static inline void wv_lds_or(uint32_t *p, uint32_t v) { *p |= v; }
static inline uint32_t wv_mul24(uint32_t a, uint32_t b) { return (a & 0xFFFFFFu) * (b & 0xFFFFFFu); }
static inline uint32_t wv_perm(uint32_t hi, uint32_t lo, uint32_t s)
{
    const uint64_t v = ((uint64_t)hi << 32) | lo;
    uint32_t r = 0;
    for (int i = 0; i < 4; ++i) {
        const uint32_t k = (s >> (8 * i)) & 0xFFu;
        uint32_t b;
        if (k < 8) b = (uint32_t)(v >> (8 * k)) & 0xFFu;
        else if (k == 12) b = 0;
        else { fprintf(stderr, "wave emulator: unsupported perm selector\n"); exit(3); }
        r |= b << (8 * i);
    }
    return r;
}
static inline int wv_incl_scan(int x)
{
    const uint64_t *v = emu::exchange((uint32_t)x, 0);
    int s = 0;
    for (int l = 0; l <= emu::lane(); ++l) s += (int)(uint32_t)v[l];
    return s;
}
static inline int wv_mbcnt(uint64_t m)
{
    const int l = emu::lane();
    return __builtin_popcountll(l == 0 ? 0ull : (m & (~0ull >> (64 - l))));
}
