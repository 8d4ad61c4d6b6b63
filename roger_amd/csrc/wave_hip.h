// Wavefront primitives used by the tier-1 cores (t1_core.h), mapped onto the
// gfx950 intrinsics.  The cores are written against these names only, so the
// same source also builds for the host against a lock-step lane emulator
// (tests/emu/wave_emu.h) that checks the kernels' logic on the CPU.
//
// Contract shared by both mappings: one wavefront of 64 lanes per workgroup
// (the tier-1 kernels launch 64 threads; the MQ coders' two-wavefront form
// launches 128, wv_wave() telling the two apart), and every cross-lane call
// is made by all 64 lanes of a wavefront at the same program point
// (wave-uniform control flow).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define RG_DEV __device__ __forceinline__
#define RG_CONST __constant__

using i32x4 = int4;
using u32x4 = uint4;

RG_DEV i32x4 mk_i32x4(int x, int y, int z, int w) { return make_int4(x, y, z, w); }
RG_DEV u32x4 mk_u32x4(uint32_t x, uint32_t y, uint32_t z, uint32_t w) { return make_uint4(x, y, z, w); }
RG_DEV int wv_lane() { return (int)(threadIdx.x & 63); }
// wavefront of the workgroup (0 in the one-wavefront kernels)
RG_DEV int wv_wave() { return (int)(threadIdx.x >> 6); }
RG_DEV void wv_sync() { __syncthreads(); }
// barrier of the whole workgroup (every wavefront must reach it)
RG_DEV void wg_sync() { __syncthreads(); }
RG_DEV int wv_clz(uint32_t x) { return __clz((int)x); }
RG_DEV bool wv_any(bool p) { return __any(p); }
RG_DEV uint64_t wv_ballot(bool p) { return __ballot(p); }
// value of lane j (j uniform)
RG_DEV int wv_readlane(int v, int j) { return __builtin_amdgcn_readlane(v, j); }
// v with lane j's value replaced by x (j and x uniform)
RG_DEV int wv_writelane(int v, int j, int x) { return wv_lane() == j ? x : v; }
// a value the caller knows to be wave-uniform, moved to a scalar register
RG_DEV int wv_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
template <class T> RG_DEV T wv_shfl_up(T v, int d) { return __shfl_up(v, d, 64); }
template <class T> RG_DEV T wv_shfl_down(T v, int d) { return __shfl_down(v, d, 64); }
template <class T> RG_DEV T wv_shfl_xor(T v, int m) { return __shfl_xor(v, m, 64); }
// Neighbour exchange across the whole wavefront as a DPP row operation (no
// LDS round trip): lane i receives lane i-1's value (lane 0 receives 0) /
// lane i+1's value (lane 63 receives 0).
RG_DEV uint32_t wv_from_left(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x138 /* wave_shr:1 */, 0xF, 0xF, true);
}
RG_DEV uint32_t wv_from_right(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x130 /* wave_shl:1 */, 0xF, 0xF, true);
}
// Inclusive prefix sum over the wavefront's lanes (lane i: v[0] + ... + v[i]),
// by DPP alone: Hillis-Steele inside each 16-lane row (row_shr 1, 2, 4, 8
// with zero fill), then row 0's total into row 1 and row 2's into row 3
// (row_bcast:15 to rows 1 and 3), then the first half's total into rows 2
// and 3 (row_bcast:31).
RG_DEV int wv_incl_scan(int v)
{
    v += __builtin_amdgcn_update_dpp(0, v, 0x111 /* row_shr:1 */, 0xF, 0xF, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x112 /* row_shr:2 */, 0xF, 0xF, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x114 /* row_shr:4 */, 0xF, 0xF, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x118 /* row_shr:8 */, 0xF, 0xF, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x142 /* row_bcast:15 */, 0xA, 0xF, false);
    v += __builtin_amdgcn_update_dpp(0, v, 0x143 /* row_bcast:31 */, 0xC, 0xF, false);
    return v;
}
// 24-bit multiply (v_mul_u32_u24): low 32 bits of a * b for a, b < 2^24.
RG_DEV uint32_t wv_mul24(uint32_t a, uint32_t b) { return __umul24(a, b); }
// Byte permute (v_perm_b32): byte i of the result is byte s.byte[i] of the
// 8-byte value hi:lo (selector 12 gives 0x00).
RG_DEV uint32_t wv_perm(uint32_t hi, uint32_t lo, uint32_t s) { return __builtin_amdgcn_perm(hi, lo, s); }
// Compiler scheduling fence: no instruction moves across it (used to issue
// a memory read before independent work that should hide its latency).
RG_DEV void wv_sched_fence() { __builtin_amdgcn_sched_barrier(0); }
// Byte p of read-only global memory through the scalar cache (p wave-
// uniform): the aligned dword holding it is loaded, so the read never leaves
// the page of a valid byte.
RG_DEV uint32_t wv_sload_u8(const uint8_t *p)
{
    const uintptr_t a = (uintptr_t)p;
    const uint32_t w = *(const __attribute__((address_space(4))) uint32_t *)(a & ~(uintptr_t)3);
    return (w >> (8 * (a & 3))) & 0xFFu;
}
// Atomic OR into workgroup memory.
RG_DEV void wv_lds_or(uint32_t *p, uint32_t v) { atomicOr(p, v); }
// Number of set bits of a 64-bit lane mask below this lane.
RG_DEV int wv_mbcnt(uint64_t m)
{
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Byte stores into [base, base + size) through a raw buffer resource: the
// hardware range check discards a store whose lane offset lies outside, so a
// lane with nothing to write aims at SINK_DROP instead of branching on the
// exec mask (base and size wave-uniform).  wv_sink_store writes byte v at
// off + k, k a small constant added after the range check (the scalar offset
// field), so off + k must lie inside whenever off does.
struct ByteSink {
    __amdgpu_buffer_rsrc_t r;
};
constexpr int SINK_DROP = 0x40000000;   // beyond every sink size
RG_DEV ByteSink wv_byte_sink(uint8_t *base, int size)
{
    return ByteSink{__builtin_amdgcn_make_buffer_rsrc(base, (short)0, size, 0x00020000)};
}
RG_DEV void wv_sink_store(const ByteSink &s, int off, int k, uint32_t v)
{
    __builtin_amdgcn_raw_buffer_store_b8((uint8_t)v, s.r, off, k, 0);
}

// 16-byte global-memory load / store (p 16-byte aligned).  The explicit global
// address space gives global_load/store (vector-memory counter only) where a
// generic pointer would give flat instructions, whose completion every later
// LDS wait would also have to cover.
typedef unsigned int rg_v4u __attribute__((ext_vector_type(4)));
RG_DEV u32x4 wv_load16(const uint8_t *p)
{
    const rg_v4u v = *(const __attribute__((address_space(1))) rg_v4u *)p;
    return make_uint4(v.x, v.y, v.z, v.w);
}
RG_DEV void wv_store16(uint8_t *p, uint32_t a, uint32_t b, uint32_t c, uint32_t d)
{
    const rg_v4u v = {a, b, c, d};
    *(__attribute__((address_space(1))) rg_v4u *)p = v;
}
