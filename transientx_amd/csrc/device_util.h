// device_util.h -- small gfx950 device helpers shared by the kernel files.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace txgpu {

// Copies len 32-bit words from global memory to LDS with the gfx950 direct
// global->LDS load (no register round trip, many loads in flight).  Wave
// `wave` of `nw` copies the 64-word chunks wave, wave+nw, ...; the LDS
// destination of an instruction is wave-uniform and each lane lands at
// +4*lane.  The caller waits for the loads and barriers before reading.
__device__ inline void lds_fill(const uint32_t *__restrict__ src, uint32_t *dst, int len, int wave,
                                int nw, int lane)
{
    for (int c = wave * 64; c < len; c += nw * 64)
        if (c + lane < len) __builtin_amdgcn_global_load_lds(src + c + lane, dst + c, 4, 0, 0);
}

// The same with 16 bytes per lane (gfx950's global_load_lds_dwordx4; the
// source need only be 4-byte aligned): chunks of 256 words, len rounded up
// to a multiple of 4 words, so the destination must have room for that and
// start 16-byte aligned, and the source must be readable that far.
__device__ inline void lds_fill16(const uint32_t *__restrict__ src, uint32_t *dst, int len, int wave,
                                  int nw, int lane)
{
    for (int c = wave * 256; c < len; c += nw * 256)
        if (c + 4 * lane < len) __builtin_amdgcn_global_load_lds(src + c + 4 * lane, dst + c, 16, 0, 0);
}

__device__ inline void lds_fill_done()
{
    __builtin_amdgcn_s_waitcnt(0); // vmcnt(0): every global->LDS load has landed
    __syncthreads();
}

// Wavefront sum in FP64, fixed butterfly order (deterministic).
__device__ inline double wave_sum_d(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// One DPP move of a double (two 32-bit moves with the same control): lanes
// the control gives no source, or rows outside row_mask, read 0.
template <int CTRL, int ROW_MASK, bool BOUND_ZERO>
__device__ __forceinline__ double dpp_d(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROW_MASK, 0xF, BOUND_ZERO);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xF, BOUND_ZERO);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Inclusive wavefront scan in FP64 with DPP (gfx9 controls): shifts by 1, 2,
// 4 and 8 lanes inside each row of 16, then row 0's total into row 1 and
// row 2's into row 3 (row_bcast:15), then lane 31's into rows 2 and 3
// (row_bcast:31).  Fixed order, every lane active.
__device__ __forceinline__ double wave_scan_incl_d(double x)
{
    x += dpp_d<0x111, 0xF, true>(x); // row_shr:1
    x += dpp_d<0x112, 0xF, true>(x); // row_shr:2
    x += dpp_d<0x114, 0xF, true>(x); // row_shr:4
    x += dpp_d<0x118, 0xF, true>(x); // row_shr:8
    x += dpp_d<0x142, 0xA, false>(x); // row_bcast:15
    x += dpp_d<0x143, 0xC, false>(x); // row_bcast:31
    return x;
}

// Lane l - 1's value (0 on lane 0): wave_shr:1.
__device__ __forceinline__ double wave_shr1_d(double x) { return dpp_d<0x138, 0xF, true>(x); }

typedef float f4v __attribute__((ext_vector_type(4)));

// Four floats of LDS at a 16-byte aligned address, as one 128-bit read.
__device__ __forceinline__ f4v lds_f4(const float *p) { return *(const f4v *)__builtin_assume_aligned(p, 16); }

// Reads a table the kernel never writes through the constant address space,
// so that uniform indices become scalar (SMEM) loads into SGPRs even where
// the kernel also stores to global memory or LDS.
typedef __attribute__((address_space(4))) const int const_int_t;
__device__ inline int ldc(const int *p, long i) { return ((const_int_t *)p)[i]; }

// XCD-aware workgroup order.  The dispatcher hands workgroup b to XCD b mod 8;
// this returns a logical index such that each XCD receives one contiguous run
// of logical indices, so neighbouring tiles (which share input) meet in the
// same L2.  A bijection of [0, n).
__device__ inline unsigned xcd_logical(unsigned b, unsigned n)
{
    const unsigned x = b & 7u, q = n >> 3, rem = n & 7u;
    return x * q + (x < rem ? x : rem) + (b >> 3);
}

} // namespace txgpu
