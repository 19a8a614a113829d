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
