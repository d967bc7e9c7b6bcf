// Shared device/host helpers for the nefct HIP kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/nefct_hip.h"

#define NEFCT_WAVE 64

// Raw buffer resource over [p, p + nbytes): loads past the end return 0 and
// atomics past the end are dropped by the texture address unit, so a wrong
// index can never fault the device.  Dword 3 = 0x00020000 is the raw-buffer
// format word for the gfx9 family (CDNA3/4).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* p, uint32_t nbytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)nbytes, 0x00020000);
}

__device__ __forceinline__ float buf_load(__amdgpu_buffer_rsrc_t r, int byte_off)
{
    // the b32 builtin returns the raw bits as an unsigned int
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, byte_off, 0, 0));
}

__device__ __forceinline__ void buf_atomic_add(__amdgpu_buffer_rsrc_t r, int byte_off, float v)
{
    __builtin_amdgcn_raw_ptr_buffer_atomic_fadd_f32(v, r, byte_off, 0, 0);
}

__device__ __forceinline__ int wave_min_i32(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, NEFCT_WAVE));
    return v;
}

__device__ __forceinline__ int wave_max_i32(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, NEFCT_WAVE));
    return v;
}

// 1/x in FP64 from the hardware reciprocal and two Newton steps (a few ulp at
// most, against ~10 instructions for a correctly rounded division).
__device__ __forceinline__ double rcp64(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// v unchanged, but opaque to the optimiser: it cannot fuse the product that
// made v into a later addition (an FMA contraction rounds once instead of
// twice), which it may do in one kernel and not in another.
__device__ __forceinline__ float no_contract(float v)
{
    asm("" : "+v"(v));
    return v;
}

#define NEFCT_CHECK_LAUNCH()                                        \
    do {                                                            \
        hipError_t e_ = hipGetLastError();                          \
        if (e_ != hipSuccess) return NEFCT_ERR_LAUNCH;              \
    } while (0)
