// Host/device qualifier used by the few helpers that both the gfx950 kernels
// and the host-side population builder need (Philox stream, project math).
#pragma once
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PS_HD __host__ __device__
#define PS_INLINE __host__ __device__ __forceinline__
#else
#define PS_HD
#define PS_INLINE inline
#endif
