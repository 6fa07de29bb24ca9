// Shared helpers for the mifdtd HIP kernels and their C ABI (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/mifdtd.h"

namespace mifdtd {

// ---------------------------------------------------------------- error plumbing
void set_error(const std::string &msg);
int fail(int code, const char *fmt, ...);
int hip_status(hipError_t e, const char *where);

#define MIFDTD_CHECK_LAUNCH(where)                                  \
    do {                                                            \
        hipError_t e_ = hipGetLastError();                          \
        if (e_ != hipSuccess) return ::mifdtd::hip_status(e_, where); \
    } while (0)

inline bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

inline hipStream_t to_stream(mifdtd_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// Validates the common (nx, ny, nz_pitch, pad) geometry; returns 0 or an error code.
int check_grid(int nx, int ny, int nzp, int pad);

// ---------------------------------------------------------------- device helpers
struct Q4 {
    float v[4];
};

__device__ __forceinline__ Q4 ld4(const float *__restrict__ p)
{
    const float4 t = *reinterpret_cast<const float4 *>(p);
    Q4 q;
    q.v[0] = t.x;
    q.v[1] = t.y;
    q.v[2] = t.z;
    q.v[3] = t.w;
    return q;
}

__device__ __forceinline__ void st4(float *__restrict__ p, const Q4 &q)
{
    *reinterpret_cast<float4 *>(p) = make_float4(q.v[0], q.v[1], q.v[2], q.v[3]);
}

__device__ __forceinline__ Q4 splat4(float x)
{
    Q4 q;
    q.v[0] = q.v[1] = q.v[2] = q.v[3] = x;
    return q;
}

}  // namespace mifdtd
