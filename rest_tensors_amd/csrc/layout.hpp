// Device strided block movers shared by the entry points.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace rta {

// dst[x*dx + y*dy + z*dz] = src[x*sx + y*sy + z*sz], x<X, y<Y, z<Z
int copy3d(int64_t X, int64_t Y, int64_t Z, const double* src, int64_t sx,
           int64_t sy, int64_t sz, double* dst, int64_t dx, int64_t dy,
           int64_t dz, hipStream_t s);

// dst[...] = src[...] + beta * dst[...] over the same 3-D index space
int axpby3d(int64_t X, int64_t Y, int64_t Z, const double* src, int64_t sx,
            int64_t sy, int64_t sz, double* dst, int64_t dx, int64_t dy,
            int64_t dz, double beta, hipStream_t s);

}  // namespace rta
