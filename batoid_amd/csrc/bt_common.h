// bt_common.h -- host-side helpers shared by the library's translation units:
// error reporting, launch geometry, XCD-aware block mapping.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include "../../include/batoid_amd.h"

namespace bt {

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
// Checks the launch that was just issued; returns 0 or the hipError_t.
int check_launch(const char* what);

constexpr int kBlock = 256;
constexpr int kXcds = 8;   // MI355X: 8 XCDs, blocks dispatched round-robin

inline unsigned grid_for(int64_t n, int block = kBlock)
{
    int64_t g = (n + block - 1) / block;
    return (unsigned)(g < 1 ? 1 : g);
}

// Workgroups are dispatched to the 8 XCDs round-robin (block b -> XCD b%8).
// Remap so each XCD works on one contiguous range of logical blocks: rays
// that are neighbours in the input (and on the pupil) then share an L2.
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nblocks)
{
    const unsigned q = nblocks / kXcds, r = nblocks % kXcds;
    const unsigned xcd = b % kXcds, slot = b / kXcds;
    return xcd * q + (xcd < r ? xcd : r) + slot;
}

struct Xform {
    double dr[3];
    double drot[9];
};

inline Xform make_xform(const double* dr, const double* drot)
{
    Xform x;
    for (int i = 0; i < 3; i++) x.dr[i] = dr[i];
    for (int i = 0; i < 9; i++) x.drot[i] = drot[i];
    return x;
}

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline bool rays_ok(const bt_rays& r)
{
    return r.x && r.y && r.z && r.vx && r.vy && r.vz && r.t && r.wavelength &&
           r.flux && r.vignetted && r.failed && r.n >= 0;
}

}  // namespace bt
