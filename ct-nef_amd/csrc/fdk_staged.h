// Internal entry into the LDS-staged voxel-driven backprojection of fdk.hip
// for the voxel-driven op without weights (pixel.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace nefct_internal {

// The pixel_3d op on the staged 16 x 16 x 16 mapping: per column and view
// the detector row coordinate in FP64 (a flat detector's column too; a
// cylindrical one's from the in-plane angle by FP32 atan2, as the op), the
// op's integer n/2 and half-cell offsets, no distance weight, a view skipped where the voxel's
// 2 x 2 neighbourhood leaves the detector.  Returns false, launching nothing,
// when the geometry is not stageable (a detector side below 2 or beyond the
// 24-bit row offsets, or a voxel centre within 1e-3 SID of the source's
// plane); the caller then gathers.
bool pixel_bp_staged(bool cyl, const float* pv, const int grid[3], const float center[3], const float size[3],
                     const float* angles, int nv, float SID, float SAD, int na, float da, float ai, int nb, float db,
                     float bi, float* image, hipStream_t stream);

}  // namespace nefct_internal
