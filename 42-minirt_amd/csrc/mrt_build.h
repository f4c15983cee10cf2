// BVH construction on the device (mrt_build.hip).
#pragma once

#include "mrt.h"
#include "mrt_host.h"

#include <hip/hip_runtime.h>

namespace mrt {

// Device arrays of a built scene, in the layouts the kernels read: 4-wide
// nodes (8 float4 each, root at 0), primitive records (3 float4 each, leaf
// order) and shading records (9 floats each).  The caller frees them.
struct DeviceBvh {
    float4  *nodes;
    float4  *prims;
    float   *nrm;
    uint32_t n_nodes, n_prims, depth, iterations;
};

// Builds s's BVH on the device on stream st.  MRT_E_UNSUPPORTED (with a
// reason in *why) when the scene is too small for it or the tree would
// exceed the traversal stack's depth bound; the caller then builds on the
// host.
int device_bvh_build(hipStream_t st, const mrt_scene *s, DeviceBvh *out,
                     const char **why);

}  // namespace mrt
