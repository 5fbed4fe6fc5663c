// api_common.h -- error plumbing shared by the C ABI translation units.
#pragma once

#include "rise_hip.h"

#include <hip/hip_runtime_api.h>

#include <string>
#include <vector>

namespace rise_hip {

// Records msg as the calling thread's last error and returns s.
rise_hip_status fail(rise_hip_status s, const std::string& msg);
rise_hip_status fail_hip(hipError_t e, const char* what);
void clear_error();

// Host threads for the library's own pools (BVH build, table generation,
// host-side checks): OMP_NUM_THREADS when set, else the hardware threads, at
// most 16 either way, so a pool never outgrows the CPU share of one GPU.
int host_threads();

struct WideBVH;
// Uploads a built wide BVH as a ray-cast scene (rise_hip_scene_upload's
// device part); bvh_source goes to the scene's statistics.  tri64: NULL, or
// the leaf records' double vertices (9 per record, leaf order), which the
// certifier then uses; attr: NULL, or the records' vertex normals and
// texture coordinates (kAttrDoubles per record) for hit records.
rise_hip_status upload_scene(const WideBVH& bvh, int64_t total, int32_t bvh_source, rise_hip_scene** out_scene,
                             const std::vector<double>* tri64 = nullptr, const std::vector<double>* attr = nullptr);

}  // namespace rise_hip
