// risemesh_internal.h -- library-internal access to a read .risemesh file.
#pragma once

#include "rise_hip_risemesh.h"

#include <vector>

namespace rise_hip {

struct Bvh2Tree;

// The file's BVH2 cache when present and valid, else NULL (owned by the mesh).
const Bvh2Tree* risemesh_tree(const rise_hip_risemesh* m);
// The triangles as 9 doubles each (the file's points), in file order.
std::vector<double> risemesh_triangles_f64(const rise_hip_risemesh* m);
bool risemesh_double_sided(const rise_hip_risemesh* m);

}  // namespace rise_hip
