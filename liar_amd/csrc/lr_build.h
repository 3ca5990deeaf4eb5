// lr_build.h -- host-side QBVH build interface (internal to the library).
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/liar_amd.h"

namespace lr {

struct BuiltQbvh {
    std::vector<lr_node> nodes;
    std::vector<lr_prim> prims;
    uint32_t n_triangles = 0, n_spheres = 0, max_depth = 0;
    double sah_cost = 0.0, build_seconds = 0.0;
    float scene_lo[3] = {0, 0, 0}, scene_hi[3] = {0, 0, 0};
};

// Returns 0 on success, -1 with `err` set on invalid input.
int build_qbvh(const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt, const float* spheres,
               uint32_t ns, BuiltQbvh& out, std::string& err);

// The same SAH build over boxes (n x {lo[3], hi[3]}): fills out.nodes (leaves
// reference box slots in leaf order) and order[slot] = input box index.
int build_boxes(const float* boxes, uint32_t n, BuiltQbvh& out, std::vector<uint32_t>& order, std::string& err);

// The device node format: each node's child boxes quantised to 8 bits per
// plane, rounded outward (lr_node_q8).  Indices and child encodings are
// copied, so this applies to any node array, concatenated ones included.
void quantize_nodes(const lr_node* in, size_t n, lr_node_q8* out);
std::vector<lr_node_q8> quantize_nodes(const std::vector<lr_node>& in);

}  // namespace lr
