// lr_internal.h -- handle layouts shared by the library's HIP translation
// units.  Not part of the C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <unordered_map>
#include <vector>

#include "../../include/liar_amd.h"
#include "lr_build.h"
#include "lr_util.h"

// Grow-only device scratch keyed by stream, so that calls on distinct
// streams never share a buffer.
struct lr_scratch {
    void* ptr = nullptr;
    size_t bytes = 0;
};

struct lr_context {
    int device = 0;
    int n_cu = 0;
    lr_trace_config cfg{256, 8, 8, 32};
    std::mutex mu;
    std::unordered_map<hipStream_t, lr_scratch> spill;

    // Returns a device buffer of at least `bytes` for `stream` (locked).
    int scratch(hipStream_t stream, size_t bytes, void** out);
};

struct lr_qbvh {
    lr_context* ctx = nullptr;
    lr_node* d_nodes = nullptr;
    lr_prim* d_prims = nullptr;
    lr::BuiltQbvh host;
    lr_qbvh_info info{};
};

namespace lr {
// Sets the device of `ctx` current on the calling thread.
int bind(const lr_context* ctx);

// Closest-hit over a device ray queue whose length (<= n_max) is read from
// device memory `n_dev` by the kernel; no host synchronisation.
int trace_closest_queue(lr_qbvh* bvh, const lr_rays_soa* rays, uint32_t n_max, const uint32_t* n_dev,
                        lr_hits_soa* hits, hipStream_t stream);
}
