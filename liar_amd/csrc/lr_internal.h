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
    lr_trace_config cfg{256, 8, 8, 32, LR_TRACE_SCHEDULE_STEP};
    std::mutex mu;
    std::unordered_map<hipStream_t, lr_scratch> spill;
    // resident workgroups per CU of each traversal kernel at its block size
    // (queried once: launches under graph capture make no runtime queries)
    std::unordered_map<const void*, int> occupancy;
    // device address of this GPU's fixed traversal scratch (a file-scope
    // __device__ array of trace.hip), used by lr_trace_closest_noalloc
    void* fixed_scratch = nullptr;

    // Returns a device buffer of at least `bytes` for `stream` (locked).
    int scratch(hipStream_t stream, size_t bytes, void** out);
    // Frees the buffer of `stream` (still valid: called before the stream
    // is destroyed) and forgets the stream.
    void release(hipStream_t stream);
};

struct lr_qbvh {
    lr_context* ctx = nullptr;
    lr_node_q8* d_nodes = nullptr;  // the device traversal's nodes (quantised boxes)
    lr_prim* d_prims = nullptr;
    lr::BuiltQbvh host;
    std::vector<lr_node_q8> host_q8;  // host copy of d_nodes
    lr_qbvh_info info{};
};

namespace lr {
// Sets the device of `ctx` current on the calling thread.
int bind(const lr_context* ctx);

// Closest-hit over a device ray queue whose length (<= n_max) is read from
// device memory `n_dev` by the kernel; no host synchronisation.  root
// (optional, per ray): the node each ray's traversal starts at, for node
// arrays that hold several trees.
int trace_closest_queue(lr_qbvh* bvh, const lr_rays_soa* rays, uint32_t n_max, const uint32_t* n_dev,
                        lr_hits_soa* hits, hipStream_t stream, const uint32_t* root = nullptr);
// Any-hit over a device ray queue of length *n_dev (<= n_max): occluded[i] = 0 / 1.
int trace_any_queue(lr_qbvh* bvh, const lr_rays_soa* rays, uint32_t n_max, const uint32_t* n_dev, uint8_t* occluded,
                    hipStream_t stream);
// Grows the per-stream traversal scratch for queues of up to n_max rays, so
// that later trace_closest_queue calls (e.g. under graph capture) allocate
// nothing.
int trace_reserve(lr_qbvh* bvh, uint32_t n_max, hipStream_t stream);
}

// One Adam step on `batch` rows drawn uniformly with replacement from the
// min(*count, cap) training records (r_in 8, r_dir 3, r_w 1 floats each;
// Philox keyed by seed, counter (row, step, chunk)).  The record count stays
// on the device and nothing is allocated, so the step can be captured.
// lr_guide_infer with the mixture in 48 columns of ld (>= n) floats: value c
// of row r at mixture[c * ld + r].
int lr_guide_infer_columns(lr_guide* g, const float* inputs, uint32_t n, const uint32_t* n_dev, float* mixture,
                           size_t ld, hipStream_t stream);
// The radiance loss weighs each row by the guide's radiance weight times the
// records' mean weight, as lr_guide_prepare_records last computed it.
int lr_guide_train_records(lr_guide* g, const float* r_in, const float* r_dir, const float* r_w,
                           const uint32_t* count, uint32_t cap, uint64_t seed, uint32_t chunk, uint32_t step,
                           uint32_t batch, hipStream_t stream);
// Before a chunk's training steps, over the min(*count, cap) record weights:
// clamps each at weight_clamp times their mean (when weight_clamp > 0), then
// stores the radiance-loss scale (radiance weight times the clamped mean
// weight).  Fixed-order sums; nothing allocated, no host synchronisation.
int lr_guide_prepare_records(lr_guide* g, float* r_w, const uint32_t* count, uint32_t cap, float weight_clamp,
                             hipStream_t stream);
