// qbvh_build.cpp -- host SAH builder for the 4-wide QBVH consumed by the
// gfx950 traversal kernels.
//
// Replaces the Lass QbvhTree build that LiAR runs inside the TriangleMesh ctor
// (src/liar/scenery/triangle_mesh.h:53, triangle_mesh.cpp:69-85) and the
// object-level QbvhTree::doPreProcess (src/liar/scenery/object_trees.h:100-108).
//
// Method: binned SAH (32 bins, all three axes) on primitive-AABB centroids.
// A 4-wide node is formed by splitting its range, then repeatedly splitting
// the child range with the largest surface area, until four children exist
// or no child holds more than LR_LEAF_MAX primitives.  Ranges of at most
// LR_LEAF_MAX primitives become leaves.  Past depth 48 the split falls back
// to the object median so that the traversal stack stays bounded.
//
// Large subtrees are built on std::threads; nodes are then renumbered in
// depth-first preorder so the output is identical for any thread schedule.
#include "lr_build.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "lr_geom.h"

namespace lr {
namespace {

constexpr int kBins = 32;
constexpr int kLeafMax = 4;
constexpr int kMedianDepth = 48;
constexpr uint32_t kParallelMin = 1u << 15;

struct Box {
    float lo[3], hi[3];
    void reset() {
        for (int k = 0; k < 3; ++k) {
            lo[k] = INFINITY;
            hi[k] = -INFINITY;
        }
    }
    void grow(const Box& b) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], b.lo[k]);
            hi[k] = std::max(hi[k], b.hi[k]);
        }
    }
    void grow(const float* p) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], p[k]);
            hi[k] = std::max(hi[k], p[k]);
        }
    }
    double area() const {
        const double x = std::max(0.0, (double)hi[0] - lo[0]);
        const double y = std::max(0.0, (double)hi[1] - lo[1]);
        const double z = std::max(0.0, (double)hi[2] - lo[2]);
        return 2.0 * (x * y + y * z + z * x);
    }
};

struct Ref {
    Box box;
    float c[3];
    uint32_t prim;  // index into the input prim list (triangles then spheres)
};

struct Range {
    uint32_t begin, end;
    Box box;  // union of prim boxes
    uint32_t count() const { return end - begin; }
};

struct TmpNode {
    Box box[4];
    uint32_t kind[4];  // 0 empty, 1 inner (child = tmp node index), 2 leaf
    uint32_t child[4];
    uint32_t count[4];
};

struct Builder {
    std::vector<Ref>& refs;
    std::vector<TmpNode> nodes;
    std::atomic<uint32_t> next{0};
    std::atomic<uint32_t> max_depth{0};
    std::atomic<int> threads_left;

    Builder(std::vector<Ref>& r, uint32_t capacity, int threads)
        : refs(r), nodes(capacity), threads_left(threads) {}

    uint32_t alloc() { return next.fetch_add(1); }

    Box centroid_box(const Range& r) const {
        Box cb;
        cb.reset();
        for (uint32_t i = r.begin; i < r.end; ++i) cb.grow(refs[i].c);
        return cb;
    }

    Box prim_box(uint32_t b, uint32_t e) const {
        Box bb;
        bb.reset();
        for (uint32_t i = b; i < e; ++i) bb.grow(refs[i].box);
        return bb;
    }

    // Split r into two non-empty ranges.  Returns false only if r.count() < 2.
    bool split(const Range& r, int depth, Range& left, Range& right) {
        const uint32_t n = r.count();
        if (n < 2) return false;
        uint32_t mid = r.begin + n / 2;
        bool done = false;
        if (depth < kMedianDepth) {
            const Box cb = centroid_box(r);
            double best_cost = INFINITY;
            int best_axis = -1, best_bin = -1;
            Box bin_box[3][kBins];
            uint32_t bin_cnt[3][kBins];
            for (int a = 0; a < 3; ++a) {
                const float ext = cb.hi[a] - cb.lo[a];
                if (!(ext > 0.0f)) continue;
                for (int b = 0; b < kBins; ++b) {
                    bin_box[a][b].reset();
                    bin_cnt[a][b] = 0;
                }
                const float scale = (float)kBins / ext;
                for (uint32_t i = r.begin; i < r.end; ++i) {
                    int b = (int)((refs[i].c[a] - cb.lo[a]) * scale);
                    b = std::min(std::max(b, 0), kBins - 1);
                    bin_box[a][b].grow(refs[i].box);
                    bin_cnt[a][b]++;
                }
                double right_area[kBins];
                uint32_t right_cnt[kBins];
                Box acc;
                acc.reset();
                uint32_t cnt = 0;
                for (int b = kBins - 1; b > 0; --b) {
                    acc.grow(bin_box[a][b]);
                    cnt += bin_cnt[a][b];
                    right_area[b] = acc.area();
                    right_cnt[b] = cnt;
                }
                acc.reset();
                cnt = 0;
                for (int b = 0; b < kBins - 1; ++b) {
                    acc.grow(bin_box[a][b]);
                    cnt += bin_cnt[a][b];
                    if (cnt == 0 || right_cnt[b + 1] == 0) continue;
                    const double cost = acc.area() * cnt + right_area[b + 1] * right_cnt[b + 1];
                    if (cost < best_cost) {
                        best_cost = cost;
                        best_axis = a;
                        best_bin = b;
                    }
                }
            }
            if (best_axis >= 0) {
                const int a = best_axis;
                const float ext = cb.hi[a] - cb.lo[a];
                const float scale = (float)kBins / ext;
                const float lo = cb.lo[a];
                auto it = std::partition(refs.begin() + r.begin, refs.begin() + r.end, [&](const Ref& x) {
                    int b = (int)((x.c[a] - lo) * scale);
                    b = std::min(std::max(b, 0), kBins - 1);
                    return b <= best_bin;
                });
                mid = (uint32_t)(it - refs.begin());
                done = mid > r.begin && mid < r.end;
            }
        }
        if (!done) {
            // Object median along the largest centroid extent (or input order
            // when every centroid coincides).
            mid = r.begin + n / 2;
            const Box cb = centroid_box(r);
            int a = 0;
            float ext = -1.0f;
            for (int k = 0; k < 3; ++k) {
                if (cb.hi[k] - cb.lo[k] > ext) {
                    ext = cb.hi[k] - cb.lo[k];
                    a = k;
                }
            }
            if (ext > 0.0f) {
                std::nth_element(refs.begin() + r.begin, refs.begin() + mid, refs.begin() + r.end,
                                 [a](const Ref& x, const Ref& y) {
                                     if (x.c[a] != y.c[a]) return x.c[a] < y.c[a];
                                     return x.prim < y.prim;
                                 });
            }
        }
        left = {r.begin, mid, prim_box(r.begin, mid)};
        right = {mid, r.end, prim_box(mid, r.end)};
        return true;
    }

    void build(uint32_t node_index, const Range& r, int depth) {
        {
            uint32_t d = (uint32_t)depth;
            uint32_t cur = max_depth.load();
            while (d > cur && !max_depth.compare_exchange_weak(cur, d)) {
            }
        }
        Range parts[4];
        int np = 1;
        parts[0] = r;
        while (np < 4) {
            int pick = -1;
            double best_area = -1.0;
            for (int i = 0; i < np; ++i) {
                if (parts[i].count() > (uint32_t)kLeafMax && parts[i].box.area() > best_area) {
                    best_area = parts[i].box.area();
                    pick = i;
                }
            }
            if (pick < 0) break;
            Range a, b;
            split(parts[pick], depth, a, b);
            parts[pick] = a;
            parts[np++] = b;
        }
        // Keep children in range order for a deterministic layout.
        for (int i = 1; i < np; ++i)
            for (int j = i; j > 0 && parts[j].begin < parts[j - 1].begin; --j) std::swap(parts[j], parts[j - 1]);
        TmpNode& nd = nodes[node_index];
        std::vector<std::thread> spawned;
        for (int i = 0; i < 4; ++i) {
            if (i >= np) {
                nd.kind[i] = 0;
                nd.box[i].reset();
                nd.child[i] = 0;
                nd.count[i] = 0;
                continue;
            }
            nd.box[i] = parts[i].box;
            if (parts[i].count() <= (uint32_t)kLeafMax) {
                nd.kind[i] = 2;
                nd.child[i] = parts[i].begin;
                nd.count[i] = parts[i].count();
            } else {
                const uint32_t c = alloc();
                nd.kind[i] = 1;
                nd.child[i] = c;
                nd.count[i] = parts[i].count();
                const Range sub = parts[i];
                if (sub.count() >= kParallelMin && threads_left.fetch_sub(1) > 0) {
                    spawned.emplace_back([this, c, sub, depth]() { build(c, sub, depth + 1); });
                } else {
                    build(c, sub, depth + 1);
                }
            }
        }
        for (auto& t : spawned) t.join();
    }
};

void build_nodes(std::vector<Ref>& refs, const Box& scene, BuiltQbvh& out);

}  // namespace

int build_qbvh(const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt, const float* spheres,
               uint32_t ns, BuiltQbvh& out, std::string& err) {
    const auto t_start = std::chrono::steady_clock::now();
    const uint64_t n64 = (uint64_t)nt + ns;
    if (n64 >= (1ull << 29)) {
        err = "too many primitives (limit 2^29 - 1)";
        return -1;
    }
    const uint32_t n = (uint32_t)n64;
    out.nodes.clear();
    out.prims.clear();
    out.n_triangles = nt;
    out.n_spheres = ns;
    out.max_depth = 0;
    out.sah_cost = 0.0;
    for (int k = 0; k < 3; ++k) {
        out.scene_lo[k] = 0.0f;
        out.scene_hi[k] = 0.0f;
    }
    if (n == 0) {
        out.build_seconds = 0.0;
        return 0;
    }

    std::vector<Ref> refs(n);
    for (uint32_t i = 0; i < nt; ++i) {
        Ref& r = refs[i];
        r.box.reset();
        for (int k = 0; k < 3; ++k) {
            const uint32_t vi = tris[3 * i + k];
            if (vi >= nv) {
                err = "triangle " + std::to_string(i) + " references vertex " + std::to_string(vi) +
                      " >= n_vertices " + std::to_string(nv);
                return -1;
            }
            const float* p = verts + 3 * (size_t)vi;
            if (!std::isfinite(p[0]) || !std::isfinite(p[1]) || !std::isfinite(p[2])) {
                err = "non-finite vertex " + std::to_string(vi) + " in triangle " + std::to_string(i);
                return -1;
            }
            r.box.grow(p);
        }
        for (int k = 0; k < 3; ++k) r.c[k] = 0.5f * (r.box.lo[k] + r.box.hi[k]);
        r.prim = i;
    }
    for (uint32_t j = 0; j < ns; ++j) {
        Ref& r = refs[nt + j];
        const float* s = spheres + 4 * (size_t)j;
        if (!std::isfinite(s[0]) || !std::isfinite(s[1]) || !std::isfinite(s[2]) || !std::isfinite(s[3])) {
            err = "sphere " + std::to_string(j) + " has non-finite centre or radius";
            return -1;
        }
        if (!(s[3] >= 0.0f)) {
            err = "sphere " + std::to_string(j) + " has negative or NaN radius";
            return -1;
        }
        for (int k = 0; k < 3; ++k) {
            r.box.lo[k] = s[k] - s[3];
            r.box.hi[k] = s[k] + s[3];
            r.c[k] = s[k];
        }
        r.prim = nt + j;
    }
    Box scene;
    scene.reset();
    for (const Ref& r : refs) scene.grow(r.box);
    for (int k = 0; k < 3; ++k) {
        if (!std::isfinite(scene.lo[k]) || !std::isfinite(scene.hi[k])) {
            err = "non-finite primitive coordinates";
            return -1;
        }
        out.scene_lo[k] = scene.lo[k];
        out.scene_hi[k] = scene.hi[k];
    }

    build_nodes(refs, scene, out);

    out.prims.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t p = refs[i].prim;
        lr_prim& q = out.prims[i];
        std::memset(&q, 0, sizeof(q));
        if (p < nt) {
            const float* v0 = verts + 3 * (size_t)tris[3 * p + 0];
            const float* v1 = verts + 3 * (size_t)tris[3 * p + 1];
            const float* v2 = verts + 3 * (size_t)tris[3 * p + 2];
            for (int k = 0; k < 3; ++k) {
                q.a[k] = v0[k];
                q.b[k] = v1[k] - v0[k];
                q.c[k] = v2[k] - v0[k];
            }
            q.a[3] = as_f32(p);
            q.b[3] = as_f32(LR_PRIM_TRIANGLE);
        } else {
            const float* s = spheres + 4 * (size_t)(p - nt);
            q.a[0] = s[0];
            q.a[1] = s[1];
            q.a[2] = s[2];
            q.a[3] = as_f32(p);
            q.b[0] = s[3];
            q.b[1] = s[3] * s[3];
            q.b[3] = as_f32(LR_PRIM_SPHERE);
        }
    }
    out.build_seconds =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    return 0;
}

int build_boxes(const float* boxes, uint32_t n, BuiltQbvh& out, std::vector<uint32_t>& order, std::string& err) {
    const auto t_start = std::chrono::steady_clock::now();
    out = BuiltQbvh();
    order.clear();
    if (n >= (1u << 29)) {
        err = "too many boxes (limit 2^29 - 1)";
        return -1;
    }
    if (n == 0) return 0;
    std::vector<Ref> refs(n);
    Box scene;
    scene.reset();
    for (uint32_t i = 0; i < n; ++i) {
        Ref& r = refs[i];
        for (int k = 0; k < 3; ++k) {
            r.box.lo[k] = boxes[6 * (size_t)i + k];
            r.box.hi[k] = boxes[6 * (size_t)i + 3 + k];
            if (!std::isfinite(r.box.lo[k]) || !std::isfinite(r.box.hi[k]) || !(r.box.lo[k] <= r.box.hi[k])) {
                err = "box " + std::to_string(i) + " is not finite or not ordered";
                return -1;
            }
            r.c[k] = 0.5f * (r.box.lo[k] + r.box.hi[k]);
        }
        r.prim = i;
        scene.grow(r.box);
    }
    for (int k = 0; k < 3; ++k) {
        out.scene_lo[k] = scene.lo[k];
        out.scene_hi[k] = scene.hi[k];
    }
    build_nodes(refs, scene, out);
    order.resize(n);
    for (uint32_t i = 0; i < n; ++i) order[i] = refs[i].prim;
    out.build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    return 0;
}

// Per node and axis: origin = the smallest lo of the valid children, step
// s = the smallest power of two with fmaf(255, s, origin) >= the largest hi.
// A lo plane takes the largest q whose decode fmaf(q, s, origin) is <= the
// FP32 plane (q = 0 decodes to origin itself), a hi plane the smallest q whose
// decode is >= it (q = 255 always is).  q * s is exact, so std::fmaf here and
// the kernel's fmaf give the same floats.
void quantize_nodes(const lr_node* in, size_t n, lr_node_q8* out) {
    for (size_t k = 0; k < n; ++k) {
        const lr_node& a = in[k];
        lr_node_q8& o = out[k];
        std::memset(&o, 0, sizeof(o));
        const float* lo[3] = {a.lo_x, a.lo_y, a.lo_z};
        const float* hi[3] = {a.hi_x, a.hi_y, a.hi_z};
        uint8_t* qlo[3] = {o.qlo_x, o.qlo_y, o.qlo_z};
        uint8_t* qhi[3] = {o.qhi_x, o.qhi_y, o.qhi_z};
        for (int i = 0; i < 4; ++i) o.child[i] = a.child[i];
        for (int ax = 0; ax < 3; ++ax) {
            float org = INFINITY, top = -INFINITY;
            for (int i = 0; i < 4; ++i) {
                if (a.child[i] == LR_CHILD_EMPTY) continue;
                org = std::min(org, lo[ax][i]);
                top = std::max(top, hi[ax][i]);
            }
            if (!(org <= top)) {  // no child: any grid (the kernel skips empty children)
                org = 0.0f;
                top = 0.0f;
            }
            const double ext = (double)top - (double)org;
            int e = !(ext < 1e38) ? 127 : ext > 0.0 ? (int)std::ceil(std::log2(ext / 255.0)) : -126;
            e = std::max(-126, std::min(127, e));
            while (e < 127 && std::fmaf(255.0f, std::ldexp(1.0f, e), org) < top) ++e;
            const float sc = std::ldexp(1.0f, e);
            o.origin[ax] = org;
            o.scale[ax] = sc;
            auto dec = [&](int q) { return std::fmaf((float)q, sc, org); };
            for (int i = 0; i < 4; ++i) {
                if (a.child[i] == LR_CHILD_EMPTY) {
                    qlo[ax][i] = 255;
                    qhi[ax][i] = 0;
                    continue;
                }
                int q = (int)std::floor(((double)lo[ax][i] - org) / sc);
                q = std::max(0, std::min(255, q));
                while (q > 0 && dec(q) > lo[ax][i]) --q;
                while (q < 255 && dec(q + 1) <= lo[ax][i]) ++q;
                qlo[ax][i] = (uint8_t)q;
                q = (int)std::ceil(((double)hi[ax][i] - org) / sc);
                q = std::max(0, std::min(255, q));
                while (q < 255 && dec(q) < hi[ax][i]) ++q;
                while (q > 0 && dec(q - 1) >= hi[ax][i]) --q;
                qhi[ax][i] = (uint8_t)q;
            }
        }
    }
}

std::vector<lr_node_q8> quantize_nodes(const std::vector<lr_node>& in) {
    std::vector<lr_node_q8> out(in.size());
    if (!in.empty()) quantize_nodes(in.data(), in.size(), out.data());
    return out;
}

namespace {

void build_nodes(std::vector<Ref>& refs, const Box& scene, BuiltQbvh& out) {
    const uint32_t n = (uint32_t)refs.size();
    // Every inner node has >= 2 children and every leaf >= 1 prim, so there
    // are at most n - 1 inner nodes (1 when n == 1).
    const uint32_t capacity = std::max(n, 1u);
    const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
    Builder b(refs, capacity, hw);
    const uint32_t root = b.alloc();
    b.build(root, Range{0, n, scene}, 0);
    const uint32_t n_tmp = b.next.load();
    out.max_depth = b.max_depth.load();

    // Absolute pad: absorbs FP32 rounding of triangle hit distances near the
    // ray origin; relative pad: one part in 2^20 of each coordinate.
    float mag = 0.0f;
    for (int k = 0; k < 3; ++k) mag = std::max(mag, std::max(std::fabs(scene.lo[k]), std::fabs(scene.hi[k])));
    const float abs_pad = mag * (1.0f / 1048576.0f);
    auto pad_lo = [&](float x) { return x - std::fabs(x) * (1.0f / 1048576.0f) - abs_pad; };
    auto pad_hi = [&](float x) { return x + std::fabs(x) * (1.0f / 1048576.0f) + abs_pad; };

    // Renumber in depth-first preorder.
    out.nodes.resize(n_tmp);
    std::vector<std::pair<uint32_t, uint32_t>> work;  // (tmp index, final index)
    uint32_t next_final = 0;
    work.push_back({root, next_final++});
    const double root_area = std::max(scene.area(), 1e-30);
    double sah = 1.0;  // root node
    while (!work.empty()) {
        const auto w = work.back();
        work.pop_back();
        const TmpNode& t = b.nodes[w.first];
        lr_node& o = out.nodes[w.second];
        std::memset(&o, 0, sizeof(o));
        // Push inner children in reverse so child 0 is numbered first.
        uint32_t finals[4] = {0, 0, 0, 0};
        for (int i = 0; i < 4; ++i)
            if (t.kind[i] == 1) finals[i] = next_final++;
        for (int i = 3; i >= 0; --i)
            if (t.kind[i] == 1) work.push_back({t.child[i], finals[i]});
        for (int i = 0; i < 4; ++i) {
            if (t.kind[i] == 0) {
                o.lo_x[i] = o.lo_y[i] = o.lo_z[i] = 1.0f;
                o.hi_x[i] = o.hi_y[i] = o.hi_z[i] = -1.0f;
                o.child[i] = LR_CHILD_EMPTY;
                continue;
            }
            const Box& bx = t.box[i];
            o.lo_x[i] = pad_lo(bx.lo[0]);
            o.lo_y[i] = pad_lo(bx.lo[1]);
            o.lo_z[i] = pad_lo(bx.lo[2]);
            o.hi_x[i] = pad_hi(bx.hi[0]);
            o.hi_y[i] = pad_hi(bx.hi[1]);
            o.hi_z[i] = pad_hi(bx.hi[2]);
            const double rel = bx.area() / root_area;
            if (t.kind[i] == 1) {
                o.child[i] = finals[i];
                sah += rel;
            } else {
                o.child[i] = 0x80000000u | (t.child[i] << 2) | (t.count[i] - 1);
                sah += rel * t.count[i];
            }
        }
    }
    out.sah_cost = sah;
}

}  // namespace

}  // namespace lr
