// build.hip -- acceleration-structure builds on gfx950.
//
// Replaces the driver work behind
//   AccelerationStructure::build_blas   src/util_structs.rs:140-224
//   build_tlas                          src/util_functions.rs:453-510
//   AccelerationStructure::update_tlas  src/util_structs.rs:285-357 (refit)
//   AccelerationStructure::clone        src/util_structs.rs:359-405
// with a linear BVH (Karras 2012): primitive boxes and centres, 30-bit Morton
// codes, a stable LSD radix sort written here (LDS digit ranking with wave
// ballots), the radix-tree split search and a bottom-up box merge with
// atomic arrival counters, then a top-down collapse of the binary tree into
// 4-wide nodes, one breadth-first level per launch (node numbers from a
// prefix sum, so they do not depend on scheduling), and the slot boxes of
// the wide nodes from the deepest level up. The blobs are bit-identical to
// the host oracle's: every box operation is an exact min/max or a fixed FP32
// sequence.
#include <algorithm>
#include <cstring>
#include <vector>

#include "rtg_device.hpp"
#include "rtg_internal.hpp"

namespace rtg {

// min/max of box planes as selects: the first operand wins ties, so that
// +0/-0 resolve the same way on every device.
__device__ inline float mins(float a, float b) { return b < a ? b : a; }
__device__ inline float maxs(float a, float b) { return b > a ? b : a; }
__device__ inline Box box_union(const Box& a, const Box& b) {
    Box r;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        r.lo[k] = mins(a.lo[k], b.lo[k]);
        r.hi[k] = maxs(a.hi[k], b.hi[k]);
    }
    return r;
}
__device__ inline Box box_empty() {
    Box b;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        b.lo[k] = __builtin_inff();
        b.hi[k] = -__builtin_inff();
    }
    return b;
}

// Order-preserving map of float to uint (for atomic min/max of centres).
__device__ inline uint32_t f2ord(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float ord2f(uint32_t o) { return __uint_as_float((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o); }

__device__ inline float centre(const Box& b, int a) { return (b.lo[a] + b.hi[a]) * 0.5f + 0.0f; }

__device__ inline void set_slot(Node& nd, int s, const Box& b) {
    nd.lo_x[s] = b.lo[0];
    nd.hi_x[s] = b.hi[0];
    nd.lo_y[s] = b.lo[1];
    nd.hi_y[s] = b.hi[1];
    nd.lo_z[s] = b.lo[2];
    nd.hi_z[s] = b.hi[2];
}

__device__ inline Box slot_box(const Node& nd, int s) {
    Box b;
    b.lo[0] = nd.lo_x[s];
    b.hi[0] = nd.hi_x[s];
    b.lo[1] = nd.lo_y[s];
    b.hi[1] = nd.hi_y[s];
    b.lo[2] = nd.lo_z[s];
    b.hi[2] = nd.hi_z[s];
    return b;
}

// Box of a node: its four slot boxes folded in slot order.
__device__ inline Box node_box(const Node& nd) {
    Box b = box_empty();
#pragma unroll
    for (int s = 0; s < 4; ++s) b = box_union(b, slot_box(nd, s));
    return b;
}

// Slot box of child ref c: its leaf's primitive boxes (leaf_boxes[order[k]],
// or leaf_boxes[k] without an order) folded in leaf order, or its node's box.
__device__ inline Box ref_box(const Node* nodes, uint32_t c, const Box* leaf_boxes, const uint32_t* order) {
    Box b = box_empty();
    if (c == kEmpty) return b;
    if (c & kLeafBit) {
        uint32_t f = c & kFirstMask, cnt = ((c >> 27) & 0xfu) + 1u;
        for (uint32_t k = f; k < f + cnt; ++k) b = box_union(b, leaf_boxes[order ? order[k] : k]);
        return b;
    }
    return node_box(nodes[c]);
}

// Surface area of a box (half of it: the collapse only compares areas).
__device__ inline float box_area(const Box& b) {
    float ex = b.hi[0] - b.lo[0], ey = b.hi[1] - b.lo[1], ez = b.hi[2] - b.lo[2];
    return (ex * ey + ey * ez) + ez * ex;
}

// Block-wide reduction of centre bounds, then one atomic per block and axis.
template <int BS>
__device__ inline void reduce_centre_bounds(float c[3], bool valid[3], uint32_t* cb) {
    __shared__ uint32_t s_lo[3][BS / 64], s_hi[3][BS / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        uint32_t lo = valid[a] ? f2ord(c[a]) : 0xffffffffu;
        uint32_t hi = valid[a] ? f2ord(c[a]) : 0u;
        for (int off = 32; off > 0; off >>= 1) {
            lo = min(lo, (uint32_t)__shfl_xor((int)lo, off));
            hi = max(hi, (uint32_t)__shfl_xor((int)hi, off));
        }
        if (lane == 0) {
            s_lo[a][w] = lo;
            s_hi[a][w] = hi;
        }
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        int a = threadIdx.x;
        uint32_t lo = 0xffffffffu, hi = 0u;
        for (int k = 0; k < BS / 64; ++k) {
            lo = min(lo, s_lo[a][k]);
            hi = max(hi, s_hi[a][k]);
        }
        if (lo != 0xffffffffu) atomicMin(&cb[a], lo);
        if (hi != 0u) atomicMax(&cb[3 + a], hi);
    }
}

struct GeomDesc {
    const uint8_t* vertices;
    const uint32_t* indices;
    uint32_t stride;
    uint32_t max_vertex;
    uint32_t first_tri;
    uint32_t flags;
};

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_init_bounds(uint32_t* cb) {
    if (threadIdx.x < 3) {
        cb[threadIdx.x] = 0xffffffffu;
        cb[3 + threadIdx.x] = 0u;
    }
}

// Triangle records, primitive boxes, centre bounds.
__global__ __launch_bounds__(256) void k_tri_setup(const GeomDesc* geoms, uint32_t ngeom, uint32_t ntri, TriRec* tris,
                                                   Box* boxes, uint32_t* cb, uint32_t* status) {
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    float c[3] = {0, 0, 0};
    bool valid[3] = {false, false, false};
    if (i < ntri) {
        uint32_t lo = 0, hi = ngeom;  // last geometry with first_tri <= i
        while (hi - lo > 1) {
            uint32_t mid = (lo + hi) >> 1;
            if (geoms[mid].first_tri <= i) lo = mid;
            else hi = mid;
        }
        const GeomDesc g = geoms[lo];
        uint32_t t = i - g.first_tri;
        uint32_t i0 = g.indices[3ull * t], i1 = g.indices[3ull * t + 1], i2 = g.indices[3ull * t + 2];
        TriRec r;
        Box b;
        if (i0 > g.max_vertex || i1 > g.max_vertex || i2 > g.max_vertex) {
            atomicOr(status, (uint32_t)kStatusBadIndex);
            i0 = i1 = i2 = 0;
        }
        const float* a = reinterpret_cast<const float*>(g.vertices + (uint64_t)i0 * g.stride);
        const float* bb = reinterpret_cast<const float*>(g.vertices + (uint64_t)i1 * g.stride);
        const float* cc = reinterpret_cast<const float*>(g.vertices + (uint64_t)i2 * g.stride);
        float A[3] = {a[0], a[1], a[2]}, B[3] = {bb[0], bb[1], bb[2]}, Cv[3] = {cc[0], cc[1], cc[2]};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            r.v0[k] = A[k];
            r.v1[k] = B[k];
            r.v2[k] = Cv[k];
            b.lo[k] = mins(mins(A[k], B[k]), Cv[k]);
            b.hi[k] = maxs(maxs(A[k], B[k]), Cv[k]);
        }
        r.prim = t;
        r.geom = lo | ((g.flags & RTG_GEOMETRY_OPAQUE) ? 0u : kNonOpaqueBit);
        r.pad = 0;
        tris[i] = r;
        boxes[i] = b;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            c[k] = centre(b, k);
            valid[k] = isfinite(c[k]);
        }
    }
    reduce_centre_bounds<256>(c, valid, cb);
}

__device__ inline bool registry_has(const uint64_t* reg, uint32_t n, uint64_t a) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (reg[mid] < a) lo = mid + 1;
        else hi = mid;
    }
    return lo < n && reg[lo] == a;
}

// World box of an instance: the 8 corners of its BLAS root box through the
// object-to-world transform; empty for an empty or unknown BLAS.
__device__ inline Box instance_world_box(const rtg_instance& in, const AsHeader* bh) {
    Box r = box_empty();
    if (bh == nullptr || bh->prim_count == 0) return r;
    float m[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) m[k] = (&in.transform[0][0])[k];
    for (int c = 0; c < 8; ++c) {
        V3 p = v3((c & 1) ? bh->hi[0] : bh->lo[0], (c & 2) ? bh->hi[1] : bh->lo[1], (c & 4) ? bh->hi[2] : bh->lo[2]);
        V3 w = m34_point(m, p);
        r.lo[0] = mins(r.lo[0], w.x);
        r.hi[0] = maxs(r.hi[0], w.x);
        r.lo[1] = mins(r.lo[1], w.y);
        r.hi[1] = maxs(r.hi[1], w.y);
        r.lo[2] = mins(r.lo[2], w.z);
        r.hi[2] = maxs(r.hi[2], w.z);
    }
    return r;
}

__device__ inline void fill_inst_rec(InstRec& rec, const rtg_instance& in, uint32_t index, uint64_t blas) {
    float m[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) m[k] = (&in.transform[0][0])[k];
    m34_inverse(m, rec.w2o);
    rec.blas = blas;
    rec.index = index;
    rec.mask_flags = (in.custom_index_and_mask >> 24) | ((in.sbt_offset_and_flags >> 24) << 8);
}

__device__ inline const AsHeader* checked_blas(const rtg_instance& in, const uint64_t* reg, uint32_t nreg,
                                               uint32_t* status) {
    uint64_t a = in.acceleration_structure_reference;
    if (a == 0) return nullptr;
    if (!registry_has(reg, nreg, a)) {
        atomicOr(status, (uint32_t)kStatusBadBlas);
        return nullptr;
    }
    return reinterpret_cast<const AsHeader*>(a);
}

__global__ __launch_bounds__(256) void k_inst_setup(const rtg_instance* inst, uint32_t n, const uint64_t* reg,
                                                    uint32_t nreg, InstRec* recs, Box* boxes, uint32_t* cb,
                                                    uint32_t* status) {
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    float c[3] = {0, 0, 0};
    bool valid[3] = {false, false, false};
    if (i < n) {
        rtg_instance in = inst[i];
        const AsHeader* bh = checked_blas(in, reg, nreg, status);
        Box b = instance_world_box(in, bh);
        InstRec r;
        fill_inst_rec(r, in, i, bh ? in.acceleration_structure_reference : 0);
        recs[i] = r;
        boxes[i] = b;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            c[k] = centre(b, k);
            valid[k] = isfinite(c[k]);
        }
    }
    reduce_centre_bounds<256>(c, valid, cb);
}

__device__ inline uint32_t spread3(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
__device__ inline uint32_t quantize10(float c, float lo, float hi) {
    float ext = hi - lo;
    float f = ext > 0.0f ? (c - lo) / ext : 0.0f;
    float q = fminf(fmaxf(f * 1024.0f, 0.0f), 1023.0f);
    return (uint32_t)q;
}

__global__ __launch_bounds__(256) void k_morton(const Box* boxes, uint32_t n, const uint32_t* cb, uint32_t* keys,
                                                uint32_t* vals) {
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float lo[3], hi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        lo[a] = cb[a] == 0xffffffffu ? __builtin_inff() : ord2f(cb[a]);
        hi[a] = cb[3 + a] == 0u ? -__builtin_inff() : ord2f(cb[3 + a]);
    }
    Box b = boxes[i];
    uint32_t k = (spread3(quantize10(centre(b, 0), lo[0], hi[0])) << 2) |
                 (spread3(quantize10(centre(b, 1), lo[1], hi[1])) << 1) | spread3(quantize10(centre(b, 2), lo[2], hi[2]));
    keys[i] = k;
    vals[i] = i;
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort, 8-bit digits, tiles of 256 threads x 16 keys.
// ---------------------------------------------------------------------------
constexpr int kRsThreads = 256;
constexpr int kRsItems = 16;
constexpr int kRsTile = kRsThreads * kRsItems;

__global__ __launch_bounds__(kRsThreads) void k_rs_hist(const uint32_t* keys, uint32_t n, int shift, uint32_t* hist,
                                                        uint32_t ntiles) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    uint32_t base = blockIdx.x * kRsTile;
    for (int r = 0; r < kRsItems; ++r) {
        uint32_t i = base + r * kRsThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 0xffu], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of m entries in one block of 1024 threads.
__global__ __launch_bounds__(1024) void k_scan_exclusive(uint32_t* data, uint32_t m) {
    __shared__ uint32_t part[1024];
    uint32_t chunk = (m + 1023) / 1024;
    uint32_t b = threadIdx.x * chunk, e = min(m, b + chunk);
    uint32_t s = 0;
    for (uint32_t i = b; i < e; ++i) s += data[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        uint32_t v = threadIdx.x >= (uint32_t)off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - s;
    for (uint32_t i = b; i < e; ++i) {
        uint32_t v = data[i];
        data[i] = run;
        run += v;
    }
}

__global__ __launch_bounds__(kRsThreads) void k_rs_scatter(const uint32_t* kin, const uint32_t* vin, uint32_t* kout,
                                                           uint32_t* vout, uint32_t n, int shift,
                                                           const uint32_t* offsets, uint32_t ntiles) {
    constexpr int kWaves = kRsThreads / 64;
    __shared__ uint32_t base_off[256];
    __shared__ uint32_t running[256];
    __shared__ uint32_t wave_cnt[kWaves][256];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    base_off[threadIdx.x] = offsets[threadIdx.x * ntiles + blockIdx.x];
    running[threadIdx.x] = 0;
    for (int k = 0; k < kWaves; ++k) wave_cnt[k][threadIdx.x] = 0;
    __syncthreads();
    const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    uint32_t tile = blockIdx.x * kRsTile;
    for (int r = 0; r < kRsItems; ++r) {
        uint32_t i = tile + r * kRsThreads + threadIdx.x;
        bool valid = i < n;
        uint32_t key = valid ? kin[i] : 0u;
        uint32_t val = valid ? vin[i] : 0u;
        uint32_t digit = (key >> shift) & 0xffu;
        // lanes of this wave holding the same digit
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            uint64_t bal = __ballot((digit >> bit) & 1u);
            peers &= ((digit >> bit) & 1u) ? bal : ~bal;
        }
        uint32_t rank_in_wave = (uint32_t)__popcll(peers & lt_mask);
        bool leader = valid && (peers & lt_mask) == 0;
        if (leader) wave_cnt[w][digit] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t off = running[digit] + rank_in_wave;
            for (int k = 0; k < w; ++k) off += wave_cnt[k][digit];
            uint32_t dst = base_off[digit] + off;
            kout[dst] = key;
            vout[dst] = val;
        }
        __syncthreads();
        {
            uint32_t d = threadIdx.x;
            uint32_t s = 0;
            for (int k = 0; k < kWaves; ++k) {
                s += wave_cnt[k][d];
                wave_cnt[k][d] = 0;
            }
            running[d] += s;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Radix tree (Karras 2012, Figure 4)
// ---------------------------------------------------------------------------
__device__ inline int delta(const uint32_t* keys, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint32_t a = keys[i], b = keys[j];
    if (a == b) return 32 + __clz((int)((uint32_t)i ^ (uint32_t)j));
    return __clz((int)(a ^ b));
}

__global__ __launch_bounds__(256) void k_karras(const uint32_t* keys, int n, uint32_t* left, uint32_t* right,
                                                uint32_t* parent, uint32_t* rfirst, uint32_t* rcount) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n - 1) return;
    int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = delta(keys, n, i, j);
    int s = 0;
    int t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int gamma = i + s * d + (d < 0 ? d : 0);
    int lo = min(i, j), hi = max(i, j);
    uint32_t lc = (lo == gamma) ? (kLeafBit | (uint32_t)gamma) : (uint32_t)gamma;
    uint32_t rc = (hi == gamma + 1) ? (kLeafBit | (uint32_t)(gamma + 1)) : (uint32_t)(gamma + 1);
    left[i] = lc;
    right[i] = rc;
    rfirst[i] = (uint32_t)lo;
    rcount[i] = (uint32_t)(hi - lo + 1);
    // parent array: internal nodes at [0, n-1), leaves at [n-1, 2n-1)
    parent[(lc & kLeafBit) ? (n - 1) + (lc & kFirstMask) : lc] = (uint32_t)i;
    parent[(rc & kLeafBit) ? (n - 1) + (rc & kFirstMask) : rc] = (uint32_t)i;
    if (i == 0) parent[0] = 0xffffffffu;
}

// Depth of every internal node of the binary tree (the root has depth 0),
// by walking its parent chain; the largest depth is folded into *max_depth.
__global__ __launch_bounds__(256) void k_node_depth(int n, const uint32_t* parent, uint32_t* depth,
                                                    uint32_t* max_depth) {
    int i = blockIdx.x * 256 + threadIdx.x;
    uint32_t d = 0;
    if (i < n - 1) {
        for (uint32_t p = parent[i]; p != 0xffffffffu; p = parent[p]) ++d;
        depth[i] = d;
    }
    for (int off = 32; off > 0; off >>= 1) d = max(d, (uint32_t)__shfl_xor((int)d, off));
    if ((threadIdx.x & 63) == 0 && d) atomicMax(max_depth, d);
}

// Boxes and surface areas of the internal nodes at depth `d` (the deeper
// ones are done).
__global__ __launch_bounds__(256) void k_boxes_level(uint32_t d, const uint32_t* depth, const Box* boxes,
                                                     const uint32_t* order, int n, const uint32_t* left,
                                                     const uint32_t* right, Box* ibox, float* area) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n - 1 || depth[i] != d) return;
    uint32_t c[2] = {left[i], right[i]};
    Box cbx[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) cbx[s] = (c[s] & kLeafBit) ? boxes[order[c[s] & kFirstMask]] : ibox[c[s]];
    Box u = box_union(cbx[0], cbx[1]);
    ibox[i] = u;
    area[i] = box_area(u);
}

// ---------------------------------------------------------------------------
// Treelet restructuring (Karras & Aila 2013, HPG, "Fast Parallel
// Construction of High-Quality Bounding Volume Hierarchies", section 4) of
// the binary tree, for single-primitive leaves. The treelet of internal node
// r: r's two children, then, while fewer than kTreeletLeaves, the treelet
// leaf with the largest area that is an internal node replaced by its left
// child in place and its right child appended. Its topology becomes the one
// that minimises the summed surface area of its internal nodes (SAH cost with
// fixed leaves), by dynamic programming over subsets of treelet leaves:
// cost(S) = area(S) + min over partitions (P, S\P), P holding S's lowest
// leaf, of cost(P) + cost(S\P); partitions tried in decreasing order of P,
// the first minimum kept. New internal nodes reuse the treelet's numbers:
// r for the whole set, then the expanded nodes in expansion order, handed
// out as subsets are split from a stack (P's subtree first). A node is
// processed after its whole subtree (arrival counters), so the result does
// not depend on thread timing.
// ---------------------------------------------------------------------------
constexpr int kTreeletLeaves = 7;
constexpr int kTreeletRounds = 5;  // passes with RTG_BUILD_PREFER_FAST_TRACE

__device__ inline Box bin_box(const Box* boxes, const uint32_t* order, const Box* ibox, uint32_t c) {
    return (c & kLeafBit) ? boxes[order[c & kFirstMask]] : ibox[c];
}

// The cheapest split of leaf set S: the partition (P, S \ P) with P holding
// S's lowest leaf that minimises cost(P) + cost(S \ P), partitions tried in
// decreasing order of P and the first minimum kept. P = low | Q for Q a
// proper submask of S \ low (2^m - 1 of them for m bits), and decreasing Q is
// decreasing P. Four partitions at a time, so that their LDS reads are in
// flight together; the comparisons stay in order.
__device__ inline uint32_t best_split(uint32_t S, const float* cost, int stride, float& best) {
    const uint32_t low = S & (0u - S), R = S ^ low;
    const uint32_t count = (1u << __builtin_popcount(R)) - 1u;
    uint32_t Q = (R - 1u) & R;
    uint32_t bp = 0;
    best = 0.0f;
    uint32_t i = 0;
    for (; i + 4u <= count; i += 4u) {
        uint32_t P[4];
        float c[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            P[j] = Q | low;
            c[j] = cost[P[j] * stride] + cost[(S ^ P[j]) * stride];
            Q = (Q - 1u) & R;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (bp == 0 || c[j] < best) {
                best = c[j];
                bp = P[j];
            }
    }
    for (; i < count; ++i) {
        const uint32_t P = Q | low;
        const float c = cost[P * stride] + cost[(S ^ P) * stride];
        if (bp == 0 || c < best) {
            best = c;
            bp = P;
        }
        Q = (Q - 1u) & R;
    }
    return bp;
}

// `cost` is this lane's column of a per-workgroup LDS table (entry S at
// cost[S * stride]); LDS keeps the dynamic programme's 2^7 subset costs
// close, where a per-lane array would live in scratch memory.
__device__ void treelet_optimize(uint32_t r, const Box* boxes, const uint32_t* order, uint32_t* left,
                                 uint32_t* right, uint32_t* rcount, Box* ibox, float* area, float* cost,
                                 int stride) {
    uint32_t leaf[kTreeletLeaves], inner[kTreeletLeaves];
    int nl = 2, ni = 0;
    leaf[0] = left[r];
    leaf[1] = right[r];
    while (nl < kTreeletLeaves) {
        int best = -1;
        float best_area = 0.0f;
        for (int k = 0; k < nl; ++k) {
            const uint32_t c = leaf[k];
            if (c & kLeafBit) continue;
            const float a = area[c];
            if (best < 0 || a > best_area) {
                best = k;
                best_area = a;
            }
        }
        if (best < 0) break;
        const uint32_t c = leaf[best];
        inner[ni++] = c;
        leaf[best] = left[c];
        leaf[nl++] = right[c];
    }
    if (nl < 3) return;
    // leaf boxes in registers: the loops over them are unrolled to constant
    // indices (set bits of S beyond nl never occur)
    Box lb[kTreeletLeaves];
#pragma unroll
    for (int k = 0; k < kTreeletLeaves; ++k) lb[k] = k < nl ? bin_box(boxes, order, ibox, leaf[k]) : box_empty();
    const uint32_t full = (1u << nl) - 1u;
    for (uint32_t Sl = 1; Sl <= full; ++Sl) {
        // lanes go through the sets in step (lanes past their `full` have
        // left the loop), so S is uniform: its table reads are scalar
        const uint32_t S = (uint32_t)__builtin_amdgcn_readfirstlane((int)Sl);
        if ((S & (S - 1u)) == 0) {
            cost[S * stride] = 0.0f;
            continue;
        }
        Box b = box_empty();
#pragma unroll
        for (int k = 0; k < kTreeletLeaves; ++k)
            if (S & (1u << k)) b = box_union(b, lb[k]);
        float best;
        best_split(S, cost, stride, best);
        cost[S * stride] = box_area(b) + best;
    }
    uint32_t stk_set[kTreeletLeaves], stk_id[kTreeletLeaves], done[kTreeletLeaves];
    int sp = 0, no = 0, next = 0;
    stk_set[sp] = full;
    stk_id[sp++] = r;
    while (sp > 0) {
        --sp;
        const uint32_t S = stk_set[sp], id = stk_id[sp];
        done[no++] = id;
        float unused;
        const uint32_t P = best_split(S, cost, stride, unused);
        const uint32_t part[2] = {P, S ^ P};
        uint32_t ref[2];
        for (int h = 0; h < 2; ++h)
            ref[h] = (part[h] & (part[h] - 1u)) ? inner[next++] : leaf[__builtin_ctz(part[h])];
        left[id] = ref[0];
        right[id] = ref[1];
        for (int h = 1; h >= 0; --h)
            if (part[h] & (part[h] - 1u)) {
                stk_set[sp] = part[h];
                stk_id[sp++] = ref[h];
            }
    }
    for (int k = no - 1; k >= 0; --k) {
        const uint32_t id = done[k], c0 = left[id], c1 = right[id];
        const Box u = box_union(bin_box(boxes, order, ibox, c0), bin_box(boxes, order, ibox, c1));
        ibox[id] = u;
        area[id] = box_area(u);
        rcount[id] = ((c0 & kLeafBit) ? 1u : rcount[c0]) + ((c1 & kLeafBit) ? 1u : rcount[c1]);
    }
}

// One depth level of a pass: every internal node at depth `d` of the tree
// the pass started from. Launched from the deepest level up, so a node is
// restructured after its whole subtree; nodes of one level have disjoint
// subtrees, and a restructured subtree keeps its node set, so the result is
// the one of any order that visits subtrees first.
constexpr int kTreeletThreads = 64;

// Restructures the `count` internal nodes of one depth level listed in
// `nodes` (the level's run of the nodes sorted by depth), so that every
// workgroup is full of treelets.
__global__ __launch_bounds__(kTreeletThreads) void k_treelet_level(const uint32_t* nodes, uint32_t count,
                                                                   const Box* boxes, const uint32_t* order,
                                                                   uint32_t* left, uint32_t* right, uint32_t* rcount,
                                                                   Box* ibox, float* area) {
    __shared__ float s_cost[(1 << kTreeletLeaves) * kTreeletThreads];
    uint32_t k = blockIdx.x * kTreeletThreads + threadIdx.x;
    if (k >= count) return;
    treelet_optimize(nodes[k], boxes, order, left, right, rcount, ibox, area, s_cost + threadIdx.x,
                     kTreeletThreads);
}

__global__ __launch_bounds__(256) void k_iota(uint32_t* p, uint32_t n) {
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = i;
}

// First position of each depth in the depth-sorted node list: where the
// depth differs from its predecessor's. Every depth from 0 to the tree's
// depth holds a node (the ancestors of a deepest one); the entry past the
// deepest is the list length, filled beforehand.
__global__ __launch_bounds__(256) void k_level_starts(const uint32_t* sorted_depth, uint32_t m, uint32_t* start) {
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    const uint32_t d = sorted_depth[i];
    if (i == 0 || sorted_depth[i - 1] != d) start[d] = i;
}

// Parents of the restructured tree (the root keeps none).
__global__ __launch_bounds__(256) void k_parents(int n, const uint32_t* left, const uint32_t* right,
                                                 uint32_t* parent) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n - 1) return;
    const uint32_t lc = left[i], rc = right[i];
    parent[(lc & kLeafBit) ? (n - 1) + (lc & kFirstMask) : lc] = (uint32_t)i;
    parent[(rc & kLeafBit) ? (n - 1) + (rc & kFirstMask) : rc] = (uint32_t)i;
}

// The single node of a structure with at most leaf_max primitives: one leaf
// of all of them in input order (or none).
__global__ void k_wide_single(const Box* boxes, uint32_t n, Node* nodes, uint32_t* order) {
    if (threadIdx.x == 0) {
        Node nd;
        const uint32_t ref = n ? (kLeafBit | ((n - 1) << 27)) : kEmpty;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            nd.child[s] = s == 0 ? ref : kEmpty;
            nd.pad[s] = 0;
            set_slot(nd, s, s == 0 ? ref_box(nullptr, ref, boxes, nullptr) : box_empty());
        }
        nodes[0] = nd;
    }
    for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) order[k] = k;
}

// Slots (binary refs) of each wide node of a level: the binary root's two
// children, then the largest-area openable slot replaced in place by its
// two children while fewer than four; and the number of slots that become
// wide nodes of the next level.
__global__ __launch_bounds__(256) void k_wide_expand(uint32_t lb, uint32_t m, const uint32_t* wroot,
                                                     const uint32_t* left, const uint32_t* right,
                                                     const uint32_t* rcount, const float* area, uint32_t leaf_max,
                                                     uint4* slots, uint32_t* cnt) {
    uint32_t w = blockIdx.x * 256 + threadIdx.x;
    if (w >= m) return;
    uint32_t r = wroot[lb + w];
    uint32_t s[4] = {left[r], right[r], kEmpty, kEmpty};
    int nslot = 2;
    while (nslot < 4) {
        int best = -1;
        float best_area = 0.0f;
        for (int k = 0; k < nslot; ++k) {
            uint32_t c = s[k];
            if ((c & kLeafBit) || rcount[c] <= leaf_max) continue;
            float a = area[c];
            if (best < 0 || a > best_area) {
                best = k;
                best_area = a;
            }
        }
        if (best < 0) break;
        uint32_t c = s[best];
        for (int k = nslot; k > best + 1; --k) s[k] = s[k - 1];
        s[best] = left[c];
        s[best + 1] = right[c];
        ++nslot;
    }
    uint32_t internal = 0;
    for (int k = 0; k < nslot; ++k) internal += (!(s[k] & kLeafBit) && rcount[s[k]] > leaf_max) ? 1u : 0u;
    slots[w] = make_uint4(s[0], s[1], s[2], s[3]);
    cnt[w] = internal;
}

// Child refs of each wide node of a level; opened slots get the next
// level's node numbers from the exclusive prefix sum `base` (node le + base).
__global__ __launch_bounds__(256) void k_wide_emit(uint32_t lb, uint32_t le, uint32_t m, const uint4* slots,
                                                   const uint32_t* base, const uint32_t* rfirst,
                                                   const uint32_t* rcount, uint32_t leaf_max, uint32_t cap,
                                                   Node* nodes, uint32_t* wroot, uint32_t* next_count) {
    uint32_t w = blockIdx.x * 256 + threadIdx.x;
    if (w >= m) return;
    uint4 sv = slots[w];
    uint32_t s[4] = {sv.x, sv.y, sv.z, sv.w};
    uint32_t next = le + base[w];
    uint32_t internal = 0;
    Node& nd = nodes[lb + w];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        uint32_t c = s[k], ref = kEmpty;
        if (c != kEmpty) {
            if (c & kLeafBit) {
                ref = c;
            } else if (rcount[c] <= leaf_max) {
                ref = kLeafBit | ((rcount[c] - 1u) << 27) | rfirst[c];
            } else {
                // a wide tree over n leaves has at most n - 1 nodes (each has
                // two or more slots); the host checks the count against cap
                ref = next;
                if (next < cap) wroot[next] = c;
                ++next;
                ++internal;
            }
        }
        nd.child[k] = ref;
    }
    if (w == m - 1) *next_count = base[w] + internal;
}

// Slot boxes of the wide nodes of a level (the levels below are done).
__global__ __launch_bounds__(256) void k_wide_boxes(uint32_t lb, uint32_t m, Node* nodes, const Box* boxes,
                                                    const uint32_t* order) {
    uint32_t w = blockIdx.x * 256 + threadIdx.x;
    if (w >= m) return;
    Node& nd = nodes[lb + w];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        set_slot(nd, k, ref_box(nodes, nd.child[k], boxes, order));
        nd.pad[k] = 0;
    }
}

__global__ void k_header(AsHeader* h, uint32_t kind, uint32_t node_count, uint32_t prim_count, uint32_t geom_count,
                         uint32_t levels, uint64_t nodes_off, uint64_t prims_off, uint64_t extra_off,
                         uint64_t level_off, uint64_t leafbox_off, uint64_t total) {
    if (threadIdx.x != 0) return;
    AsHeader hd;
    memset(&hd, 0, sizeof(hd));
    hd.magic = kMagic;
    hd.kind = kind;
    hd.node_count = node_count;
    hd.prim_count = prim_count;
    hd.geom_count = geom_count;
    hd.max_depth = levels;
    hd.node_width = 4;
    hd.flags = 0;
    const Node* nodes = reinterpret_cast<const Node*>(reinterpret_cast<const uint8_t*>(h) + nodes_off);
    Box r = node_box(nodes[0]);
    for (int a = 0; a < 3; ++a) {
        hd.lo[a] = r.lo[a];
        hd.hi[a] = r.hi[a];
    }
    hd.nodes_off = nodes_off;
    hd.prims_off = prims_off;
    hd.extra_off = extra_off;
    hd.level_off = level_off;
    hd.total_size = total;
    hd.leafbox_off = leafbox_off;
    *h = hd;
}

__global__ __launch_bounds__(256) void k_gather_tris(const uint32_t* order, const TriRec* in, TriRec* out, uint32_t n) {
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = in[order[i]];
}

__global__ __launch_bounds__(256) void k_gather_insts(const uint32_t* order, const InstRec* in, const Box* boxes,
                                                      InstRec* out, Box* leafbox, uint32_t n) {
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        uint32_t s = order[i];
        out[i] = in[s];
        leafbox[i] = boxes[s];
    }
}

// In-place TLAS UPDATE (src/util_structs.rs:285-357): new instance records,
// same topology. One workgroup walks the level lists from the deepest level
// to the root with a barrier between levels. 256 lanes, one wave per SIMD:
// while another frame's traversal waves fill the GPU, such a workgroup finds
// room on a CU long before a 1024-lane one (four waves per SIMD) would.
__global__ __launch_bounds__(256) void k_tlas_refit(uint8_t* base, const rtg_instance* inst, uint32_t n,
                                                    const uint64_t* reg, uint32_t nreg, uint32_t* status) {
    AsHeader* h = reinterpret_cast<AsHeader*>(base);
    Node* nodes = reinterpret_cast<Node*>(base + h->nodes_off);
    InstRec* recs = reinterpret_cast<InstRec*>(base + h->prims_off);
    rtg_instance* src = reinterpret_cast<rtg_instance*>(base + h->extra_off);
    Box* leafbox = reinterpret_cast<Box*>(base + h->leafbox_off);
    for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) {
        uint32_t s = recs[k].index;
        rtg_instance in = inst[s];
        const AsHeader* bh = checked_blas(in, reg, nreg, status);
        InstRec r;
        fill_inst_rec(r, in, s, bh ? in.acceleration_structure_reference : 0);
        recs[k] = r;
        leafbox[k] = instance_world_box(in, bh);
        src[s] = in;
    }
    __syncthreads();
    // same topology: slot boxes again, deepest level first
    const uint32_t* lv = reinterpret_cast<const uint32_t*>(base + h->level_off);
    const uint32_t levels = lv[0];
    const uint32_t* offs = lv + 1;
    for (uint32_t L = levels; L-- > 0;) {
        for (uint32_t q = offs[L] + threadIdx.x; q < offs[L + 1]; q += blockDim.x) {
            Node& nd = nodes[q];
#pragma unroll
            for (int s = 0; s < 4; ++s) set_slot(nd, s, ref_box(nodes, nd.child[s], leafbox, nullptr));
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        Box r = node_box(nodes[0]);
        for (int a = 0; a < 3; ++a) {
            h->lo[a] = r.lo[a];
            h->hi[a] = r.hi[a];
        }
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static inline uint64_t align64(uint64_t x) { return (x + 63) & ~uint64_t(63); }
static inline unsigned blocks_for(uint64_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

struct Layout {
    uint64_t nodes_off, prims_off, extra_off, level_off, leafbox_off, total;
    uint32_t node_count;
};

// Blob layouts for a tree of node_count nodes in `levels` levels.
static Layout blas_layout(uint32_t n, uint32_t ngeom, uint32_t node_count) {
    Layout L;
    L.node_count = node_count;
    L.nodes_off = 128;
    L.prims_off = align64(L.nodes_off + (uint64_t)node_count * sizeof(Node));
    L.extra_off = align64(L.prims_off + (uint64_t)n * 48);
    L.level_off = 0;
    L.leafbox_off = 0;
    L.total = align64(L.extra_off + (uint64_t)ngeom * 4);
    return L;
}

static Layout tlas_layout(uint32_t n, uint32_t node_count, uint32_t levels) {
    Layout L;
    L.node_count = node_count;
    L.nodes_off = 128;
    L.prims_off = align64(L.nodes_off + (uint64_t)node_count * sizeof(Node));
    L.extra_off = align64(L.prims_off + (uint64_t)n * 64);
    L.level_off = align64(L.extra_off + (uint64_t)n * 64);
    L.leafbox_off = align64(L.level_off + 4ull * (2 + levels));
    L.total = align64(L.leafbox_off + (uint64_t)n * 24);
    return L;
}

// Node capacity of a wide tree over n primitives (its node count is at most
// n - 1: every node has two or more slots).
static inline uint32_t node_cap(uint32_t n) { return n >= 2 ? n - 1 : 1; }

// Scratch needs of the wide-tree build over n primitives.
static size_t lbvh_scratch_bytes(uint32_t n, size_t rec_bytes) {
    size_t off = 0;
    uint32_t ntiles = (n + kRsTile - 1) / kRsTile;
    uint32_t cap = node_cap(n);
    carve_size<uint8_t>(off, (size_t)n * rec_bytes);  // unsorted records
    carve_size<Box>(off, n);                          // primitive boxes
    carve_size<uint32_t>(off, 8);                     // centre bounds
    carve_size<uint32_t>(off, n);                     // keys a
    carve_size<uint32_t>(off, n);                     // keys b
    carve_size<uint32_t>(off, n);                     // vals a
    carve_size<uint32_t>(off, n);                     // vals b
    carve_size<uint32_t>(off, (size_t)256 * std::max(1u, ntiles));
    carve_size<uint32_t>(off, n);                     // left
    carve_size<uint32_t>(off, n);                     // right
    carve_size<uint32_t>(off, 2ull * n);              // parent
    carve_size<uint32_t>(off, n);                     // arrivals (node depths)
    carve_size<uint32_t>(off, n);                     // range first
    carve_size<uint32_t>(off, n);                     // range count
    carve_size<float>(off, n);                        // binary box areas
    carve_size<Box>(off, n);                          // binary boxes
    carve_size<Node>(off, cap);                       // wide nodes
    carve_size<uint32_t>(off, cap);                   // binary root of each wide node
    carve_size<uint4>(off, cap);                      // slots of a level
    carve_size<uint32_t>(off, cap);                   // opened slots per node / prefix sums
    carve_size<uint32_t>(off, 4 + 257);               // next level size, tree depth, level starts
    carve_size<GeomDesc>(off, 1);
    return off + 4096;
}

struct LbvhBufs {
    uint8_t* recs;
    Box* boxes;
    uint32_t* cb;
    uint32_t *ka, *kb, *va, *vb, *hist, *left, *right, *parent, *arrivals, *rfirst, *rcount;
    float* area;
    Box* ibox;
    Node* nodes;
    uint32_t *wroot, *cnt, *next;
    uint4* slots;
    GeomDesc* geoms;
};

static LbvhBufs carve_lbvh(void* base, uint32_t n, size_t rec_bytes, uint32_t ngeom = 1) {
    Carve c(base);
    LbvhBufs b;
    b.geoms = c.take<GeomDesc>(std::max(1u, ngeom));
    uint32_t ntiles = (n + kRsTile - 1) / kRsTile;
    uint32_t cap = node_cap(n);
    b.recs = c.take<uint8_t>((size_t)n * rec_bytes);
    b.boxes = c.take<Box>(n);
    b.cb = c.take<uint32_t>(8);
    b.ka = c.take<uint32_t>(n);
    b.kb = c.take<uint32_t>(n);
    b.va = c.take<uint32_t>(n);
    b.vb = c.take<uint32_t>(n);
    b.hist = c.take<uint32_t>((size_t)256 * std::max(1u, ntiles));
    b.left = c.take<uint32_t>(n);
    b.right = c.take<uint32_t>(n);
    b.parent = c.take<uint32_t>(2ull * n);
    b.arrivals = c.take<uint32_t>(n);
    b.rfirst = c.take<uint32_t>(n);
    b.rcount = c.take<uint32_t>(n);
    b.area = c.take<float>(n);
    b.ibox = c.take<Box>(n);
    b.nodes = c.take<Node>(cap);
    b.wroot = c.take<uint32_t>(cap);
    b.slots = c.take<uint4>(cap);
    b.cnt = c.take<uint32_t>(cap);
    b.next = c.take<uint32_t>(4 + 257);
    return b;
}

// Sorts the primitives, builds the binary radix tree and collapses it into
// the wide tree in b.nodes. On return the leaf order is in b.va and
// level_off holds the level offsets (level_off.back() = node count). Reads
// each level's size back from the device (synchronises `s`).
static int wide_hierarchy(LbvhBufs& b, uint32_t n, uint32_t leaf_max, int treelet_rounds, hipStream_t s,
                          std::vector<uint32_t>& level_off) {
    level_off.assign(1, 0u);
    if (n <= leaf_max || n <= 1) {
        hipLaunchKernelGGL(k_wide_single, dim3(1), dim3(256), 0, s, b.boxes, n, b.nodes, b.va);
        RTG_HIP(hipGetLastError());
        level_off.push_back(1u);
        return RTG_OK;
    }
    hipLaunchKernelGGL(k_morton, dim3(blocks_for(n, 256)), dim3(256), 0, s, b.boxes, n, b.cb, b.ka, b.va);
    uint32_t ntiles = (n + kRsTile - 1) / kRsTile;
    uint32_t *kin = b.ka, *kout = b.kb, *vin = b.va, *vout = b.vb;
    for (int pass = 0; pass < 4; ++pass) {
        int shift = pass * 8;
        hipLaunchKernelGGL(k_rs_hist, dim3(ntiles), dim3(kRsThreads), 0, s, kin, n, shift, b.hist, ntiles);
        hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, s, b.hist, 256u * ntiles);
        hipLaunchKernelGGL(k_rs_scatter, dim3(ntiles), dim3(kRsThreads), 0, s, kin, vin, kout, vout, n, shift, b.hist,
                           ntiles);
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    // after 4 passes the sorted data is back in (ka, va)
    hipLaunchKernelGGL(k_karras, dim3(blocks_for(n - 1, 256)), dim3(256), 0, s, b.ka, (int)n, b.left, b.right,
                       b.parent, b.rfirst, b.rcount);
    // Bottom-up passes go one depth level per launch, deepest first: the
    // kernel boundaries order the levels, where arrival counters would need
    // device-scope atomics and fences, which MI355X serves past its per-XCD
    // L2s. `b.arrivals` holds the depths; the host reads the tree's depth.
    uint32_t* depth = b.arrivals;
    uint32_t* max_depth = b.next + 1;
    auto tree_depth = [&](uint32_t& out) -> int {
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, max_depth, 1ull, 0u);
        hipLaunchKernelGGL(k_node_depth, dim3(blocks_for(n - 1, 256)), dim3(256), 0, s, (int)n, b.parent, depth,
                           max_depth);
        RTG_HIP(hipGetLastError());
        RTG_HIP(hipMemcpyAsync(&out, max_depth, 4, hipMemcpyDeviceToHost, s));
        RTG_HIP(hipStreamSynchronize(s));
        RTG_REQUIRE(out < n, RTG_ERROR_DEVICE, "BVH build: tree depth %u over %u primitives", out, n);
        return RTG_OK;
    };
    uint32_t dmax = 0;
    int st = tree_depth(dmax);
    if (st != RTG_OK) return st;
    for (uint32_t d = dmax + 1; d-- > 0;)
        hipLaunchKernelGGL(k_boxes_level, dim3(blocks_for(n - 1, 256)), dim3(256), 0, s, d, depth, b.boxes, b.va,
                           (int)n, b.left, b.right, b.ibox, b.area);
    // the treelet passes need single-primitive leaves: a restructured
    // subtree no longer covers a contiguous range of sorted primitives.
    // Each pass sorts the internal nodes by depth (one 8-bit radix pass:
    // depth values in b.arrivals, node numbers from b.ka, which the Morton
    // codes no longer need, into b.kb / b.vb) and restructures each depth's
    // run with full workgroups, deepest first.
    const uint32_t m = n - 1;
    const uint32_t mtiles = (m + kRsTile - 1) / kRsTile;
    std::vector<uint32_t> start;
    for (int round = 0; round < treelet_rounds && leaf_max == 1; ++round) {
        if (round > 0) {
            st = tree_depth(dmax);
            if (st != RTG_OK) return st;
        }
        RTG_REQUIRE(dmax < 256u, RTG_ERROR_LIMIT, "BVH build: binary tree depth %u beyond 255", dmax);
        hipLaunchKernelGGL(k_iota, dim3(blocks_for(m, 256)), dim3(256), 0, s, b.ka, m);
        hipLaunchKernelGGL(k_rs_hist, dim3(mtiles), dim3(kRsThreads), 0, s, depth, m, 0, b.hist, mtiles);
        hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, s, b.hist, 256u * mtiles);
        hipLaunchKernelGGL(k_rs_scatter, dim3(mtiles), dim3(kRsThreads), 0, s, depth, b.ka, b.kb, b.vb, m, 0, b.hist,
                           mtiles);
        uint32_t* d_start = b.next + 4;  // 257 words after the collapse's counters
        hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(dmax + 2, 256)), dim3(256), 0, s, d_start,
                           (uint64_t)(dmax + 2), m);
        hipLaunchKernelGGL(k_level_starts, dim3(blocks_for(m, 256)), dim3(256), 0, s, b.kb, m, d_start);
        RTG_HIP(hipGetLastError());
        start.assign(dmax + 2, m);
        RTG_HIP(hipMemcpyAsync(start.data(), d_start, (dmax + 2) * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        RTG_HIP(hipStreamSynchronize(s));
        for (uint32_t d = dmax + 1; d-- > 0;) {
            const uint32_t count = start[d + 1] - start[d];
            if (count == 0) continue;
            hipLaunchKernelGGL(k_treelet_level, dim3(blocks_for(count, kTreeletThreads)), dim3(kTreeletThreads), 0, s,
                               b.vb + start[d], count, b.boxes, b.va, b.left, b.right, b.rcount, b.ibox, b.area);
        }
        hipLaunchKernelGGL(k_parents, dim3(blocks_for(n - 1, 256)), dim3(256), 0, s, (int)n, b.left, b.right,
                           b.parent);
    }
    // top-down collapse, one breadth-first level per step
    const uint32_t cap = node_cap(n);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, b.wroot, 1ull, 0u);
    uint32_t lb = 0, le = 1;
    while (lb < le) {
        const uint32_t m = le - lb;
        hipLaunchKernelGGL(k_wide_expand, dim3(blocks_for(m, 256)), dim3(256), 0, s, lb, m, b.wroot, b.left, b.right,
                           b.rcount, b.area, leaf_max, b.slots, b.cnt);
        hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, s, b.cnt, m);
        hipLaunchKernelGGL(k_wide_emit, dim3(blocks_for(m, 256)), dim3(256), 0, s, lb, le, m, b.slots, b.cnt,
                           b.rfirst, b.rcount, leaf_max, cap, b.nodes, b.wroot, b.next);
        RTG_HIP(hipGetLastError());
        uint32_t next = 0;
        RTG_HIP(hipMemcpyAsync(&next, b.next, 4, hipMemcpyDeviceToHost, s));
        RTG_HIP(hipStreamSynchronize(s));
        RTG_REQUIRE(next <= cap - le, RTG_ERROR_DEVICE, "BVH collapse: %u nodes exceed the capacity %u", le + next,
                    cap);
        level_off.push_back(le);
        lb = le;
        le += next;
    }
    const uint32_t levels = (uint32_t)level_off.size() - 1;
    for (uint32_t L = levels; L-- > 0;) {
        const uint32_t m = level_off[L + 1] - level_off[L];
        hipLaunchKernelGGL(k_wide_boxes, dim3(blocks_for(m, 256)), dim3(256), 0, s, level_off[L], m, b.nodes,
                           b.boxes, b.va);
    }
    RTG_HIP(hipGetLastError());
    return RTG_OK;
}

int refresh_registry(rtg_ctx* ctx) {
    std::vector<uint64_t> addrs;
    for (auto& kv : ctx->structures)
        if (kv.second.kind == kKindBlas) addrs.push_back(kv.first);
    std::sort(addrs.begin(), addrs.end());
    uint64_t* d = nullptr;
    RTG_HIP(hipMalloc(&d, std::max<size_t>(1, addrs.size()) * sizeof(uint64_t)));
    if (!addrs.empty()) RTG_HIP(hipMemcpy(d, addrs.data(), addrs.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
    if (ctx->d_registry) ctx->retired.push_back(ctx->d_registry);
    ctx->d_registry = d;
    ctx->registry_count = (uint32_t)addrs.size();
    return RTG_OK;
}

int launch_refit(rtg_ctx* ctx, uint8_t* tlas, const rtg_instance* instances, uint32_t count, hipStream_t s) {
    LaunchTimer timer(ctx, RTG_KERNEL_REFIT, s);
    hipLaunchKernelGGL(k_tlas_refit, dim3(1), dim3(256), 0, s, tlas, instances, count, ctx->d_registry,
                       ctx->registry_count, ctx->d_status);
    RTG_HIP(hipGetLastError());
    return RTG_OK;
}

int check_status(rtg_ctx* ctx, hipStream_t stream, uint32_t mask, const char* what) {
    uint32_t st = 0;
    RTG_HIP(hipStreamSynchronize(stream));
    RTG_HIP(hipMemcpy(&st, ctx->d_status, 4, hipMemcpyDeviceToHost));
    if (st & mask) {
        uint32_t cleared = st & ~mask;
        RTG_HIP(hipMemcpy(ctx->d_status, &cleared, 4, hipMemcpyHostToDevice));
        set_error("%s: device reported status 0x%x (1: index > max_vertex, 2: reference to an unknown BLAS, "
                  "4: traversal stack overflow)",
                  what, st & mask);
        return RTG_ERROR_INVALID_ARGUMENT;
    }
    return RTG_OK;
}

}  // namespace rtg

using namespace rtg;

extern "C" int rtg_blas_query_sizes(rtg_ctx* ctx, const rtg_geometry* geoms, uint32_t n_geom, rtg_build_sizes* out) {
    clear_error();
    RTG_REQUIRE(ctx && out && (geoms || n_geom == 0), RTG_ERROR_INVALID_ARGUMENT, "rtg_blas_query_sizes: null argument");
    uint64_t total = 0;
    for (uint32_t g = 0; g < n_geom; ++g) total += geoms[g].index_count / 3;
    RTG_REQUIRE(total < (1ull << 27), RTG_ERROR_LIMIT, "BLAS has %llu triangles (limit 2^27)",
                (unsigned long long)total);
    // an upper bound: the blob of a build holds its actual node count
    Layout L = blas_layout((uint32_t)total, n_geom, node_cap((uint32_t)total));
    out->acceleration_structure_size = L.total;
    out->build_scratch_size = lbvh_scratch_bytes((uint32_t)total, sizeof(TriRec)) + n_geom * sizeof(GeomDesc);
    out->update_scratch_size = 0;
    return RTG_OK;
}

extern "C" int rtg_tlas_query_sizes(rtg_ctx* ctx, uint32_t count, rtg_build_sizes* out) {
    clear_error();
    RTG_REQUIRE(ctx && out, RTG_ERROR_INVALID_ARGUMENT, "rtg_tlas_query_sizes: null argument");
    RTG_REQUIRE(count < (1u << 27), RTG_ERROR_LIMIT, "TLAS has %u instances (limit 2^27)", count);
    Layout L = tlas_layout(count, node_cap(count), node_cap(count));
    out->acceleration_structure_size = L.total;
    out->build_scratch_size = lbvh_scratch_bytes(count, sizeof(InstRec));
    out->update_scratch_size = 0;
    return RTG_OK;
}

extern "C" int rtg_blas_build(rtg_ctx* ctx, const rtg_geometry* geoms, uint32_t n_geom, uint32_t flags, rtg_blas* out,
                              void* stream) {
    clear_error();
    RTG_REQUIRE(ctx && out && (geoms || n_geom == 0), RTG_ERROR_INVALID_ARGUMENT, "rtg_blas_build: null argument");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    RTG_HIP(hipSetDevice(ctx->device));
    uint64_t total = 0;
    std::vector<GeomDesc> gd(std::max(1u, n_geom));
    for (uint32_t g = 0; g < n_geom; ++g) {
        const rtg_geometry& G = geoms[g];
        RTG_REQUIRE(G.index_count % 3 == 0, RTG_ERROR_INVALID_ARGUMENT,
                    "geometry %u: index_count %u is not a multiple of 3", g, G.index_count);
        RTG_REQUIRE(G.vertex_stride >= 12 && G.vertex_stride % 4 == 0, RTG_ERROR_INVALID_ARGUMENT,
                    "geometry %u: vertex_stride %u (need >= 12, multiple of 4)", g, G.vertex_stride);
        RTG_REQUIRE(G.index_count == 0 || (G.vertex_data && G.index_data), RTG_ERROR_INVALID_ARGUMENT,
                    "geometry %u: null vertex or index data", g);
        gd[g].vertices = reinterpret_cast<const uint8_t*>(G.vertex_data);
        gd[g].indices = reinterpret_cast<const uint32_t*>(G.index_data);
        gd[g].stride = G.vertex_stride;
        gd[g].max_vertex = G.max_vertex;
        gd[g].first_tri = (uint32_t)total;
        gd[g].flags = G.flags;
        total += G.index_count / 3;
        RTG_REQUIRE(total < (1ull << 27), RTG_ERROR_LIMIT, "BLAS exceeds 2^27 triangles");
    }
    uint32_t n = (uint32_t)total;
    size_t need = lbvh_scratch_bytes(n, sizeof(TriRec)) + std::max(1u, n_geom) * sizeof(GeomDesc) + 256;
    if (ctx->scratch.ensure(need) != RTG_OK) return RTG_ERROR_OUT_OF_MEMORY;
    LbvhBufs b = carve_lbvh(ctx->scratch.ptr, n, sizeof(TriRec), n_geom);
    GeomDesc* d_geoms = b.geoms;
    // the copy source must outlive the asynchronous copy: the build
    // synchronises its stream before returning
    RTG_HIP(hipMemcpyAsync(d_geoms, gd.data(), std::max(1u, n_geom) * sizeof(GeomDesc), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_init_bounds, dim3(1), dim3(64), 0, s, b.cb);
    if (n > 0)
        hipLaunchKernelGGL(k_tri_setup, dim3(blocks_for(n, 256)), dim3(256), 0, s, d_geoms, n_geom, n,
                           reinterpret_cast<TriRec*>(b.recs), b.boxes, b.cb, ctx->d_status);
    std::vector<uint32_t> level_off;
    int st = wide_hierarchy(b, n, kBlasLeafMax, (flags & RTG_BUILD_PREFER_FAST_TRACE) ? kTreeletRounds : 0, s,
                            level_off);
    if (st != RTG_OK) {
        hipStreamSynchronize(s);  // pending copies read host memory of this frame
        return st;
    }
    const uint32_t levels = (uint32_t)level_off.size() - 1;
    Layout L = blas_layout(n, n_geom, level_off.back());
    uint8_t* blob = nullptr;
    RTG_HIP(hipMalloc(&blob, L.total));
    hipError_t e = hipMemsetAsync(blob, 0, L.total, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(blob + L.nodes_off, b.nodes, (size_t)L.node_count * sizeof(Node), hipMemcpyDeviceToDevice,
                           s);
    if (e != hipSuccess) {
        hipFree(blob);
        set_error("rtg_blas_build: %s", hipGetErrorString(e));
        return RTG_ERROR_DEVICE;
    }
    if (n > 0)
        hipLaunchKernelGGL(k_gather_tris, dim3(blocks_for(n, 256)), dim3(256), 0, s, b.va,
                           reinterpret_cast<const TriRec*>(b.recs), reinterpret_cast<TriRec*>(blob + L.prims_off), n);
    std::vector<uint32_t> gflags(std::max(1u, n_geom));
    for (uint32_t g = 0; g < n_geom; ++g) gflags[g] = geoms[g].flags;
    if (n_geom) RTG_HIP(hipMemcpyAsync(blob + L.extra_off, gflags.data(), n_geom * 4, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_header, dim3(1), dim3(64), 0, s, reinterpret_cast<AsHeader*>(blob), kKindBlas, L.node_count,
                       n, n_geom, levels, L.nodes_off, L.prims_off, L.extra_off, L.level_off, L.leafbox_off, L.total);
    RTG_HIP(hipGetLastError());
    st = check_status(ctx, s, kStatusBadIndex, "rtg_blas_build");
    if (st != RTG_OK) {
        hipFree(blob);
        return st;
    }
    uint64_t addr = reinterpret_cast<uint64_t>(blob);
    ctx->structures[addr] = AsRecord{kKindBlas, n, L.node_count, L.total};
    st = refresh_registry(ctx);
    if (st != RTG_OK) return st;
    *out = addr;
    return RTG_OK;
}

extern "C" int rtg_tlas_build(rtg_ctx* ctx, const rtg_instance* instances, uint32_t count, uint32_t flags,
                              rtg_tlas* out, void* stream) {
    clear_error();
    RTG_REQUIRE(ctx && out && (instances || count == 0), RTG_ERROR_INVALID_ARGUMENT, "rtg_tlas_build: null argument");
    RTG_REQUIRE(count < (1u << 27), RTG_ERROR_LIMIT, "TLAS has %u instances (limit 2^27)", count);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    RTG_HIP(hipSetDevice(ctx->device));
    uint32_t n = count;
    size_t need = lbvh_scratch_bytes(n, sizeof(InstRec));
    if (ctx->scratch.ensure(need) != RTG_OK) return RTG_ERROR_OUT_OF_MEMORY;
    LbvhBufs b = carve_lbvh(ctx->scratch.ptr, n, sizeof(InstRec));
    hipLaunchKernelGGL(k_init_bounds, dim3(1), dim3(64), 0, s, b.cb);
    if (n > 0)
        hipLaunchKernelGGL(k_inst_setup, dim3(blocks_for(n, 256)), dim3(256), 0, s, instances, n, ctx->d_registry,
                           ctx->registry_count, reinterpret_cast<InstRec*>(b.recs), b.boxes, b.cb, ctx->d_status);
    std::vector<uint32_t> level_off;
    int st = wide_hierarchy(b, n, kTlasLeafMax, (flags & RTG_BUILD_PREFER_FAST_TRACE) ? kTreeletRounds : 0, s,
                            level_off);
    if (st != RTG_OK) {
        hipStreamSynchronize(s);
        return st;
    }
    const uint32_t levels = (uint32_t)level_off.size() - 1;
    Layout L = tlas_layout(n, level_off.back(), levels);
    // level data: level_count, level_offsets[level_count + 1]
    std::vector<uint32_t> level_data(1, levels);
    level_data.insert(level_data.end(), level_off.begin(), level_off.end());
    uint8_t* blob = nullptr;
    RTG_HIP(hipMalloc(&blob, L.total));
    hipError_t e = hipMemsetAsync(blob, 0, L.total, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(blob + L.nodes_off, b.nodes, (size_t)L.node_count * sizeof(Node), hipMemcpyDeviceToDevice,
                           s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(blob + L.level_off, level_data.data(), level_data.size() * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && n > 0)
        e = hipMemcpyAsync(blob + L.extra_off, instances, (size_t)n * 64, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) {
        hipStreamSynchronize(s);
        hipFree(blob);
        set_error("rtg_tlas_build: %s", hipGetErrorString(e));
        return RTG_ERROR_DEVICE;
    }
    if (n > 0)
        hipLaunchKernelGGL(k_gather_insts, dim3(blocks_for(n, 256)), dim3(256), 0, s, b.va,
                           reinterpret_cast<const InstRec*>(b.recs), b.boxes,
                           reinterpret_cast<InstRec*>(blob + L.prims_off), reinterpret_cast<Box*>(blob + L.leafbox_off),
                           n);
    hipLaunchKernelGGL(k_header, dim3(1), dim3(64), 0, s, reinterpret_cast<AsHeader*>(blob), kKindTlas, L.node_count,
                       n, 0u, levels, L.nodes_off, L.prims_off, L.extra_off, L.level_off, L.leafbox_off, L.total);
    RTG_HIP(hipGetLastError());
    // also keeps level_data alive until its copy has finished
    st = check_status(ctx, s, kStatusBadBlas, "rtg_tlas_build");
    if (st != RTG_OK) {
        hipFree(blob);
        return st;
    }
    uint64_t addr = reinterpret_cast<uint64_t>(blob);
    ctx->structures[addr] = AsRecord{kKindTlas, n, L.node_count, L.total};
    *out = addr;
    return RTG_OK;
}

extern "C" int rtg_tlas_refit(rtg_ctx* ctx, rtg_tlas tlas, const rtg_instance* instances, uint32_t count,
                              void* stream) {
    clear_error();
    RTG_REQUIRE(ctx && (instances || count == 0), RTG_ERROR_INVALID_ARGUMENT, "rtg_tlas_refit: null argument");
    auto it = ctx->structures.find(tlas);
    RTG_REQUIRE(it != ctx->structures.end() && it->second.kind == kKindTlas, RTG_ERROR_NOT_FOUND,
                "rtg_tlas_refit: 0x%llx is not a TLAS of this context", (unsigned long long)tlas);
    RTG_REQUIRE(it->second.prim_count == count, RTG_ERROR_INVALID_ARGUMENT,
                "rtg_tlas_refit: %u instances, the TLAS was built with %u", count, it->second.prim_count);
    return launch_refit(ctx, reinterpret_cast<uint8_t*>(tlas), instances, count, reinterpret_cast<hipStream_t>(stream));
}

extern "C" int rtg_tlas_clone(rtg_ctx* ctx, rtg_tlas src, rtg_tlas* out, void* stream) {
    clear_error();
    RTG_REQUIRE(ctx && out, RTG_ERROR_INVALID_ARGUMENT, "rtg_tlas_clone: null argument");
    auto it = ctx->structures.find(src);
    RTG_REQUIRE(it != ctx->structures.end() && it->second.kind == kKindTlas, RTG_ERROR_NOT_FOUND,
                "rtg_tlas_clone: 0x%llx is not a TLAS of this context", (unsigned long long)src);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    RTG_HIP(hipSetDevice(ctx->device));
    uint8_t* blob = nullptr;
    RTG_HIP(hipMalloc(&blob, it->second.size));
    RTG_HIP(hipMemcpyAsync(blob, reinterpret_cast<const void*>(src), it->second.size, hipMemcpyDeviceToDevice, s));
    uint64_t addr = reinterpret_cast<uint64_t>(blob);
    ctx->structures[addr] = it->second;
    *out = addr;
    return RTG_OK;
}

static int destroy_as(rtg_ctx* ctx, uint64_t addr, uint32_t kind, const char* what) {
    clear_error();
    RTG_REQUIRE(ctx, RTG_ERROR_INVALID_ARGUMENT, "%s: null context", what);
    auto it = ctx->structures.find(addr);
    RTG_REQUIRE(it != ctx->structures.end() && it->second.kind == kind, RTG_ERROR_NOT_FOUND,
                "%s: 0x%llx is not a live structure of this kind", what, (unsigned long long)addr);
    RTG_HIP(hipDeviceSynchronize());
    RTG_HIP(hipFree(reinterpret_cast<void*>(addr)));
    ctx->structures.erase(it);
    if (kind == kKindBlas) return refresh_registry(ctx);
    return RTG_OK;
}

extern "C" int rtg_blas_destroy(rtg_ctx* ctx, rtg_blas blas) { return destroy_as(ctx, blas, kKindBlas, "rtg_blas_destroy"); }
extern "C" int rtg_tlas_destroy(rtg_ctx* ctx, rtg_tlas tlas) { return destroy_as(ctx, tlas, kKindTlas, "rtg_tlas_destroy"); }

extern "C" int rtg_as_get_info(rtg_ctx* ctx, uint64_t addr, rtg_as_info* out) {
    clear_error();
    RTG_REQUIRE(ctx && out, RTG_ERROR_INVALID_ARGUMENT, "rtg_as_get_info: null argument");
    auto it = ctx->structures.find(addr);
    RTG_REQUIRE(it != ctx->structures.end(), RTG_ERROR_NOT_FOUND, "rtg_as_get_info: unknown structure 0x%llx",
                (unsigned long long)addr);
    AsHeader h;
    RTG_HIP(hipDeviceSynchronize());
    RTG_HIP(hipMemcpy(&h, reinterpret_cast<const void*>(addr), sizeof(h), hipMemcpyDeviceToHost));
    out->kind = h.kind;
    out->node_count = h.node_count;
    out->prim_count = h.prim_count;
    out->geometry_count = h.geom_count;
    out->max_depth = h.max_depth;
    out->node_width = h.node_width;
    for (int a = 0; a < 3; ++a) {
        out->bounds[a] = h.lo[a];
        out->bounds[3 + a] = h.hi[a];
    }
    out->size_bytes = h.total_size;
    return RTG_OK;
}
