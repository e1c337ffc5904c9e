// rtg_device.hpp -- device code shared by the build and trace kernels (gfx950).
//
// Everything here is compiled with -ffp-contract=off and correctly rounded
// FP32 division and square root, so each written operation is one IEEE-754
// binary32 operation rounded to nearest even. The operation order of every
// function is part of the contract: the CPU oracle evaluates the same
// operations, and hit IDs and occlusion must agree bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rtg.h"

namespace rtg {

// ---------------------------------------------------------------------------
// Acceleration-structure blob layout (one device allocation per structure).
//   header (128 B) | nodes (128 B, breadth first) | prims | extra
//   [| level data | leaf boxes]
// BLAS prims: 48-B triangle records in leaf order
//   {v0.xyz, primitive id} {v1.xyz, geometry id | non-opaque bit 31} {v2.xyz, 0}
// TLAS prims: 64-B instance leaf records in leaf order
//   {world-to-object 3x4 row-major} {BLAS address, instance index, mask | flags<<8}
// TLAS extra: copy of the source rtg_instance records (original order).
// TLAS level data: level_count, level_offsets[level_count+1] (level L is
//   nodes [off[L], off[L+1]) of the breadth-first order, root level 0).
// TLAS leaf boxes: world box of each leaf (24 B), scratch of the refit.
// Node: four child slots as lo_x[4], hi_x[4], lo_y[4], hi_y[4], lo_z[4],
//   hi_z[4], child refs[4], 4 zero words; unused slots are EMPTY with the
//   empty box. A slot box is the fold, from the empty box, of its leaf's
//   primitive boxes in leaf order or of its child node's four slot boxes.
// Child ref: internal node index, or LEAF_BIT | (count-1) << 27 | first,
//   or EMPTY.
// The tree is the binary radix tree of the build (Karras 2012) collapsed
// top-down into 4-wide nodes: a node rooted at binary node r takes r's two
// children and, while it has fewer than four slots, opens the slot whose
// subtree box has the largest surface area (first slot on ties) in place;
// subtrees of at most kBlasLeafMax (BLAS) or kTlasLeafMax (TLAS) primitives
// are leaves and are not opened.
// ---------------------------------------------------------------------------
constexpr uint32_t kMagic = 0x53415452u;  // 'RTAS'
constexpr uint32_t kKindBlas = 0u;
constexpr uint32_t kKindTlas = 1u;
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr uint32_t kSentinel = 0xFFFFFFFEu;
constexpr uint32_t kFirstMask = 0x07FFFFFFu;
constexpr uint32_t kNonOpaqueBit = 0x80000000u;
// 1 + 2*gamma(3) rounded up: far-plane inflation of the robust slab test
// (Ize 2013, "Robust BVH Ray Traversal", JCGT 2(2)).
constexpr uint32_t kFarScaleBits = 0x3f800003u;
constexpr uint32_t kBlasLeafMax = 1u;
constexpr uint32_t kTlasLeafMax = 1u;

struct AsHeader {
    uint32_t magic, kind, node_count, prim_count;
    uint32_t geom_count, max_depth, node_width, flags;
    float lo[3];
    float pad0;
    float hi[3];
    float pad1;
    uint64_t nodes_off, prims_off, extra_off, level_off;
    uint64_t total_size;
    uint64_t leafbox_off;  // TLAS: refit scratch of leaf boxes
    uint64_t reserved[2];
};
static_assert(sizeof(AsHeader) == 128, "header");

struct Node {
    float lo_x[4], hi_x[4];
    float lo_y[4], hi_y[4];
    float lo_z[4], hi_z[4];
    uint32_t child[4];
    uint32_t pad[4];
};
static_assert(sizeof(Node) == 128, "node");

struct TriRec {
    float v0[3];
    uint32_t prim;
    float v1[3];
    uint32_t geom;
    float v2[3];
    uint32_t pad;
};
static_assert(sizeof(TriRec) == 48, "tri");

struct InstRec {
    float w2o[12];
    uint64_t blas;
    uint32_t index;
    uint32_t mask_flags;
};
static_assert(sizeof(InstRec) == 64, "inst");

struct Box {
    float lo[3], hi[3];
};

// ---------------------------------------------------------------------------
// Scalar math (operation order fixed; see the file comment)
// ---------------------------------------------------------------------------
struct V3 {
    float x, y, z;
};
struct V2 {
    float x, y;
};
__host__ __device__ inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__host__ __device__ inline V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__host__ __device__ inline V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__host__ __device__ inline V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__host__ __device__ inline float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__host__ __device__ inline V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// normalize(v) = v * (1 / sqrt(dot(v, v)))
__device__ inline V3 normalize(V3 v) {
    float s = 1.0f / sqrtf(dot(v, v));
    return v * s;
}
// interpolate(a, b, c, bc) = (a*bc.x + b*bc.y) + c*bc.z  (hit_shader_common.glsl:79-81)
__device__ inline V3 interp3(V3 a, V3 b, V3 c, V3 bc) { return (a * bc.x + b * bc.y) + c * bc.z; }
__device__ inline V2 interp2(V2 a, V2 b, V2 c, V3 bc) {
    return V2{(a.x * bc.x + b.x * bc.y) + c.x * bc.z, (a.y * bc.x + b.y * bc.y) + c.y * bc.z};
}
__device__ inline float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
__device__ inline float fractf(float x) { return x - floorf(x); }

// Row-major 3x4 affine: ((m0*x + m1*y) + m2*z) + m3
__device__ inline V3 m34_point(const float* m, V3 p) {
    return v3(((m[0] * p.x + m[1] * p.y) + m[2] * p.z) + m[3], ((m[4] * p.x + m[5] * p.y) + m[6] * p.z) + m[7],
              ((m[8] * p.x + m[9] * p.y) + m[10] * p.z) + m[11]);
}
__device__ inline V3 m34_dir(const float* m, V3 d) {
    return v3((m[0] * d.x + m[1] * d.y) + m[2] * d.z, (m[4] * d.x + m[5] * d.y) + m[6] * d.z,
              (m[8] * d.x + m[9] * d.y) + m[10] * d.z);
}
// mat3(gl_WorldToObject3x4EXT) * n with m the row-major world-to-object matrix
__device__ inline V3 m34_tdir(const float* m, V3 n) {
    return v3((m[0] * n.x + m[4] * n.y) + m[8] * n.z, (m[1] * n.x + m[5] * n.y) + m[9] * n.z,
              (m[2] * n.x + m[6] * n.y) + m[10] * n.z);
}
// Inverse of a row-major 3x4 affine transform by cofactors.
__device__ inline void m34_inverse(const float* m, float* inv) {
    float a = m[0], b = m[1], c = m[2];
    float d = m[4], e = m[5], f = m[6];
    float g = m[8], h = m[9], i = m[10];
    float c00 = e * i - f * h;
    float c01 = f * g - d * i;
    float c02 = d * h - e * g;
    float det = (a * c00 + b * c01) + c * c02;
    float id = 1.0f / det;
    inv[0] = c00 * id;
    inv[1] = (c * h - b * i) * id;
    inv[2] = (b * f - c * e) * id;
    inv[4] = c01 * id;
    inv[5] = (a * i - c * g) * id;
    inv[6] = (c * d - a * f) * id;
    inv[8] = c02 * id;
    inv[9] = (b * g - a * h) * id;
    inv[10] = (a * e - b * d) * id;
#pragma unroll
    for (int r = 0; r < 3; ++r)
        inv[4 * r + 3] = -((inv[4 * r + 0] * m[3] + inv[4 * r + 1] * m[7]) + inv[4 * r + 2] * m[11]);
}
// Column-major 4x4 times vec4: ((c0*x + c1*y) + c2*z) + c3*w
__device__ inline void mat4_mul(const float* m, const float* v, float* out) {
#pragma unroll
    for (int r = 0; r < 4; ++r) out[r] = ((m[0 + r] * v[0] + m[4 + r] * v[1]) + m[8 + r] * v[2]) + m[12 + r] * v[3];
}

__device__ inline float u2f(uint32_t u) { return __uint_as_float(u); }
__device__ inline uint32_t f2u(float f) { return __float_as_uint(f); }

// Loads through the global address space. Pointers read from memory (BLAS
// addresses, table entries) are generic to the compiler, and generic (FLAT)
// loads share a wait counter with the LDS traversal stack; these casts make
// them global loads.
#define RTG_GLOBAL __attribute__((address_space(1)))
typedef float rtg_f4v __attribute__((ext_vector_type(4)));
typedef uint32_t rtg_u4v __attribute__((ext_vector_type(4)));
typedef uint32_t rtg_u2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float4 gld_f4(const void* p) {
    rtg_f4v v = *(const RTG_GLOBAL rtg_f4v*)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint4 gld_u4(const void* p) {
    rtg_u4v v = *(const RTG_GLOBAL rtg_u4v*)p;
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint2 gld_u2(const void* p) {
    rtg_u2v v = *(const RTG_GLOBAL rtg_u2v*)p;
    return make_uint2(v.x, v.y);
}
__device__ __forceinline__ uint32_t gld_u32(const void* p) { return *(const RTG_GLOBAL uint32_t*)p; }
__device__ __forceinline__ uint64_t gld_u64(const void* p) { return *(const RTG_GLOBAL uint64_t*)p; }
__device__ __forceinline__ float gld_f32(const void* p) { return *(const RTG_GLOBAL float*)p; }

constexpr float kPi = 3.14159265358979323846f;

// sin and cos: Cody-Waite reduction by pi/2 (parts of 12, 12 and 24 bits),
// Taylor series of sin to x^9 and cos to x^10 on [-pi/4, pi/4].
__device__ inline void sincos_own(float x, float* s_out, float* c_out) {
    const float two_over_pi = u2f(0x3f22f983u);
    const float p1 = u2f(0x3fc90000u);
    const float p2 = u2f(0x39fda000u);
    const float p3 = u2f(0x33a22169u);
    float qf = rintf(x * two_over_pi);
    int q = (int)qf;
    float r = ((x - qf * p1) - qf * p2) - qf * p3;
    float r2 = r * r;
    float sp = r + (r * r2) * (-1.66666667e-1f + r2 * (8.33333333e-3f + r2 * (-1.98412698e-4f + r2 * 2.75573192e-6f)));
    float cp = 1.0f + r2 * (-0.5f + r2 * (4.16666667e-2f + r2 * (-1.38888889e-3f + r2 * (2.48015873e-5f + r2 * -2.75573192e-7f))));
    float s, c;
    switch (q & 3) {
        case 0: s = sp; c = cp; break;
        case 1: s = cp; c = -sp; break;
        case 2: s = -sp; c = -cp; break;
        default: s = -cp; c = sp; break;
    }
    *s_out = s;
    *c_out = c;
}
__device__ inline float cos_own(float x) {
    float s, c;
    sincos_own(x, &s, &c);
    return c;
}

// log2 via x = m 2^e, m in [sqrt(1/2), sqrt(2)), 2 atanh((m-1)/(m+1)) / ln 2.
__device__ inline float log2_own(float x) {
    if (!(x > 0.0f)) return x == 0.0f ? -__builtin_inff() : __builtin_nanf("");
    if (x == __builtin_inff()) return x;
    int e = 0;
    uint32_t u = f2u(x);
    if ((u >> 23) == 0) {
        x = x * 8388608.0f;
        u = f2u(x);
        e = -23;
    }
    e += (int)((u >> 23) & 0xffu) - 127;
    float m = u2f((u & 0x007fffffu) | 0x3f800000u);
    if (m > u2f(0x3fb504f3u)) {
        m = m * 0.5f;
        e += 1;
    }
    float s = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float series = 1.0f + s2 * (3.33333333e-1f + s2 * (2.0e-1f + s2 * (1.42857143e-1f + s2 * (1.11111111e-1f + s2 * 9.09090909e-2f))));
    float ln_m = (2.0f * s) * series;
    return (float)e + ln_m * u2f(0x3fb8aa3bu);
}
// 2^y via y = n + f, Taylor series of e^(f ln 2) to degree 7, two-step scaling.
__device__ inline float exp2_own(float y) {
    if (y != y) return y;
    if (y >= 128.0f) return __builtin_inff();
    if (y < -150.0f) return 0.0f;
    float nf = rintf(y);
    float f = y - nf;
    float t = f * u2f(0x3f317218u);
    float p = 1.0f + t * (1.0f + t * (0.5f + t * (1.66666667e-1f + t * (4.16666667e-2f + t * (8.33333333e-3f + t * (1.38888889e-3f + t * 1.98412698e-4f))))));
    int n = (int)nf;
    int n1 = n / 2, n2 = n - n1;
    float s1 = u2f((uint32_t)(n1 + 127) << 23);
    float s2 = u2f((uint32_t)(n2 + 127) << 23);
    return (p * s1) * s2;
}
__device__ inline float pow_own(float x, float y) {
    if (x == 0.0f) return y > 0.0f ? 0.0f : (y == 0.0f ? 1.0f : __builtin_inff());
    return exp2_own(y * log2_own(x));
}
// pow(x, 5.0) as ((x*x)*(x*x))*x
__device__ inline float pow5(float x) {
    float x2 = x * x;
    return (x2 * x2) * x;
}

// ---------------------------------------------------------------------------
// Textures: sampler2D lookups (LOD 0, REPEAT wrap, own filtering).
// ---------------------------------------------------------------------------
struct SceneView {
    const rtg_model_info* models;
    const rtg_texture* textures;
    const float* srgb_lut;  // 256 entries (IEC 61966-2-1 decode)
    uint32_t model_count;
    uint32_t texture_count;
};

// The texel index f (an integer-valued float) wrapped into [0, size): the
// non-negative remainder of f modulo size as an int; 0 for |f| >= 2^24.
// |f| < 2^24 fits an int, so 32-bit arithmetic gives the remainder exactly:
// a mask for power-of-two sizes (two's complement wraps negative indices),
// an unsigned remainder of |f| (of -f - 1 for negative f) otherwise.
__device__ inline int wrap_coord(float f, uint32_t size) {
    if (!(fabsf(f) < 16777216.0f)) return 0;
    const int i = (int)f;
    if ((size & (size - 1u)) == 0u) return (int)((uint32_t)i & (size - 1u));
    if (i >= 0) return (int)((uint32_t)i % size);
    return (int)(size - 1u - (uint32_t)(-(i + 1)) % size);
}

// c / 255 correctly rounded for an 8-bit code c: the product with the
// rounded reciprocal, corrected by one fused residual step (exact for all
// 256 codes, which a test checks against the division).
__device__ inline float unorm8(uint32_t c) {
    const float cf = (float)c;
    const float q = cf * (1.0f / 255.0f);
    const float e = fmaf(-q, 255.0f, cf);
    return fmaf(e, 1.0f / 255.0f, q);
}

__device__ inline float4 fetch_texel(const rtg_texture& t, const float* lut, int x, int y) {
    uint64_t idx = (uint64_t)y * t.width + (uint64_t)x;
    if (t.format == RTG_FORMAT_RGBA32_FLOAT) return gld_f4(reinterpret_cast<const float4*>(t.data) + idx);
    uint32_t p = gld_u32(reinterpret_cast<const uint32_t*>(t.data) + idx);
    uint32_t r = p & 0xffu, g = (p >> 8) & 0xffu, b = (p >> 16) & 0xffu, a = p >> 24;
    float4 o;
    if (t.format == RTG_FORMAT_RGBA8_SRGB) {
        o.x = gld_f32(lut + r);
        o.y = gld_f32(lut + g);
        o.z = gld_f32(lut + b);
    } else {
        o.x = unorm8(r);
        o.y = unorm8(g);
        o.z = unorm8(b);
    }
    o.w = unorm8(a);
    return o;
}

__device__ inline float4 sample_tex(const rtg_texture& t, const float* lut, float u, float v) {
    if (t.width == 0 || t.height == 0) return make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    if (t.filter == RTG_FILTER_NEAREST) {
        int x = wrap_coord(floorf(u * (float)t.width), t.width);
        int y = wrap_coord(floorf(v * (float)t.height), t.height);
        return fetch_texel(t, lut, x, y);
    }
    float x = u * (float)t.width - 0.5f;
    float y = v * (float)t.height - 0.5f;
    float fx = floorf(x), fy = floorf(y);
    float a = x - fx, b = y - fy;
    int x0 = wrap_coord(fx, t.width), y0 = wrap_coord(fy, t.height);
    int x1 = x0 + 1 == (int)t.width ? 0 : x0 + 1;
    int y1 = y0 + 1 == (int)t.height ? 0 : y0 + 1;
    float4 t00 = fetch_texel(t, lut, x0, y0), t10 = fetch_texel(t, lut, x1, y0);
    float4 t01 = fetch_texel(t, lut, x0, y1), t11 = fetch_texel(t, lut, x1, y1);
    float ia = 1.0f - a, ib = 1.0f - b;
    float4 o;
    o.x = (t00.x * ia + t10.x * a) * ib + (t01.x * ia + t11.x * a) * b;
    o.y = (t00.y * ia + t10.y * a) * ib + (t01.y * ia + t11.y * a) * b;
    o.z = (t00.z * ia + t10.z * a) * ib + (t01.z * ia + t11.z * a) * b;
    o.w = (t00.w * ia + t10.w * a) * ib + (t01.w * ia + t11.w * a) * b;
    return o;
}

// Descriptor of texture `index` of the table; a missing entry reads as an
// empty texture, which samples as (0, 0, 0, 1).
__device__ inline rtg_texture load_texture(const SceneView& S, uint32_t index) {
    rtg_texture t;
    if (index >= S.texture_count) {
        t.data = 0;
        t.width = t.height = t.format = t.filter = 0;
        return t;
    }
    const rtg_texture* tp = S.textures + index;
    t.data = gld_u64(&tp->data);
    uint4 rest = gld_u4(&tp->width);
    t.width = rest.x;
    t.height = rest.y;
    t.format = rest.z;
    t.filter = rest.w;
    return t;
}

__device__ inline float4 sample_index(const SceneView& S, uint32_t index, float u, float v) {
    return sample_tex(load_texture(S, index), S.srgb_lut, u, v);
}

// ---------------------------------------------------------------------------
// Rays, triangles, boxes
// ---------------------------------------------------------------------------
struct RayCtx {
    V3 o, d, inv;
    float sx, sy, sz;
    int kx, ky, kz;
};

__device__ inline float sel3(V3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

__device__ inline void ray_setup_box(RayCtx& r, V3 o, V3 d) {
    r.o = o;
    r.d = d;
    r.inv = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
}

// Box and shear setup of the watertight test: kz = the dominant axis,
// (kx, ky) the others (swapped if d[kz] < 0), sz = 1/d[kz],
// sx = d[kx]*sz, sy = d[ky]*sz.
__device__ inline void ray_setup(RayCtx& r, V3 o, V3 d) {
    ray_setup_box(r, o, d);
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = ax > ay ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    float dz = sel3(d, kz);
    if (dz < 0.0f) {
        int t = kx;
        kx = ky;
        ky = t;
    }
    r.kx = kx;
    r.ky = ky;
    r.kz = kz;
    r.sz = sel3(r.inv, kz);
    r.sx = sel3(d, kx) * r.sz;
    r.sy = sel3(d, ky) * r.sz;
}

// Watertight ray/triangle test (Woop, Benthin, Wald 2013, JCGT 2(1)).
// Returns true for a hit with tmin <= t <= tmax; (u, v) weight vertices 1, 2.
__device__ inline bool tri_test(const RayCtx& r, V3 p0, V3 p1, V3 p2, float tmin, float tmax, float& t_out,
                                float& u_out, float& v_out) {
    V3 A = p0 - r.o, B = p1 - r.o, C = p2 - r.o;
    float Akz = sel3(A, r.kz), Bkz = sel3(B, r.kz), Ckz = sel3(C, r.kz);
    float Ax = sel3(A, r.kx) - r.sx * Akz, Ay = sel3(A, r.ky) - r.sy * Akz;
    float Bx = sel3(B, r.kx) - r.sx * Bkz, By = sel3(B, r.ky) - r.sy * Bkz;
    float Cx = sel3(C, r.kx) - r.sx * Ckz, Cy = sel3(C, r.ky) - r.sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = (U + V) + W;
    if (det == 0.0f) return false;
    float Az = r.sz * Akz, Bz = r.sz * Bkz, Cz = r.sz * Ckz;
    float T = (U * Az + V * Bz) + W * Cz;
    float t = T / det;
    if (!(t >= tmin && t <= tmax)) return false;
    t_out = t;
    u_out = V / det;
    v_out = W / det;
    return true;
}

// Robust slab test of one slot (Ize 2013) against its near and far planes
// per axis (the hi plane is the near one where the sign bit of the inverse
// direction is set: 1/-0 = -inf counts as negative, and a NaN component
// makes both distances NaN); the far distance is inflated by 1 + 2 gamma(3)
// before the comparison.
__device__ inline bool slab(float nx, float fx, float ny, float fy, float nz, float fz, const RayCtx& r, float tmin,
                            float tcur, float& tn) {
    float tnx = (nx - r.o.x) * r.inv.x, tfx = (fx - r.o.x) * r.inv.x;
    float tny = (ny - r.o.y) * r.inv.y, tfy = (fy - r.o.y) * r.inv.y;
    float tnz = (nz - r.o.z) * r.inv.z, tfz = (fz - r.o.z) * r.inv.z;
    float tnear = fmaxf(fmaxf(fmaxf(tnx, tny), tnz), tmin);
    float tfar = fminf(fminf(fminf(tfx, tfy), tfz) * u2f(kFarScaleBits), tcur);
    tn = tnear;
    return tnear <= tfar;
}

// ---------------------------------------------------------------------------
// Two-level traversal
// ---------------------------------------------------------------------------
struct HitRec {
    float t, u, v;
    uint32_t inst, prim, geom;
};

struct TravCounters {
    uint32_t nodes, tris, insts, anyhits, spills;
};

struct TlasView {
    const uint8_t* base;
    const Node* nodes;
    const InstRec* recs;
    const rtg_instance* src;
};

__device__ inline TlasView tlas_view(const uint8_t* base) {
    const AsHeader* h = reinterpret_cast<const AsHeader*>(base);
    TlasView v;
    v.base = base;
    v.nodes = reinterpret_cast<const Node*>(base + gld_u64(&h->nodes_off));
    v.recs = reinterpret_cast<const InstRec*>(base + gld_u64(&h->prims_off));
    v.src = reinterpret_cast<const rtg_instance*>(base + gld_u64(&h->extra_off));
    return v;
}

__device__ inline bool key_less(uint32_t ia, uint32_t ga, uint32_t pa, uint32_t ib, uint32_t gb, uint32_t pb) {
    if (ia != ib) return ia < ib;
    if (ga != gb) return ga < gb;
    return pa < pb;
}

__device__ inline bool is_opaque(uint32_t geom_word, uint32_t inst_flags, uint32_t ray_flags) {
    bool opaque = (geom_word & kNonOpaqueBit) == 0;
    if (inst_flags & RTG_INSTANCE_FORCE_OPAQUE) opaque = true;
    else if (inst_flags & RTG_INSTANCE_FORCE_NO_OPAQUE) opaque = false;
    if (ray_flags & RTG_RAY_FLAG_OPAQUE) opaque = true;
    else if (ray_flags & RTG_RAY_FLAG_NO_OPAQUE) opaque = false;
    return opaque;
}

// any_hit_alpha_clip (shaders/any_hit_alpha_clip.glsl:11-29): keep iff alpha >= 0.5.
// Out of line: the traversal loops call it only for non-opaque geometry, and
// with its texture code inlined the loops measured up to 10 % slower.
__device__ __noinline__ bool alpha_clip_keep(const SceneView& S, const rtg_instance* src, uint32_t inst_index, uint32_t geom,
                                       uint32_t prim, float u, float v) {
    if (S.models == nullptr) return true;
    uint32_t custom = gld_u32(&src[inst_index].custom_index_and_mask) & 0xffffffu;
    if (custom >= S.model_count) return true;
    const rtg_model_info* mi = S.models + custom;
    const rtg_geometry_info* gi =
        reinterpret_cast<const rtg_geometry_info*>(gld_u64(&mi->geometry_info_address)) + geom;
    const uint32_t* idx = reinterpret_cast<const uint32_t*>(gld_u64(&gi->index_buffer_address)) + 3ull * prim;
    const float* uvs = reinterpret_cast<const float*>(gld_u64(&mi->uv_buffer_address));
    V3 bc = v3((1.0f - u) - v, u, v);
    uint32_t i0 = gld_u32(idx), i1 = gld_u32(idx + 1), i2 = gld_u32(idx + 2);
    V2 a{gld_f32(uvs + 2 * i0), gld_f32(uvs + 2 * i0 + 1)}, b{gld_f32(uvs + 2 * i1), gld_f32(uvs + 2 * i1 + 1)},
        c{gld_f32(uvs + 2 * i2), gld_f32(uvs + 2 * i2 + 1)};
    V2 uv = interp2(a, b, c, bc);
    float4 tx = sample_index(S, gld_u32(&gi->diffuse_image_index), uv.x, uv.y);
    return !(tx.w < 0.5f);
}

// Per-lane traversal stack: the first entries live in LDS (entry-major so
// that lanes never conflict), deeper ones in a per-lane global spill area of
// kSpillLaneBytes. A closest-hit entry is 8 B, a node reference and its entry
// distance (popped entries beyond the closest hit are skipped); an occlusion
// entry is the 4-B reference alone, since an occlusion query never shortens
// its interval (the first accepted hit ends it), so twice as many fit in the
// same LDS. The pair walk of two shadow rays (PairTraverser) keeps 8-B
// entries, a reference and the mask of the rays overlapping it. How many
// 8-B entries a lane keeps in LDS is a launch setting (8 to 32, 16 by
// default; a 256-lane workgroup takes entries x 8 B x 256 of LDS); 4-B
// stacks keep twice as many. All kinds hold kStackMax entries in all.
constexpr int kStackLdsMin = 8;
constexpr int kStackLdsMax = 32;
constexpr int kStackLdsDefault = 16;
constexpr int kStackMax = 96;
// bytes of one lane's spill area: the closest-hit entries past the fewest
// kept in LDS
constexpr size_t kSpillLaneBytes = (size_t)(kStackMax - kStackLdsMin) * 8;
static_assert((kStackMax - 2 * kStackLdsMin) * 4 <= (int)kSpillLaneBytes, "occlusion spill fits the lane's area");
// LDS bytes of a 256-lane workgroup's stacks with `entries` closest-hit
// entries per lane (either kind)
__host__ __device__ constexpr size_t stack_lds_bytes(int entries) { return (size_t)entries * 8 * 256; }

template <bool OCC>
struct Stack;

template <>
struct Stack<false> {
    uint2* lds;      // lds[k * stride + lane]
    uint2* spill;    // spill[k - lds_n], per lane
    int stride;
    int lds_n;       // entries in LDS
    int sp;
    // `entries`: the closest-hit entries per lane in LDS (the launch setting)
    __device__ inline void init(void* lds_words, int entries, uint8_t* spill_area, uint32_t gtid, int lane,
                                int lanes) {
        lds = static_cast<uint2*>(lds_words) + lane;
        stride = lanes;
        lds_n = entries;
        spill = reinterpret_cast<uint2*>(spill_area + (size_t)gtid * kSpillLaneBytes);
        sp = 0;
    }
    __device__ inline void push(uint32_t ref, float tn) {
        uint2 e = make_uint2(ref, f2u(tn));
        if (sp < lds_n) {
            lds[sp * stride] = e;
        } else {
            rtg_u2v v = {e.x, e.y};
            *(RTG_GLOBAL rtg_u2v*)(spill + (sp - lds_n)) = v;
        }
        ++sp;
    }
    __device__ inline uint2 pop() {
        --sp;
        return sp < lds_n ? lds[sp * stride] : gld_u2(spill + (sp - lds_n));
    }
};

template <>
struct Stack<true> {
    uint32_t* lds;
    uint32_t* spill;
    int stride;
    int lds_n;
    int sp;
    // `entries`: the closest-hit entries per lane in LDS; occlusion entries
    // are half their size, so twice as many fit
    __device__ inline void init(void* lds_words, int entries, uint8_t* spill_area, uint32_t gtid, int lane,
                                int lanes) {
        lds = static_cast<uint32_t*>(lds_words) + lane;
        stride = lanes;
        lds_n = 2 * entries;
        spill = reinterpret_cast<uint32_t*>(spill_area + (size_t)gtid * kSpillLaneBytes);
        sp = 0;
    }
    __device__ inline void push(uint32_t ref, float) {
        if (sp < lds_n) lds[sp * stride] = ref;
        else *(RTG_GLOBAL uint32_t*)(spill + (sp - lds_n)) = ref;
        ++sp;
    }
    // (reference, -inf): an occlusion entry is never skipped
    __device__ inline uint2 pop() {
        --sp;
        uint32_t r = sp < lds_n ? lds[sp * stride] : gld_u32(spill + (sp - lds_n));
        return make_uint2(r, 0xff800000u);
    }
};

// Final states of a traversal step.
constexpr int kRayActive = 0;
constexpr int kRayHit = 1;
constexpr int kRayMiss = 2;
constexpr int kRayOverflow = 3;

// Two-level traversal as a resumable state machine: start() sets up a ray,
// each step() does one visit (an internal node, an instance entry or a leaf
// with the pops that follow it). The order of visits is fixed: depth first,
// nearer child first (ties to child 0) for closest-hit queries, children in
// slot order for occlusion queries (any accepted hit ends them, so the order
// only changes how soon), popped entries skipped once their entry distance
// exceeds the closest hit.
template <bool OCCLUSION, bool COUNT>
struct Traverser {
    RayCtx world, obj;
    const Node* nodes;
    const TriRec* tris;
    float tmin, tcur;
    uint32_t ray_flags, cull_mask, sbt_offset;
    uint32_t inst_index, inst_flags;
    uint32_t cur;
    bool in_blas;
    HitRec best;  // best.inst stays RTG_NO_HIT until a hit is accepted

    using StackT = Stack<OCCLUSION>;

    __device__ inline void start(const TlasView& T, V3 o, V3 d, float tmin_, float tmax_, uint32_t flags,
                                 uint32_t mask, uint32_t sbt, StackT& st) {
        ray_setup_box(world, o, d);
        in_blas = false;
        nodes = T.nodes;
        tris = nullptr;
        tmin = tmin_;
        tcur = tmax_;
        ray_flags = flags;
        cull_mask = mask;
        sbt_offset = sbt;
        inst_index = inst_flags = 0;
        best.t = tmax_;
        best.u = best.v = 0.0f;
        best.inst = best.prim = best.geom = RTG_NO_HIT;
        st.sp = 0;
        cur = 0;
    }

    __device__ inline int done() const { return best.inst != RTG_NO_HIT ? kRayHit : kRayMiss; }

    // Next entry to visit; false once the stack is empty.
    __device__ inline bool pop(const TlasView& T, StackT& st) {
        while (st.sp > 0) {
            uint2 e = st.pop();
            if (e.x == kSentinel) {
                in_blas = false;
                nodes = T.nodes;
                continue;
            }
            if (!OCCLUSION && u2f(e.y) > tcur) continue;
            cur = e.x;
            return true;
        }
        return false;
    }

    // pushes from stack position sp0 on that went past the LDS part
    __device__ static inline uint32_t spilled(int sp0, int sp1, int lds) {
        const int from = sp0 > lds ? sp0 : lds;
        return sp1 > from ? (uint32_t)(sp1 - from) : 0u;
    }

    // An internal node: slab tests of its four slots; the nearest overlapped
    // child (the first in slot order for occlusion queries) is next, the
    // others are pushed.
    __device__ inline int node_step(const TlasView& T, StackT& st, TravCounters& cnt) {
        const RayCtx& cr = in_blas ? obj : world;
        // Each axis's lo and hi float4s lie side by side in the node, so the
        // near and far planes of all four slots are picked by address (the
        // sign bit of the inverse direction) instead of by a select per slot.
        const float4* np = reinterpret_cast<const float4*>(nodes + cur);
        const uint32_t ix = f2u(cr.inv.x) >> 31, iy = f2u(cr.inv.y) >> 31, iz = f2u(cr.inv.z) >> 31;
        float4 NX = gld_f4(np + ix), FX = gld_f4(np + (ix ^ 1u));
        float4 NY = gld_f4(np + 2 + iy), FY = gld_f4(np + 2 + (iy ^ 1u));
        float4 NZ = gld_f4(np + 4 + iz), FZ = gld_f4(np + 4 + (iz ^ 1u));
        uint4 C = gld_u4(np + 6);
        if (COUNT) cnt.nodes++;
        float t0, t1, t2, t3;
        bool h0 = slab(NX.x, FX.x, NY.x, FY.x, NZ.x, FZ.x, cr, tmin, tcur, t0) && C.x != kEmpty;
        bool h1 = slab(NX.y, FX.y, NY.y, FY.y, NZ.y, FZ.y, cr, tmin, tcur, t1) && C.y != kEmpty;
        bool h2 = slab(NX.z, FX.z, NY.z, FY.z, NZ.z, FZ.z, cr, tmin, tcur, t2) && C.z != kEmpty;
        bool h3 = slab(NX.w, FX.w, NY.w, FY.w, NZ.w, FZ.w, cr, tmin, tcur, t3) && C.w != kEmpty;
        const int nh = (int)h0 + (int)h1 + (int)h2 + (int)h3;
        if (nh == 0) {
            if (!pop(T, st)) return done();
            return kRayActive;
        }
        if (st.sp + nh - 1 > kStackMax) return kRayOverflow;
        if (COUNT) cnt.spills += spilled(st.sp, st.sp + nh - 1, st.lds_n);
        if (OCCLUSION) {
            // slot order: the lowest overlapped slot is visited next, the
            // others are pushed highest first
            if (h3 && (h0 || h1 || h2)) st.push(C.w, t3);
            if (h2 && (h0 || h1)) st.push(C.z, t2);
            if (h1 && h0) st.push(C.y, t1);
            cur = h0 ? C.x : h1 ? C.y : h2 ? C.z : C.w;
            return kRayActive;
        }
        // rank of each overlapped slot in ascending (entry distance, slot)
        // order; the rank-0 slot is visited next, ranks nh-1 .. 1 are
        // pushed. A rank held by none of slots 0-2 is slot 3's.
        const int r0 = (h1 && t1 < t0) + (h2 && t2 < t0) + (h3 && t3 < t0);
        const int r1 = (h0 && t0 <= t1) + (h2 && t2 < t1) + (h3 && t3 < t1);
        const int r2 = (h0 && t0 <= t2) + (h1 && t1 <= t2) + (h3 && t3 < t2);
#pragma unroll
        for (int q = 3; q >= 1; --q) {
            if (q < nh) {
                uint32_t ref = (h0 && r0 == q) ? C.x : (h1 && r1 == q) ? C.y : (h2 && r2 == q) ? C.z : C.w;
                float tq = (h0 && r0 == q) ? t0 : (h1 && r1 == q) ? t1 : (h2 && r2 == q) ? t2 : t3;
                st.push(ref, tq);
            }
        }
        cur = (h0 && r0 == 0) ? C.x : (h1 && r1 == 0) ? C.y : (h2 && r2 == 0) ? C.z : C.w;
        return kRayActive;
    }

    // A TLAS leaf: enter its instance's BLAS (the object-space ray, a
    // sentinel on the stack) or, if the cull mask rejects it, skip it.
    __device__ inline int instance_step(const TlasView& T, StackT& st, TravCounters& cnt) {
        const InstRec* rec = T.recs + (cur & kFirstMask);
        const float4* rp = reinterpret_cast<const float4*>(rec);
        uint4 tail = gld_u4(rp + 3);  // w2o[8..11] is rp[2]; rp[3] = blas lo, blas hi, index, mask_flags
        uint32_t mf = tail.w;
        const uint8_t* bb = reinterpret_cast<const uint8_t*>((uint64_t)tail.x | ((uint64_t)tail.y << 32));
        if (((mf & 0xffu) & cull_mask) == 0 || bb == nullptr) {
            if (!pop(T, st)) return done();
            return kRayActive;
        }
        if (COUNT) cnt.insts++;
        if (st.sp >= kStackMax) return kRayOverflow;
        if (COUNT) cnt.spills += spilled(st.sp, st.sp + 1, st.lds_n);
        st.push(kSentinel, -__builtin_inff());
        float w2o[12];
        float4 m0 = gld_f4(rp), m1 = gld_f4(rp + 1), m2 = gld_f4(rp + 2);
        w2o[0] = m0.x; w2o[1] = m0.y; w2o[2] = m0.z; w2o[3] = m0.w;
        w2o[4] = m1.x; w2o[5] = m1.y; w2o[6] = m1.z; w2o[7] = m1.w;
        w2o[8] = m2.x; w2o[9] = m2.y; w2o[10] = m2.z; w2o[11] = m2.w;
        ray_setup(obj, m34_point(w2o, world.o), m34_dir(w2o, world.d));
        const AsHeader* bh = reinterpret_cast<const AsHeader*>(bb);
        nodes = reinterpret_cast<const Node*>(bb + gld_u64(&bh->nodes_off));
        tris = reinterpret_cast<const TriRec*>(bb + gld_u64(&bh->prims_off));
        inst_index = tail.z;
        inst_flags = (mf >> 8) & 0xffu;
        in_blas = true;
        cur = 0;
        return kRayActive;
    }

    // A BLAS leaf: its triangles, then the next stack entry.
    __device__ inline int leaf_step(const TlasView& T, const SceneView& S, StackT& st, TravCounters& cnt) {
        uint32_t first = cur & kFirstMask;
        uint32_t count = ((cur >> 27) & 0xfu) + 1u;
        for (uint32_t k = first; k < first + count; ++k) {
            const float4* tp = reinterpret_cast<const float4*>(tris + k);
            float4 a = gld_f4(tp), b = gld_f4(tp + 1), c = gld_f4(tp + 2);
            if (COUNT) cnt.tris++;
            float t, u, v;
            if (!tri_test(obj, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(c.x, c.y, c.z), tmin, tcur, t, u, v))
                continue;
            uint32_t prim = f2u(a.w), gword = f2u(b.w);
            uint32_t geom = gword & ~kNonOpaqueBit;
            if (!OCCLUSION &&
                !(t < best.t || (t == best.t && key_less(inst_index, geom, prim, best.inst, best.geom, best.prim))))
                continue;
            if (!is_opaque(gword, inst_flags, ray_flags)) {
                uint32_t group = (gld_u32(&T.src[inst_index].sbt_offset_and_flags) & 0xffffffu) + sbt_offset;
                if (group < RTG_HIT_GROUP_COUNT) {
                    if (COUNT) cnt.anyhits++;
                    if (!alpha_clip_keep(S, T.src, inst_index, geom, prim, u, v)) continue;
                }
            }
            best.t = t;
            best.u = u;
            best.v = v;
            best.inst = inst_index;
            best.geom = geom;
            best.prim = prim;
            if (OCCLUSION || (ray_flags & RTG_RAY_FLAG_TERMINATE_ON_FIRST_HIT)) return kRayHit;
            tcur = t;
        }
        if (!pop(T, st)) return done();
        return kRayActive;
    }

    // One iteration: a BLAS leaf, then a TLAS leaf, then an internal node,
    // each if it is what the ray has to visit at that point. A ray thus
    // tests a leaf and enters an instance or visits the node that follows in
    // the same iteration, so a wavefront's lanes, which need different kinds
    // of visits at once, each advance further per iteration. The visits a ray
    // makes and their order are those of one visit per iteration.
    // leaf_turn / inst_turn (wave-uniform) let a caller hold the leaf or the
    // instance visits back for an iteration: a wave runs the code of a visit
    // kind whenever one of its lanes needs it, and leaf and instance visits
    // are rare per lane but needed by some lane in nearly every iteration.
    __device__ inline int step(const TlasView& T, const SceneView& S, StackT& st, TravCounters& cnt,
                               bool leaf_turn = true, bool inst_turn = true) {
        if ((cur & kLeafBit) && in_blas) {
            if (!leaf_turn) return kRayActive;
            const int r = leaf_step(T, S, st, cnt);
            if (r != kRayActive) return r;
        }
        if ((cur & kLeafBit) && !in_blas) {
            if (!inst_turn) return kRayActive;
            const int r = instance_step(T, st, cnt);
            if (r != kRayActive) return r;
        }
        if (!(cur & kLeafBit)) return node_step(T, st, cnt);
        return kRayActive;
    }
};

// Two shadow rays of one path (one origin and interval, directions in the
// light's cone) traced by one lane: they visit the union of the nodes either
// ray overlaps, once, instead of each in a lane of its own. Stack entries
// (8 B: reference, mask of the rays that overlap it) and visits follow the
// single-ray occlusion order (depth first, children in slot order); a ray
// leaves the pair's mask once it has an accepted hit, and the pair is done
// when both have one or the stack is empty. Each ray's result is what the
// single-ray query gives: the nodes and triangles tested for a ray are a
// superset of those it overlaps, and a hit is accepted or not per ray.
// A lane of mask 1 traces one ray alone (an odd shadow-ray count's last).
template <bool COUNT>
struct PairTraverser {
    V3 wo, wd0, wd1, winv0, winv1;  // world rays (shared origin)
    V3 oo, oinv0, oinv1;            // object-space rays in a BLAS
    float sx0, sy0, sz0, sx1, sy1, sz1;
    uint32_t k0, k1;                // shear axes kx | ky << 2 | kz << 4
    const Node* nodes;
    const TriRec* tris;
    float tmin, tcur;
    uint32_t inst_index, inst_flags;
    uint32_t cur, mask, alive;
    bool in_blas;
    using StackT = Stack<false>;

    // Shear of the watertight test (as ray_setup) for direction d.
    __device__ static inline void shear(V3 d, V3 inv, float& sx, float& sy, float& sz, uint32_t& k) {
        float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
        int kz = ax > ay ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
        int kx = kz + 1 == 3 ? 0 : kz + 1;
        int ky = kx + 1 == 3 ? 0 : kx + 1;
        if (sel3(d, kz) < 0.0f) {
            int t = kx;
            kx = ky;
            ky = t;
        }
        sz = sel3(inv, kz);
        sx = sel3(d, kx) * sz;
        sy = sel3(d, ky) * sz;
        k = (uint32_t)kx | ((uint32_t)ky << 2) | ((uint32_t)kz << 4);
    }
    __device__ static inline V3 inv3(V3 d) { return v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z); }

    // m: 3 for a pair, 1 for ray 0 alone (d1 is then unused).
    __device__ inline void start(const TlasView& T, V3 o, V3 d0, V3 d1, uint32_t m, float tmin_, float tmax_,
                                 StackT& st) {
        wo = o;
        wd0 = d0;
        wd1 = d1;
        winv0 = inv3(d0);
        winv1 = inv3(d1);
        in_blas = false;
        nodes = T.nodes;
        tris = nullptr;
        tmin = tmin_;
        tcur = tmax_;
        inst_index = inst_flags = 0;
        st.sp = 0;
        cur = 0;
        mask = m;
        alive = m;
    }
    // pushes from stack position sp0 on that went past the LDS part
    __device__ static inline uint32_t spilled(int sp0, int sp1, int lds) {
        const int from = sp0 > lds ? sp0 : lds;
        return sp1 > from ? (uint32_t)(sp1 - from) : 0u;
    }
    __device__ inline bool pop(const TlasView& T, StackT& st) {
        while (st.sp > 0) {
            uint2 e = st.pop();
            if (e.x == kSentinel) {
                in_blas = false;
                nodes = T.nodes;
                continue;
            }
            if ((e.y & alive) == 0) continue;
            cur = e.x;
            mask = e.y & alive;
            return true;
        }
        return false;
    }
    __device__ inline int node_step(const TlasView& T, StackT& st, TravCounters& cnt) {
        const V3 o = in_blas ? oo : wo;
        const V3 i0 = in_blas ? oinv0 : winv0;
        const V3 i1 = in_blas ? oinv1 : winv1;
        // slab planes picked by address for ray 0 (as in the single-ray
        // visit); ray 1 uses them as they are when its direction has the
        // same signs in every lane of the wave, else swaps them per axis
        const float4* np = reinterpret_cast<const float4*>(nodes + cur);
        const uint32_t ix = f2u(i0.x) >> 31, iy = f2u(i0.y) >> 31, iz = f2u(i0.z) >> 31;
        float4 NX = gld_f4(np + ix), FX = gld_f4(np + (ix ^ 1u));
        float4 NY = gld_f4(np + 2 + iy), FY = gld_f4(np + 2 + (iy ^ 1u));
        float4 NZ = gld_f4(np + 4 + iz), FZ = gld_f4(np + 4 + (iz ^ 1u));
        uint4 C = gld_u4(np + 6);
        if (COUNT) cnt.nodes++;
        RayCtx r0, r1;
        r0.o = o;
        r0.inv = i0;
        r1.o = o;
        r1.inv = i1;
        float t;
        const bool m0 = (mask & 1u) != 0, m1 = (mask & 2u) != 0;
        const uint32_t jx = f2u(i1.x) >> 31, jy = f2u(i1.y) >> 31, jz = f2u(i1.z) >> 31;
        uint32_t c0 = (m0 && slab(NX.x, FX.x, NY.x, FY.x, NZ.x, FZ.x, r0, tmin, tcur, t)) ? 1u : 0u;
        uint32_t c1 = (m0 && slab(NX.y, FX.y, NY.y, FY.y, NZ.y, FZ.y, r0, tmin, tcur, t)) ? 1u : 0u;
        uint32_t c2 = (m0 && slab(NX.z, FX.z, NY.z, FY.z, NZ.z, FZ.z, r0, tmin, tcur, t)) ? 1u : 0u;
        uint32_t c3 = (m0 && slab(NX.w, FX.w, NY.w, FY.w, NZ.w, FZ.w, r0, tmin, tcur, t)) ? 1u : 0u;
        if (__ballot(jx != ix || jy != iy || jz != iz) == 0) {
            c0 |= (m1 && slab(NX.x, FX.x, NY.x, FY.x, NZ.x, FZ.x, r1, tmin, tcur, t)) ? 2u : 0u;
            c1 |= (m1 && slab(NX.y, FX.y, NY.y, FY.y, NZ.y, FZ.y, r1, tmin, tcur, t)) ? 2u : 0u;
            c2 |= (m1 && slab(NX.z, FX.z, NY.z, FY.z, NZ.z, FZ.z, r1, tmin, tcur, t)) ? 2u : 0u;
            c3 |= (m1 && slab(NX.w, FX.w, NY.w, FY.w, NZ.w, FZ.w, r1, tmin, tcur, t)) ? 2u : 0u;
        } else {
            const bool sx = jx != ix, sy = jy != iy, sz = jz != iz;
            const float4 AX = sx ? FX : NX, BX = sx ? NX : FX, AY = sy ? FY : NY, BY = sy ? NY : FY;
            const float4 AZ = sz ? FZ : NZ, BZ = sz ? NZ : FZ;
            c0 |= (m1 && slab(AX.x, BX.x, AY.x, BY.x, AZ.x, BZ.x, r1, tmin, tcur, t)) ? 2u : 0u;
            c1 |= (m1 && slab(AX.y, BX.y, AY.y, BY.y, AZ.y, BZ.y, r1, tmin, tcur, t)) ? 2u : 0u;
            c2 |= (m1 && slab(AX.z, BX.z, AY.z, BY.z, AZ.z, BZ.z, r1, tmin, tcur, t)) ? 2u : 0u;
            c3 |= (m1 && slab(AX.w, BX.w, AY.w, BY.w, AZ.w, BZ.w, r1, tmin, tcur, t)) ? 2u : 0u;
        }
        if (C.x == kEmpty) c0 = 0u;
        if (C.y == kEmpty) c1 = 0u;
        if (C.z == kEmpty) c2 = 0u;
        if (C.w == kEmpty) c3 = 0u;
        const int nh = (int)(c0 != 0u) + (int)(c1 != 0u) + (int)(c2 != 0u) + (int)(c3 != 0u);
        if (nh == 0) {
            if (!pop(T, st)) return kRayMiss;
            return kRayActive;
        }
        if (st.sp + nh - 1 > kStackMax) return kRayOverflow;
        if (COUNT) cnt.spills += spilled(st.sp, st.sp + nh - 1, st.lds_n);
        // the lowest overlapped slot is visited next, the others are
        // pushed highest first, each with its mask
        if (c3 && (c0 || c1 || c2)) st.push(C.w, u2f(c3));
        if (c2 && (c0 || c1)) st.push(C.z, u2f(c2));
        if (c1 && c0) st.push(C.y, u2f(c1));
        cur = c0 ? C.x : c1 ? C.y : c2 ? C.z : C.w;
        mask = c0 ? c0 : c1 ? c1 : c2 ? c2 : c3;
        return kRayActive;
    }
    __device__ inline int instance_step(const TlasView& T, StackT& st, TravCounters& cnt) {
        const InstRec* rec = T.recs + (cur & kFirstMask);
        const float4* rp = reinterpret_cast<const float4*>(rec);
        uint4 tail = gld_u4(rp + 3);  // blas lo, blas hi, index, mask_flags
        uint32_t mf = tail.w;
        const uint8_t* bb = reinterpret_cast<const uint8_t*>((uint64_t)tail.x | ((uint64_t)tail.y << 32));
        if ((mf & 0xffu) == 0 || bb == nullptr) {  // shadow rays' cull mask is 0xff
            if (!pop(T, st)) return kRayMiss;
            return kRayActive;
        }
        if (COUNT) cnt.insts++;
        if (st.sp >= kStackMax) return kRayOverflow;
        if (COUNT) cnt.spills += spilled(st.sp, st.sp + 1, st.lds_n);
        st.push(kSentinel, u2f(0u));
        float w2o[12];
        float4 m0 = gld_f4(rp), m1 = gld_f4(rp + 1), m2 = gld_f4(rp + 2);
        w2o[0] = m0.x; w2o[1] = m0.y; w2o[2] = m0.z; w2o[3] = m0.w;
        w2o[4] = m1.x; w2o[5] = m1.y; w2o[6] = m1.z; w2o[7] = m1.w;
        w2o[8] = m2.x; w2o[9] = m2.y; w2o[10] = m2.z; w2o[11] = m2.w;
        oo = m34_point(w2o, wo);
        const V3 d0 = m34_dir(w2o, wd0), d1 = m34_dir(w2o, wd1);
        oinv0 = inv3(d0);
        oinv1 = inv3(d1);
        shear(d0, oinv0, sx0, sy0, sz0, k0);
        shear(d1, oinv1, sx1, sy1, sz1, k1);
        const AsHeader* bh = reinterpret_cast<const AsHeader*>(bb);
        nodes = reinterpret_cast<const Node*>(bb + gld_u64(&bh->nodes_off));
        tris = reinterpret_cast<const TriRec*>(bb + gld_u64(&bh->prims_off));
        inst_index = tail.z;
        inst_flags = (mf >> 8) & 0xffu;
        in_blas = true;
        cur = 0;
        return kRayActive;
    }
    __device__ inline int leaf_step(const TlasView& T, const SceneView& S, StackT& st, TravCounters& cnt) {
        uint32_t first = cur & kFirstMask;
        uint32_t count = ((cur >> 27) & 0xfu) + 1u;
        for (uint32_t k = first; k < first + count; ++k) {
            const float4* tp = reinterpret_cast<const float4*>(tris + k);
            float4 a = gld_f4(tp), b = gld_f4(tp + 1), c = gld_f4(tp + 2);
            if (COUNT) cnt.tris++;
            for (uint32_t r = 0; r < 2; ++r) {
                const uint32_t bit = 1u << r;
                if (!(mask & alive & bit)) continue;
                RayCtx rc;
                rc.o = oo;
                rc.sx = r ? sx1 : sx0;
                rc.sy = r ? sy1 : sy0;
                rc.sz = r ? sz1 : sz0;
                const uint32_t kk = r ? k1 : k0;
                rc.kx = (int)(kk & 3u);
                rc.ky = (int)((kk >> 2) & 3u);
                rc.kz = (int)((kk >> 4) & 3u);
                float t, u, v;
                if (!tri_test(rc, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(c.x, c.y, c.z), tmin, tcur, t, u, v))
                    continue;
                const uint32_t prim = f2u(a.w), gword = f2u(b.w);
                const uint32_t geom = gword & ~kNonOpaqueBit;
                if (!is_opaque(gword, inst_flags,
                               RTG_RAY_FLAG_TERMINATE_ON_FIRST_HIT | RTG_RAY_FLAG_SKIP_CLOSEST_HIT_SHADER)) {
                    // hit group offset 1 (closest_hit_textured.glsl:159-172)
                    uint32_t group = (gld_u32(&T.src[inst_index].sbt_offset_and_flags) & 0xffffffu) + 1u;
                    if (group < RTG_HIT_GROUP_COUNT) {
                        if (COUNT) cnt.anyhits++;
                        if (!alpha_clip_keep(S, T.src, inst_index, geom, prim, u, v)) continue;
                    }
                }
                alive &= ~bit;
            }
            if (alive == 0u) return kRayHit;
        }
        if (!pop(T, st)) return kRayMiss;
        return kRayActive;
    }
    // One iteration, as Traverser::step. Returns kRayHit once both rays
    // have a hit, kRayMiss once the stack is empty (occluded() then names
    // the rays that have one), or kRayOverflow.
    __device__ inline int step(const TlasView& T, const SceneView& S, StackT& st, TravCounters& cnt,
                               bool leaf_turn = true, bool inst_turn = true) {
        if ((cur & kLeafBit) && in_blas) {
            if (!leaf_turn) return kRayActive;
            const int r = leaf_step(T, S, st, cnt);
            if (r != kRayActive) return r;
        }
        if ((cur & kLeafBit) && !in_blas) {
            if (!inst_turn) return kRayActive;
            const int r = instance_step(T, st, cnt);
            if (r != kRayActive) return r;
        }
        if (!(cur & kLeafBit)) return node_step(T, st, cnt);
        return kRayActive;
    }
    // bit r set: ray r has an accepted hit (given the mask it started with)
    __device__ inline uint32_t occluded(uint32_t started) const { return started & ~alive; }
};

// Runs one ray to completion. Returns 1 for an accepted hit, 0 for none, -1
// if the stack overflowed.
template <bool OCCLUSION, bool COUNT>
__device__ inline int traverse(const TlasView& T, const SceneView& S, V3 o, V3 d, float tmin, float tmax,
                               uint32_t ray_flags, uint32_t cull_mask, uint32_t sbt_offset, HitRec& best,
                               Stack<OCCLUSION>& st, TravCounters& cnt) {
    Traverser<OCCLUSION, COUNT> tr;
    tr.start(T, o, d, tmin, tmax, ray_flags, cull_mask, sbt_offset, st);
    int s;
    while ((s = tr.step(T, S, st, cnt)) == kRayActive) {
    }
    best = tr.best;
    return s == kRayHit ? 1 : (s == kRayOverflow ? -1 : 0);
}

// Hands queue entries to the idle lanes of a wavefront. The wave takes
// chunks of `chunk` entries from a global counter (one atomic per chunk)
// and serves its lanes from the chunk, so a lane that finished its ray gets
// a new one without waiting for the rest of the wave.
struct WaveFeeder {
    uint32_t next, end;
    uint32_t chunk;  // entries taken per atomic (multiple of 64)
    bool exhausted;
    __device__ inline void init(uint32_t chunk_) {
        next = end = 0;
        chunk = chunk_;
        exhausted = false;
    }
    // `idle` is the wave's ballot of lanes asking for work (wave-uniform);
    // returns this lane's entry or ~0u.
    __device__ inline uint32_t feed(uint64_t idle, bool want, uint32_t* counter, uint32_t n) {
        const int lane = threadIdx.x & 63;
        const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        uint32_t rank = (uint32_t)__popcll(idle & lt);
        uint32_t k = (uint32_t)__popcll(idle);
        uint32_t mine = ~0u;
        while (k > 0 && !exhausted) {
            if (next >= end) {
                uint32_t base = 0;
                if (lane == 0) base = atomicAdd(counter, chunk);
                base = (uint32_t)__shfl((int)base, 0);
                if (base >= n) {
                    exhausted = true;
                    break;
                }
                next = base;
                end = min(base + chunk, n);
            }
            uint32_t take = min(k, end - next);
            if (want && mine == ~0u) {
                if (rank < take) mine = next + rank;
                else rank -= take;
            }
            next += take;
            k -= take;
        }
        return mine;
    }
};

}  // namespace rtg
