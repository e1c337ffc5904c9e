/*
 * orc_math.h -- scalar FP32 math of the CPU oracle (test infrastructure).
 *
 * THIS FILE IS PART OF THE ORACLE: only tests, the smoke check and the CPU
 * baseline of bench.py may use it. It is never linked into the product.
 *
 * Every function states its operation order. Compiled with
 * -O2 -ffp-contract=off -fno-fast-math so that each written operation is one
 * IEEE-754 binary32 operation rounded to nearest even (x86-64 SSE), with
 * correctly rounded division and square root.
 *
 * Transcendentals are written here from their series, not taken from libm,
 * so that the GPU path can evaluate the same operations:
 *   sin/cos : Cody-Waite reduction by pi/2 (pi/2 split into a 12-bit,
 *             a 12-bit and a 24-bit part), then the Taylor series of sin to
 *             x^9 and of cos to x^10 on [-pi/4, pi/4] (truncation error
 *             < 1.8e-9 and < 1.2e-10, below half an ulp).
 *   log2    : x = m * 2^e with m in [sqrt(1/2), sqrt(2)), then
 *             log2(m) = 2*atanh(s)/ln 2 with s = (m-1)/(m+1), atanh series
 *             to s^11 (truncation < 2e-11).
 *   exp2    : y = n + f with n = rint(y), |f| <= 1/2, then e^(f*ln 2) by its
 *             Taylor series to degree 7 (truncation < 5.2e-9), times 2^n.
 *   pow     : exp2(y * log2(x)) for x > 0.
 */
#ifndef ORC_MATH_H
#define ORC_MATH_H

#include <math.h>
#include <stdint.h>
#include <string.h>

typedef struct { float x, y, z; } v3;
typedef struct { float x, y; } v2;

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

static inline v3 V3(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 v3_add(v3 a, v3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 v3_sub(v3 a, v3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 v3_mul(v3 a, v3 b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 v3_scale(v3 a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
static inline v3 v3_neg(v3 a) { return V3(-a.x, -a.y, -a.z); }
/* GLSL dot: (x*x' + y*y') + z*z' */
static inline float v3_dot(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline v3 v3_cross(v3 a, v3 b) {
    return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
/* normalize(v) = v * (1 / sqrt(dot(v, v)))  (glam length_recip; GLSL normalize) */
static inline v3 v3_normalize(v3 v) {
    float s = 1.0f / sqrtf(v3_dot(v, v));
    return v3_scale(v, s);
}
/* interpolate(a, b, c, bc) = (a*bc.x + b*bc.y) + c*bc.z  (hit_shader_common.glsl:79-81) */
static inline v3 v3_interp(v3 a, v3 b, v3 c, v3 bc) {
    return v3_add(v3_add(v3_scale(a, bc.x), v3_scale(b, bc.y)), v3_scale(c, bc.z));
}
static inline v2 v2_interp(v2 a, v2 b, v2 c, v3 bc) {
    v2 r;
    r.x = (a.x * bc.x + b.x * bc.y) + c.x * bc.z;
    r.y = (a.y * bc.x + b.y * bc.y) + c.y * bc.z;
    return r;
}
/* GLSL clamp(x, lo, hi) = min(max(x, lo), hi) */
static inline float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
static inline float fractf(float x) { return x - floorf(x); }

/* Column-major 4x4 (glam Mat4) times vec4: ((c0*x + c1*y) + c2*z) + c3*w */
static inline void mat4_mul_vec4(const float m[16], const float v[4], float out[4]) {
    for (int r = 0; r < 4; ++r)
        out[r] = ((m[0 + r] * v[0] + m[4 + r] * v[1]) + m[8 + r] * v[2]) + m[12 + r] * v[3];
}
/* Row-major 3x4 affine times point: ((m0*x + m1*y) + m2*z) + m3 */
static inline v3 m34_point(const float m[12], v3 p) {
    return V3(((m[0] * p.x + m[1] * p.y) + m[2] * p.z) + m[3],
              ((m[4] * p.x + m[5] * p.y) + m[6] * p.z) + m[7],
              ((m[8] * p.x + m[9] * p.y) + m[10] * p.z) + m[11]);
}
static inline v3 m34_dir(const float m[12], v3 d) {
    return V3((m[0] * d.x + m[1] * d.y) + m[2] * d.z,
              (m[4] * d.x + m[5] * d.y) + m[6] * d.z,
              (m[8] * d.x + m[9] * d.y) + m[10] * d.z);
}
/* Transposed upper 3x3 times n: mat3(gl_WorldToObject3x4EXT) * n, where
 * m is the row-major world-to-object matrix. */
static inline v3 m34_transpose_dir(const float m[12], v3 n) {
    return V3((m[0] * n.x + m[4] * n.y) + m[8] * n.z,
              (m[1] * n.x + m[5] * n.y) + m[9] * n.z,
              (m[2] * n.x + m[6] * n.y) + m[10] * n.z);
}

/* Inverse of a row-major 3x4 affine transform by cofactors. */
static inline void m34_inverse(const float m[12], float inv[12]) {
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
    for (int r = 0; r < 3; ++r)
        inv[4 * r + 3] = -((inv[4 * r + 0] * m[3] + inv[4 * r + 1] * m[7]) + inv[4 * r + 2] * m[11]);
}

#define ORC_PI_F 3.14159265358979323846f

/* sin and cos of x (intended range |x| < 2^13 * pi/2). */
static inline void orc_sincosf(float x, float* s_out, float* c_out) {
    const float two_over_pi = u2f(0x3f22f983u);
    const float p1 = u2f(0x3fc90000u); /* 1.5703125 */
    const float p2 = u2f(0x39fda000u); /* 4.837512970e-4 */
    const float p3 = u2f(0x33a22169u); /* 7.549790126e-8 */
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
static inline float orc_cosf(float x) { float s, c; orc_sincosf(x, &s, &c); return c; }
static inline float orc_sinf(float x) { float s, c; orc_sincosf(x, &s, &c); return s; }

/* log2 of x > 0 (finite); returns -inf for 0, NaN for x < 0, inf for inf. */
static inline float orc_log2f(float x) {
    if (!(x > 0.0f)) return x == 0.0f ? -INFINITY : NAN;
    if (x == INFINITY) return INFINITY;
    int e = 0;
    uint32_t u = f2u(x);
    if ((u >> 23) == 0) { /* subnormal: scale by 2^23 */
        x = x * 8388608.0f;
        u = f2u(x);
        e = -23;
    }
    e += (int)((u >> 23) & 0xffu) - 127;
    float m = u2f((u & 0x007fffffu) | 0x3f800000u); /* [1, 2) */
    if (m > u2f(0x3fb504f3u)) { /* > sqrt(2) */
        m = m * 0.5f;
        e += 1;
    }
    float s = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float series = 1.0f + s2 * (3.33333333e-1f + s2 * (2.0e-1f + s2 * (1.42857143e-1f + s2 * (1.11111111e-1f + s2 * 9.09090909e-2f))));
    float ln_m = (2.0f * s) * series;
    return (float)e + ln_m * u2f(0x3fb8aa3bu); /* log2(e) */
}

/* 2^y. */
static inline float orc_exp2f(float y) {
    if (y != y) return y;
    if (y >= 128.0f) return INFINITY;
    if (y < -150.0f) return 0.0f;
    float nf = rintf(y);
    float f = y - nf;
    float t = f * u2f(0x3f317218u); /* ln 2 */
    float p = 1.0f + t * (1.0f + t * (0.5f + t * (1.66666667e-1f + t * (4.16666667e-2f + t * (8.33333333e-3f + t * (1.38888889e-3f + t * 1.98412698e-4f))))));
    int n = (int)nf;
    /* scale by 2^n in two exact steps so that subnormal results round once */
    int n1 = n / 2, n2 = n - n1;
    float s1 = u2f((uint32_t)(n1 + 127) << 23);
    float s2 = u2f((uint32_t)(n2 + 127) << 23);
    return (p * s1) * s2;
}

/* pow(x, y) for x >= 0 as exp2(y * log2(x)); pow(0, y>0) = 0. */
static inline float orc_powf(float x, float y) {
    if (x == 0.0f) return y > 0.0f ? 0.0f : (y == 0.0f ? 1.0f : INFINITY);
    return orc_exp2f(y * orc_log2f(x));
}

/* pow(x, 5.0) of pbr.glsl:69-76 as ((x*x)*(x*x))*x. */
static inline float orc_pow5f(float x) {
    float x2 = x * x;
    return (x2 * x2) * x;
}

#endif /* ORC_MATH_H */
