/* Oracle (test infrastructure only): f32 vector / quaternion / 3x3 matrix arithmetic with the
 * operation order of glam 0.27 as the reference uses it on x86_64 (Vec3 scalar; Quat, Vec4 and
 * Mat3A on the SSE2 path, because `scalar-math` is off: Cargo.lock:998-1001, PRECISION_SPEC
 * 153-168).  Every function states the order it follows.  Compiled with -ffp-contract=off.
 */
#ifndef OR_MATH_H
#define OR_MATH_H
#include <math.h>
#include <stdint.h>
#include <string.h>

typedef struct { float x, y, z; } V3;
typedef struct { float x, y, z, w; } Q4;
typedef struct { V3 c0, c1, c2; } M3; /* column-major, like glam's Mat3A */

static inline V3 v3(float x, float y, float z) { V3 r = {x, y, z}; return r; }
static inline V3 v3s(float s) { return v3(s, s, s); }
static inline V3 vadd(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 vsub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 vmul(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline V3 vdivs(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
static inline V3 vneg(V3 a) { return v3(-a.x, -a.y, -a.z); }
static inline V3 vabs(V3 a) { return v3(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }
static inline V3 vmin(V3 a, V3 b) { return v3(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
static inline V3 vmax(V3 a, V3 b) { return v3(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)); }
/* glam Vec3::dot: (x*x' + y*y') + z*z' */
static inline float vdot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
/* glam Vec3::cross (and Vec3A's SSE2 shuffle form, same products and differences) */
static inline V3 vcross(V3 a, V3 b) {
    return v3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
static inline float vlen2(V3 a) { return vdot(a, a); }
static inline float vlen(V3 a) { return sqrtf(vdot(a, a)); }
static inline int vfinite(V3 a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }
static inline int veq(V3 a, V3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
static inline int viszero(V3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
/* glam Vec3::normalize: self * (1 / length) */
static inline V3 vnormalize(V3 a) { float r = 1.0f / vlen(a); return vmul(a, r); }
/* glam Vec3::normalize_or_zero: rcp = 1/length; finite and > 0 ? self * rcp : 0 */
static inline V3 vnormalize_or_zero(V3 a) {
    float r = 1.0f / vlen(a);
    if (isfinite(r) && r > 0.0f) return vmul(a, r);
    return v3s(0.0f);
}
/* glam Vec3::clamp_length_max: |v|^2 > m^2 ? m * (v / sqrt(|v|^2)) : v */
static inline V3 vclamp_len_max(V3 v, float m) {
    float l2 = vlen2(v);
    if (l2 > m * m) { V3 u = vdivs(v, sqrtf(l2)); return v3(m * u.x, m * u.y, m * u.z); }
    return v;
}
static inline V3 vclamp(V3 v, V3 lo, V3 hi) { return vmin(vmax(v, lo), hi); }

/* ---- Quat (SSE2 Vec4 semantics) ---- */
static inline Q4 q4(float x, float y, float z, float w) { Q4 q = {x, y, z, w}; return q; }
static inline Q4 qident(void) { return q4(0, 0, 0, 1); }
static inline int qfinite(Q4 q) { return isfinite(q.x) && isfinite(q.y) && isfinite(q.z) && isfinite(q.w); }
static inline Q4 qneg(Q4 q) { return q4(-q.x, -q.y, -q.z, -q.w); }
static inline Q4 qconj(Q4 q) { return q4(-q.x, -q.y, -q.z, q.w); }
static inline Q4 qscale(Q4 q, float s) { return q4(q.x * s, q.y * s, q.z * s, q.w * s); }
/* SSE2 dot4: (x*x' + z*z') + (y*y' + w*w') */
static inline float qdot(Q4 a, Q4 b) { return (a.x * b.x + a.z * b.z) + (a.y * b.y + a.w * b.w); }
/* SSE2 Vec4::normalize: self / sqrt(dot4) (a division, unlike scalar Vec3) */
static inline Q4 qnormalize(Q4 q) {
    float l = sqrtf(qdot(q, q));
    return q4(q.x / l, q.y / l, q.z / l, q.w / l);
}
static inline int qeq(Q4 a, Q4 b) { return a.x == b.x && a.y == b.y && a.z == b.z && a.w == b.w; }
/* SSE2 Quat::mul_quat (rtm quat_mul form): (lw*r + lx*r_wzyx*s) + (ly*r_zwxy*s + lz*r_yxwz*s) */
static inline Q4 qmul(Q4 l, Q4 r) {
    float ax = l.w * r.x, ay = l.w * r.y, az = l.w * r.z, aw = l.w * r.w;
    float bx = l.x * r.w, by = -(l.x * r.z), bz = l.x * r.y, bw = -(l.x * r.x);
    float cx = l.y * r.z, cy = l.y * r.w, cz = -(l.y * r.x), cw = -(l.y * r.y);
    float dx = -(l.z * r.y), dy = l.z * r.x, dz = l.z * r.w, dw = -(l.z * r.z);
    return q4((ax + bx) + (cx + dx), (ay + by) + (cy + dy), (az + bz) + (cz + dz), (aw + bw) + (cw + dw));
}
/* Quat::mul_vec3: v*(w*w - b.b) + b*(v.b * 2) + (b x v)*(w*2), dot3 as (x+y)+z */
static inline V3 qrot(Q4 q, V3 v) {
    V3 b = v3(q.x, q.y, q.z);
    float w = q.w;
    float b2 = vdot(b, b);
    float s0 = w * w - b2;
    float s1 = vdot(v, b) * 2.0f;
    float s2 = w * 2.0f;
    V3 c = vcross(b, v);
    return vadd(vadd(vmul(v, s0), vmul(b, s1)), vmul(c, s2));
}

/* ---- Mat3A ---- */
static inline M3 m3cols(V3 a, V3 b, V3 c) { M3 m = {a, b, c}; return m; }
static inline M3 m3zero(void) { return m3cols(v3s(0), v3s(0), v3s(0)); }
static inline M3 m3diag(V3 d) { return m3cols(v3(d.x, 0, 0), v3(0, d.y, 0), v3(0, 0, d.z)); }
static inline int m3iszero(M3 m) { return viszero(m.c0) && viszero(m.c1) && viszero(m.c2); }
static inline int m3finite(M3 m) { return vfinite(m.c0) && vfinite(m.c1) && vfinite(m.c2); }
/* Mat3A::mul_vec3a: (c0*x + c1*y) + c2*z */
static inline V3 m3v(M3 m, V3 v) { return vadd(vadd(vmul(m.c0, v.x), vmul(m.c1, v.y)), vmul(m.c2, v.z)); }
static inline M3 m3mul(M3 a, M3 b) { return m3cols(m3v(a, b.c0), m3v(a, b.c1), m3v(a, b.c2)); }
static inline M3 m3t(M3 m) {
    return m3cols(v3(m.c0.x, m.c1.x, m.c2.x), v3(m.c0.y, m.c1.y, m.c2.y), v3(m.c0.z, m.c1.z, m.c2.z));
}
static inline M3 m3sub(M3 a, M3 b) { return m3cols(vsub(a.c0, b.c0), vsub(a.c1, b.c1), vsub(a.c2, b.c2)); }
/* Mat3A::determinant: z . (x cross y) */
static inline float m3det(M3 m) { return vdot(m.c2, vcross(m.c0, m.c1)); }
/* Mat3A::inverse: cofactor columns times 1/det, transposed */
static inline M3 m3inv(M3 m) {
    V3 t0 = vcross(m.c1, m.c2), t1 = vcross(m.c2, m.c0), t2 = vcross(m.c0, m.c1);
    float det = vdot(m.c2, t2);
    float id = 1.0f / det;
    return m3t(m3cols(vmul(t0, id), vmul(t1, id), vmul(t2, id)));
}
/* Mat3A::from_quat */
static inline M3 m3quat(Q4 q) {
    float x2 = q.x + q.x, y2 = q.y + q.y, z2 = q.z + q.z;
    float xx = q.x * x2, xy = q.x * y2, xz = q.x * z2;
    float yy = q.y * y2, yz = q.y * z2, zz = q.z * z2;
    float wx = q.w * x2, wy = q.w * y2, wz = q.w * z2;
    return m3cols(v3(1.0f - (yy + zz), xy + wz, xz - wy), v3(xy - wz, 1.0f - (xx + zz), yz + wx),
                  v3(xz + wy, yz - wx, 1.0f - (xx + yy)));
}

/* sin/cos/exp evaluated in double and rounded once to f32 (Taylor series on |r| <= pi/4 after
 * Cody-Waite reduction by pi/2; exp by 2^k * Taylor(r), |r| <= ln2/2).  The reference calls the
 * libm 0.2.15 sincosf / platform expf, which agree with a once-rounded evaluation except in rare
 * last-ulp cases; the device library uses the same definition, so oracle and device agree bit
 * for bit.  Pinned against glibc within 1 ulp (tests/test_oracle_core.py). */
static inline double or_sin_k(double r) {
    double r2 = r * r, t = 1.0 / 51090942171709440000.0;
    t = t * r2 - 1.0 / 121645100408832000.0; t = t * r2 + 1.0 / 355687428096000.0;
    t = t * r2 - 1.0 / 1307674368000.0;      t = t * r2 + 1.0 / 6227020800.0;
    t = t * r2 - 1.0 / 39916800.0;           t = t * r2 + 1.0 / 362880.0;
    t = t * r2 - 1.0 / 5040.0;               t = t * r2 + 1.0 / 120.0;
    t = t * r2 - 1.0 / 6.0;
    return r + r * (r2 * t);
}
static inline double or_cos_k(double r) {
    double r2 = r * r, t = 1.0 / 2432902008176640000.0;
    t = t * r2 - 1.0 / 6402373705728000.0; t = t * r2 + 1.0 / 20922789888000.0;
    t = t * r2 - 1.0 / 87178291200.0;      t = t * r2 + 1.0 / 479001600.0;
    t = t * r2 - 1.0 / 3628800.0;          t = t * r2 + 1.0 / 40320.0;
    t = t * r2 - 1.0 / 720.0;              t = t * r2 + 1.0 / 24.0;
    t = t * r2 - 0.5;
    return 1.0 + r2 * t;
}
static inline void or_sincos_d(float xf, float *s, float *c) {
    if (!isfinite(xf)) { *s = NAN; *c = NAN; return; }
    double x = (double)xf;
    double k = nearbyint(x * 0.63661977236758134308);
    double r = (x - k * 1.57079632673412561417e+00) - k * 6.07710050650619224932e-11;
    long q = (long)k & 3;
    double sr = or_sin_k(r), cr = or_cos_k(r), sv, cv;
    if (q == 0) { sv = sr; cv = cr; } else if (q == 1) { sv = cr; cv = -sr; }
    else if (q == 2) { sv = -sr; cv = -cr; } else { sv = -cr; cv = sr; }
    *s = (float)sv; *c = (float)cv;
}
static inline float or_expf_d(float xf) {
    if (xf != xf) return xf;
    if (xf > 88.8f) return INFINITY;
    if (xf < -104.0f) return 0.0f;
    double x = (double)xf;
    double k = nearbyint(x * 1.44269504088896338700);
    double r = (x - k * 6.93147180369123816490e-01) - k * 1.90821492927058770002e-10;
    double t = 1.0 / 87178291200.0;
    t = t * r + 1.0 / 6227020800.0; t = t * r + 1.0 / 479001600.0; t = t * r + 1.0 / 39916800.0;
    t = t * r + 1.0 / 3628800.0;    t = t * r + 1.0 / 362880.0;    t = t * r + 1.0 / 40320.0;
    t = t * r + 1.0 / 5040.0;       t = t * r + 1.0 / 720.0;       t = t * r + 1.0 / 120.0;
    t = t * r + 1.0 / 24.0;         t = t * r + 1.0 / 6.0;         t = t * r + 0.5;
    t = t * r + 1.0;                t = t * r + 1.0;
    return (float)ldexp(t, (int)k);
}

#endif
