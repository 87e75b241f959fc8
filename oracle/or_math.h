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

/* exp, and sin/cos beyond libm's medium range, evaluated in double and rounded once to f32
 * (Taylor series on |r| <= pi/4 after Cody-Waite reduction by pi/2; exp by 2^k * Taylor(r),
 * |r| <= ln2/2).  The reference's Stribeck term calls the platform expf (solver.rs:430), which
 * no golden pins; it agrees with a once-rounded evaluation except in rare last-ulp cases, and it
 * drops out bit for bit when mu_s == mu_k.  The device library uses the same definitions, so
 * oracle and device agree bit for bit.  Pinned against glibc within 1 ulp
 * (tests/test_oracle_core.py). */
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
/* libm 0.2.15 sincosf (Cargo.lock:1476-1479; glam 0.27's Quat::from_axis_angle /
 * from_rotation_* call it through its `libm` feature).  That crate ports musl's sincosf.c, whose
 * kernels are FreeBSD msun's k_sinf.c / k_cosf.c: fixed polynomials in double, rounded once to
 * f32, after reduction by multiples of pi/2 (rem_pio2f.c for |x| > 9pi/4).  Restated from that
 * published algorithm; pinned by the distance-joint goldens (distance_bits_test.rs:47-63), which
 * run every gyroscopic turn through it.  |x| >= 2^28 pi/2 (Payne-Hanek reduction in libm) falls
 * back to the once-rounded Taylor evaluation below: unpinned, and unreachable from a step. */
static inline float or_k_sinf(double x) {
    const double S1 = -0x15555554cbac77.0p-55, S2 = 0x111110896efbb2.0p-59;
    const double S3 = -0x1a00f9e2cae774.0p-65, S4 = 0x16cd878c3b46a7.0p-71;
    double z = x * x, w = z * z, r = S3 + z * S4, s = z * x;
    return (float)((x + s * (S1 + z * S2)) + s * w * r);
}
static inline float or_k_cosf(double x) {
    const double C0 = -0x1ffffffd0c5e81.0p-54, C1 = 0x155553e1053a42.0p-57;
    const double C2 = -0x16c087e80f1e27.0p-62, C3 = 0x199342e0ee5069.0p-68;
    double z = x * x, w = z * z, r = C2 + z * C3;
    return (float)(((1.0 + z * C0) + w * C1) + (w * z) * r);
}
static inline void or_sincos_d(float xf, float *s, float *c);
static inline void or_sincos_libm(float x, float *s, float *c) {
    const double P1 = 1.0 * 1.57079632679489661923, P2 = 2.0 * 1.57079632679489661923;
    const double P3 = 3.0 * 1.57079632679489661923, P4 = 4.0 * 1.57079632679489661923;
    uint32_t ix; memcpy(&ix, &x, 4);
    int sign = (int)(ix >> 31);
    ix &= 0x7fffffffu;
    double xd = (double)x;
    if (ix <= 0x3f490fdau) {                      /* |x| <= pi/4 */
        if (ix < 0x39800000u) { *s = x; *c = 1.0f; return; } /* |x| < 2^-12 */
        *s = or_k_sinf(xd); *c = or_k_cosf(xd); return;
    }
    if (ix <= 0x407b53d1u) {                      /* |x| <= 5pi/4 */
        if (ix <= 0x4016cbe3u) {                  /* |x| <= 3pi/4 */
            if (sign) { *s = -or_k_cosf(xd + P1); *c = or_k_sinf(xd + P1); }
            else { *s = or_k_cosf(P1 - xd); *c = or_k_sinf(P1 - xd); }
            return;
        }
        double y = sign ? xd + P2 : xd - P2;
        *s = -or_k_sinf(y); *c = -or_k_cosf(y); return;
    }
    if (ix <= 0x40e231d5u) {                      /* |x| <= 9pi/4 */
        if (ix <= 0x40afeddfu) {                  /* |x| <= 7pi/4 */
            if (sign) { *s = or_k_cosf(xd + P3); *c = -or_k_sinf(xd + P3); }
            else { *s = -or_k_cosf(xd - P3); *c = or_k_sinf(xd - P3); }
            return;
        }
        double y = sign ? xd + P4 : xd - P4;
        *s = or_k_sinf(y); *c = or_k_cosf(y); return;
    }
    if (ix >= 0x7f800000u) { *s = *c = x - x; return; }
    if (ix >= 0x4dc90fdbu) { or_sincos_d(x, s, c); return; }
    /* rem_pio2f, medium size: n = rint(x 2/pi) with the 1.5 2^52 trick, y = x - n (pio2_1 + pio2_1t) */
    const double TOINT = 1.5 / 2.220446049250313080847e-16;
    const double INV_PIO2 = 6.36619772367581382433e-01;
    const double PIO2_1 = 1.57079631090164184570e+00, PIO2_1T = 1.58932547735281966916e-08;
    double fn = (xd * INV_PIO2 + TOINT) - TOINT;
    int n = (int)fn;
    double y = xd - fn * PIO2_1 - fn * PIO2_1T;
    float sv = or_k_sinf(y), cv = or_k_cosf(y);
    switch (n & 3) {
        case 0: *s = sv; *c = cv; break;
        case 1: *s = cv; *c = -sv; break;
        case 2: *s = -sv; *c = -cv; break;
        default: *s = -cv; *c = sv; break;
    }
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

/* atan2 for ordering a friction patch's points by angle (solver.rs:233-236 sorts by
 * sort_key6(y.atan2(x))): a once-rounded double evaluation, the same operations on the device.
 * Only the order of the keys matters there, and a correctly rounded angle orders as any accurate
 * f32 atan2 does unless two points' angles lie within a few 1e-7 rad.  Finite inputs.
 * atan on [0, 1]: tan(pi/8) reduction, then the odd series to t^49 in Horner form (|t| <= 0.4143,
 * truncation below 1e-20). */
static inline double or_atan_unit(double t) {
    double base = 0.0;
    if (t > 0.41421356237309503) { t = (t - 1.0) / (t + 1.0); base = 0.78539816339744828; }
    const double t2 = t * t;
    double p = 1.0 / 49.0;
    for (int k = 23; k >= 0; k--) p = 1.0 / (double)(2 * k + 1) - t2 * p;
    return base + t * p;
}
static inline float or_atan2f(float y, float x) {
    if (x != x || y != y) return x + y;
    const double ay = fabs((double)y), ax = fabs((double)x);
    double a;
    if (ay == 0.0 && ax == 0.0) a = signbit(x) ? 3.141592653589793 : 0.0;
    else {
        a = ay <= ax ? or_atan_unit(ay / ax) : 1.5707963267948966 - or_atan_unit(ax / ay);
        if (signbit(x)) a = 3.141592653589793 - a;
    }
    const float r = (float)a;
    return signbit(y) ? -r : r;
}

#endif
