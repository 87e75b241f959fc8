// Device f32 arithmetic for the step path, in the operation order of glam 0.27 as the reference
// runs it on x86_64 (Vec3 scalar; Quat / Vec4 / Mat3A on the SSE2 path).  Compiled with
// -ffp-contract=off and IEEE division / square root so every operation rounds exactly as
// written.  sin/cos follow libm's sincosf (double-precision kernels rounded once to f32); exp is
// a once-rounded double evaluation (the reference calls the platform expf, which agrees with it
// except in rare ties).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define RP_HD __host__ __device__ __forceinline__

struct V3 { float x, y, z; };
struct Q4 { float x, y, z, w; };
struct M3 { V3 c0, c1, c2; };  // column-major

RP_HD V3 mk3(float x, float y, float z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
RP_HD V3 splat3(float s) { return mk3(s, s, s); }
RP_HD V3 operator+(V3 a, V3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
RP_HD V3 operator-(V3 a, V3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
RP_HD V3 operator-(V3 a) { return mk3(-a.x, -a.y, -a.z); }
RP_HD V3 operator*(V3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
RP_HD V3 operator/(V3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
RP_HD V3 v3abs(V3 a) { return mk3(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }
RP_HD V3 v3min(V3 a, V3 b) { return mk3(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
RP_HD V3 v3max(V3 a, V3 b) { return mk3(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)); }
// Vec3::dot: (x*x' + y*y') + z*z'
RP_HD float dot3(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
RP_HD V3 cross3(V3 a, V3 b) { return mk3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y); }
RP_HD float len2(V3 a) { return dot3(a, a); }
RP_HD float len3(V3 a) { return sqrtf(dot3(a, a)); }
RP_HD bool fin3(V3 a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }
RP_HD bool eq3(V3 a, V3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
RP_HD bool zero3(V3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
// Vec3::normalize: self * (1 / length)
RP_HD V3 normalize3(V3 a) { float r = 1.0f / len3(a); return a * r; }
RP_HD V3 normalize_or_zero3(V3 a) {
    float r = 1.0f / len3(a);
    return (isfinite(r) && r > 0.0f) ? a * r : splat3(0.0f);
}
// Vec3::clamp_length_max: m * (v / |v|) when |v|^2 > m^2
RP_HD V3 clamp_len_max(V3 v, float m) {
    float l2 = len2(v);
    if (l2 > m * m) { V3 u = v / sqrtf(l2); return mk3(m * u.x, m * u.y, m * u.z); }
    return v;
}
RP_HD V3 clamp3(V3 v, V3 lo, V3 hi) { return v3min(v3max(v, lo), hi); }
RP_HD float clampf_rs(float x, float lo, float hi) { if (x < lo) x = lo; if (x > hi) x = hi; return x; }

RP_HD Q4 mkq(float x, float y, float z, float w) { Q4 q; q.x = x; q.y = y; q.z = z; q.w = w; return q; }
RP_HD Q4 qid() { return mkq(0.f, 0.f, 0.f, 1.f); }
RP_HD bool finq(Q4 q) { return isfinite(q.x) && isfinite(q.y) && isfinite(q.z) && isfinite(q.w); }
RP_HD Q4 qneg(Q4 q) { return mkq(-q.x, -q.y, -q.z, -q.w); }
RP_HD Q4 qconj(Q4 q) { return mkq(-q.x, -q.y, -q.z, q.w); }
RP_HD bool eqq(Q4 a, Q4 b) { return a.x == b.x && a.y == b.y && a.z == b.z && a.w == b.w; }
// SSE2 dot4: (x*x' + z*z') + (y*y' + w*w'); SSE2 Vec4::normalize divides by sqrt(dot4)
RP_HD float qdot(Q4 a, Q4 b) { return (a.x * b.x + a.z * b.z) + (a.y * b.y + a.w * b.w); }
RP_HD Q4 qnormalize(Q4 q) { float l = sqrtf(qdot(q, q)); return mkq(q.x / l, q.y / l, q.z / l, q.w / l); }
// SSE2 Quat::mul_quat: (lw*r + lx*r_wzyx*s) + (ly*r_zwxy*s + lz*r_yxwz*s)
RP_HD Q4 qmul(Q4 l, Q4 r) {
    float ax = l.w * r.x, ay = l.w * r.y, az = l.w * r.z, aw = l.w * r.w;
    float bx = l.x * r.w, by = -(l.x * r.z), bz = l.x * r.y, bw = -(l.x * r.x);
    float cx = l.y * r.z, cy = l.y * r.w, cz = -(l.y * r.x), cw = -(l.y * r.y);
    float dx = -(l.z * r.y), dy = l.z * r.x, dz = l.z * r.w, dw = -(l.z * r.z);
    return mkq((ax + bx) + (cx + dx), (ay + by) + (cy + dy), (az + bz) + (cz + dz), (aw + bw) + (cw + dw));
}
// Quat::mul_vec3: v*(w*w - b.b) + b*((v.b)*2) + (b x v)*(w*2)
RP_HD V3 qrot(Q4 q, V3 v) {
    V3 b = mk3(q.x, q.y, q.z);
    float s0 = q.w * q.w - dot3(b, b);
    float s1 = dot3(v, b) * 2.0f;
    float s2 = q.w * 2.0f;
    return (v * s0 + b * s1) + cross3(b, v) * s2;
}

RP_HD M3 mkm(V3 a, V3 b, V3 c) { M3 m; m.c0 = a; m.c1 = b; m.c2 = c; return m; }
RP_HD M3 mzero() { return mkm(splat3(0.f), splat3(0.f), splat3(0.f)); }
RP_HD M3 mdiag(V3 d) { return mkm(mk3(d.x, 0.f, 0.f), mk3(0.f, d.y, 0.f), mk3(0.f, 0.f, d.z)); }
RP_HD bool mfinite(M3 m) { return fin3(m.c0) && fin3(m.c1) && fin3(m.c2); }
// Mat3A::mul_vec3a: (c0*x + c1*y) + c2*z
RP_HD V3 mv(M3 m, V3 v) { return (m.c0 * v.x + m.c1 * v.y) + m.c2 * v.z; }
RP_HD M3 mmul(M3 a, M3 b) { return mkm(mv(a, b.c0), mv(a, b.c1), mv(a, b.c2)); }
RP_HD M3 mtrans(M3 m) {
    return mkm(mk3(m.c0.x, m.c1.x, m.c2.x), mk3(m.c0.y, m.c1.y, m.c2.y), mk3(m.c0.z, m.c1.z, m.c2.z));
}
RP_HD M3 msub(M3 a, M3 b) { return mkm(a.c0 - b.c0, a.c1 - b.c1, a.c2 - b.c2); }
RP_HD float mdet(M3 m) { return dot3(m.c2, cross3(m.c0, m.c1)); }
RP_HD M3 minverse(M3 m) {
    V3 t0 = cross3(m.c1, m.c2), t1 = cross3(m.c2, m.c0), t2 = cross3(m.c0, m.c1);
    float id = 1.0f / dot3(m.c2, t2);
    return mtrans(mkm(t0 * id, t1 * id, t2 * id));
}
RP_HD M3 mquat(Q4 q) {
    float x2 = q.x + q.x, y2 = q.y + q.y, z2 = q.z + q.z;
    float xx = q.x * x2, xy = q.x * y2, xz = q.x * z2;
    float yy = q.y * y2, yz = q.y * z2, zz = q.z * z2;
    float wx = q.w * x2, wy = q.w * y2, wz = q.w * z2;
    return mkm(mk3(1.0f - (yy + zz), xy + wz, xz - wy), mk3(xy - wz, 1.0f - (xx + zz), yz + wx),
               mk3(xz + wy, yz - wx, 1.0f - (xx + yy)));
}

// ---- transcendental functions, evaluated in double and rounded once ----
// sin / cos on |r| <= pi/4 by their Taylor series to x^21 / x^20 (truncation < 1e-19),
// after Cody-Waite reduction by pi/2 (two-term split of pi/2).
RP_HD double rp_sin_kernel(double r) {
    double r2 = r * r;
    double t = 1.0 / 51090942171709440000.0;  // 1/21!
    t = t * r2 - 1.0 / 121645100408832000.0;  // 1/19!
    t = t * r2 + 1.0 / 355687428096000.0;     // 1/17!
    t = t * r2 - 1.0 / 1307674368000.0;       // 1/15!
    t = t * r2 + 1.0 / 6227020800.0;          // 1/13!
    t = t * r2 - 1.0 / 39916800.0;            // 1/11!
    t = t * r2 + 1.0 / 362880.0;              // 1/9!
    t = t * r2 - 1.0 / 5040.0;                // 1/7!
    t = t * r2 + 1.0 / 120.0;                 // 1/5!
    t = t * r2 - 1.0 / 6.0;                   // 1/3!
    return r + r * (r2 * t);
}
RP_HD double rp_cos_kernel(double r) {
    double r2 = r * r;
    double t = 1.0 / 2432902008176640000.0;   // 1/20!
    t = t * r2 - 1.0 / 6402373705728000.0;    // 1/18!
    t = t * r2 + 1.0 / 20922789888000.0;      // 1/16!
    t = t * r2 - 1.0 / 87178291200.0;         // 1/14!
    t = t * r2 + 1.0 / 479001600.0;           // 1/12!
    t = t * r2 - 1.0 / 3628800.0;             // 1/10!
    t = t * r2 + 1.0 / 40320.0;               // 1/8!
    t = t * r2 - 1.0 / 720.0;                 // 1/6!
    t = t * r2 + 1.0 / 24.0;                  // 1/4!
    t = t * r2 - 0.5;                         // 1/2!
    return 1.0 + r2 * t;
}
RP_HD void rp_sincosf_taylor(float xf, float *s, float *c) {
    if (!isfinite(xf)) { *s = NAN; *c = NAN; return; }
    double x = (double)xf;
    const double PIO2_HI = 1.57079632673412561417e+00;  // first 33 bits of pi/2
    const double PIO2_LO = 6.07710050650619224932e-11;  // pi/2 - PIO2_HI
    double k = nearbyint(x * 0.63661977236758134308);   // 2/pi
    double r = (x - k * PIO2_HI) - k * PIO2_LO;
    long q = (long)k & 3;
    double sr = rp_sin_kernel(r), cr = rp_cos_kernel(r);
    double sv, cv;
    switch (q) {
        case 0: sv = sr; cv = cr; break;
        case 1: sv = cr; cv = -sr; break;
        case 2: sv = -sr; cv = -cr; break;
        default: sv = -cr; cv = sr; break;
    }
    *s = (float)sv; *c = (float)cv;
}
// sincosf of libm 0.2.15 (Cargo.lock:1476-1479), which glam 0.27's Quat::from_axis_angle calls
// through its `libm` feature: musl's sincosf.c with FreeBSD msun's k_sinf / k_cosf kernels, fixed
// polynomials in double rounded once to f32, after reduction by multiples of pi/2 (rem_pio2f's
// medium-size form beyond 9pi/4).  Arguments of 2^28 pi/2 and more take the Taylor form above.
RP_HD float rp_k_sinf(double x) {
    const double S1 = -0x15555554cbac77.0p-55, S2 = 0x111110896efbb2.0p-59;
    const double S3 = -0x1a00f9e2cae774.0p-65, S4 = 0x16cd878c3b46a7.0p-71;
    double z = x * x, w = z * z, r = S3 + z * S4, s = z * x;
    return (float)((x + s * (S1 + z * S2)) + s * w * r);
}
RP_HD float rp_k_cosf(double x) {
    const double C0 = -0x1ffffffd0c5e81.0p-54, C1 = 0x155553e1053a42.0p-57;
    const double C2 = -0x16c087e80f1e27.0p-62, C3 = 0x199342e0ee5069.0p-68;
    double z = x * x, w = z * z, r = C2 + z * C3;
    return (float)(((1.0 + z * C0) + w * C1) + (w * z) * r);
}
RP_HD void rp_sincosf(float x, float *s, float *c) {
    const double P1 = 1.0 * 1.57079632679489661923, P2 = 2.0 * 1.57079632679489661923;
    const double P3 = 3.0 * 1.57079632679489661923, P4 = 4.0 * 1.57079632679489661923;
    uint32_t ix;
    __builtin_memcpy(&ix, &x, 4);
    bool neg = (ix >> 31) != 0;
    ix &= 0x7fffffffu;
    double xd = (double)x;
    if (ix <= 0x3f490fdau) {                      // |x| <= pi/4
        if (ix < 0x39800000u) { *s = x; *c = 1.0f; return; }  // |x| < 2^-12
        *s = rp_k_sinf(xd); *c = rp_k_cosf(xd); return;
    }
    if (ix <= 0x407b53d1u) {                      // |x| <= 5pi/4
        if (ix <= 0x4016cbe3u) {                  // |x| <= 3pi/4
            if (neg) { *s = -rp_k_cosf(xd + P1); *c = rp_k_sinf(xd + P1); }
            else { *s = rp_k_cosf(P1 - xd); *c = rp_k_sinf(P1 - xd); }
            return;
        }
        double y = neg ? xd + P2 : xd - P2;
        *s = -rp_k_sinf(y); *c = -rp_k_cosf(y); return;
    }
    if (ix <= 0x40e231d5u) {                      // |x| <= 9pi/4
        if (ix <= 0x40afeddfu) {                  // |x| <= 7pi/4
            if (neg) { *s = rp_k_cosf(xd + P3); *c = -rp_k_sinf(xd + P3); }
            else { *s = -rp_k_cosf(xd - P3); *c = rp_k_sinf(xd - P3); }
            return;
        }
        double y = neg ? xd + P4 : xd - P4;
        *s = rp_k_sinf(y); *c = rp_k_cosf(y); return;
    }
    if (ix >= 0x7f800000u) { *s = *c = x - x; return; }
    if (ix >= 0x4dc90fdbu) { rp_sincosf_taylor(x, s, c); return; }
    const double TOINT = 1.5 / 2.220446049250313080847e-16;   // 1.5 / DBL_EPSILON
    const double INV_PIO2 = 6.36619772367581382433e-01;
    const double PIO2_1 = 1.57079631090164184570e+00, PIO2_1T = 1.58932547735281966916e-08;
    double fn = (xd * INV_PIO2 + TOINT) - TOINT;
    int n = (int)fn;
    double y = xd - fn * PIO2_1 - fn * PIO2_1T;
    float sv = rp_k_sinf(y), cv = rp_k_cosf(y);
    switch (n & 3) {
        case 0: *s = sv; *c = cv; break;
        case 1: *s = cv; *c = -sv; break;
        case 2: *s = -sv; *c = -cv; break;
        default: *s = -cv; *c = sv; break;
    }
}
// exp(x) = 2^k exp(r), |r| <= ln2/2, Taylor to r^14
RP_HD float rp_expf(float xf) {
    if (xf != xf) return xf;
    if (xf > 88.8f) return INFINITY;
    if (xf < -104.0f) return 0.0f;
    double x = (double)xf;
    const double LN2_HI = 6.93147180369123816490e-01, LN2_LO = 1.90821492927058770002e-10;
    double k = nearbyint(x * 1.44269504088896338700);
    double r = (x - k * LN2_HI) - k * LN2_LO;
    double t = 1.0 / 87178291200.0;  // 1/14!
    t = t * r + 1.0 / 6227020800.0;
    t = t * r + 1.0 / 479001600.0;
    t = t * r + 1.0 / 39916800.0;
    t = t * r + 1.0 / 3628800.0;
    t = t * r + 1.0 / 362880.0;
    t = t * r + 1.0 / 40320.0;
    t = t * r + 1.0 / 5040.0;
    t = t * r + 1.0 / 720.0;
    t = t * r + 1.0 / 120.0;
    t = t * r + 1.0 / 24.0;
    t = t * r + 1.0 / 6.0;
    t = t * r + 0.5;
    t = t * r + 1.0;
    t = t * r + 1.0;
    return (float)ldexp(t, (int)k);
}

// ---- quantisers and keys ----
// det.rs:8-13 q6 / determinism.rs:57-65 qf32: (x / 1e-6).round() * 1e-6
RP_HD float q6div(float x) { return isfinite(x) ? roundf(x / 1.0e-6f) * 1.0e-6f : 0.0f; }
RP_HD float qf32(float x) { return isfinite(x) ? roundf(x / 1.0e-6f) * 1.0e-6f : NAN; }
// materials.rs:319-321 / gravity.rs:43-49: (x * 1e6).round() * 1e-6
RP_HD float q6mul(float x) { return roundf(x * 1.0e6f) * 1.0e-6f; }
// det.rs:21-25 sort_key6
RP_HD int64_t sort_key6(float x) {
    if (x != x) return INT64_MAX;
    double k = round((double)x * 1.0e6);
    if (k >= 9.2233720368547758e18) return INT64_MAX;
    if (k <= -9.2233720368547758e18) return INT64_MIN;
    return (int64_t)k;
}
RP_HD float finf(float x) { return isfinite(x) ? x : 0.0f; }
RP_HD V3 fin_v(V3 v) { return mk3(finf(v.x), finf(v.y), finf(v.z)); }
// dynamics.rs:14-22 fin_q: normalise, identity if non-finite, w >= 0
RP_HD Q4 fin_q(Q4 q) {
    if (!finq(q)) return qid();
    Q4 n = qnormalize(q);
    if (!finq(n)) return qid();
    return n.w < 0.0f ? qneg(n) : n;
}
// y.atan2(x) for ordering a friction patch's points by angle (solver.rs:233-236): a once-rounded
// double evaluation (atan on [0, 1] by the tan(pi/8) reduction and the odd series to t^49 in
// Horner form), the oracle's operations; finite inputs.
RP_HD double rp_atan_unit(double t) {
    double base = 0.0;
    if (t > 0.41421356237309503) { t = (t - 1.0) / (t + 1.0); base = 0.78539816339744828; }
    const double t2 = t * t;
    double p = 1.0 / 49.0;
    for (int k = 23; k >= 0; k--) p = 1.0 / (double)(2 * k + 1) - t2 * p;
    return base + t * p;
}
RP_HD float rp_atan2f(float y, float x) {
    if (x != x || y != y) return x + y;
    const double ay = fabs((double)y), ax = fabs((double)x);
    double a;
    if (ay == 0.0 && ax == 0.0) a = signbit(x) ? 3.141592653589793 : 0.0;
    else {
        a = ay <= ax ? rp_atan_unit(ay / ax) : 1.5707963267948966 - rp_atan_unit(ax / ay);
        if (signbit(x)) a = 3.141592653589793 - a;
    }
    const float r = (float)a;
    return signbit(y) ? -r : r;
}
