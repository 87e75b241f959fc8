// D6 joints on the device (articulation/src/bodies.rs:52-136, 321-390, 616-922; solver.rs:262-307,
// 336-424, 700-715, 822-832, 945-952), without armature: per axis (translational t0..t2 of frame A,
// rotational r0..r2) an enabled flag, compliance, a drive (position or velocity mode), a limit,
// damping and Coulomb friction.  Velocity level: d6_rows' dampers, friction rows, locked axes,
// motor rows and speculative limit rows, warm-started by row tag; position level: solve_d6's rows
// measured afresh from the current poses, XPBD with SHAKE directions, the hinge's twist in radians.
//
// A D6 joint is an island-list entry like a ball joint (RP_JBIT | q); its islands are solved by the
// contact-model tier (k_sweep_cm and its position passes), so the other tiers never meet one.
// Layout per joint q: jt_kind[q] (RP_JK_D6), jt_hinge[q] (stored hinge axis or RP_NONE), jt_ax[18 q
// + 3 k + c] for axis k (0..5) as (compliance, damping, friction, D6F_* bits), (drive target, kp,
// kd, max impulse), (limit min, max); this tick's rows d6_row[4 (32 q + r) + c] as (u, 1/(K +
// gamma)), (ra, target), (rb, gamma), (lower bound, upper bound, tag | angular << 8) and their count
// d6_nrows[q] and accumulators d6_racc[32 q + r]; the carried rows d6_wn[q] (bit 31 stored | count),
// d6_wtag / d6_wlam[32 q + r]; the position rows' multipliers d6_plam[6 q + k].
#pragma once

// D6 axis bits (jt_ax component 0, w)
enum { D6F_ENABLED = 1u, D6F_DRIVE = 2u, D6F_VEL = 4u, D6F_LIMIT = 8u };

struct D6AxD { uint32_t f; float comp, damp, fric, target, kp, kd, dmax, lo, hi; };
static __device__ __forceinline__ D6AxD d6_ax(const Dev &d, uint32_t q, int k) {
    const float4 a = d.jt_ax[18 * q + 3 * k], b = d.jt_ax[18 * q + 3 * k + 1], c = d.jt_ax[18 * q + 3 * k + 2];
    D6AxD x;
    x.comp = a.x; x.damp = a.y; x.fric = a.z; x.f = __float_as_uint(a.w);
    x.target = b.x; x.kp = b.y; x.kd = b.z; x.dmax = b.w; x.lo = c.x; x.hi = c.y;
    return x;
}
static __device__ __forceinline__ bool d6_is(const Dev &d, uint32_t q) { return d.d6_on && d.jt_kind[q] == RP_JK_D6; }
// a distance joint (the joint arrays above exist when a scene has D6 or distance joints)
static __device__ __forceinline__ bool dist_is(const Dev &d, uint32_t q) { return d.d6_on && d.jt_kind[q] == RP_JK_DIST; }
// a D6 or distance joint: its islands go to the contact-model tier
static __device__ __forceinline__ bool jx_is(const Dev &d, uint32_t q) { return d.d6_on && d.jt_kind[q] != 0u; }
// a distance joint's rest length and compliance (bodies.rs:273-285), in jt_ax[18 q]
static __device__ __forceinline__ float dist_rest(const Dev &d, uint32_t q) { return d.jt_ax[18 * q].x; }
static __device__ __forceinline__ float dist_compliance(const Dev &d, uint32_t q) { return d.jt_ax[18 * q].y; }
// a rotational velocity drive: its bodies keep the velocity solve's angular velocity (step.rs:353-365)
static __device__ __forceinline__ bool d6_motor(const Dev &d, uint32_t q) {
    for (int k = 3; k < 6; k++)
        if ((d6_ax(d, q, k).f & (D6F_DRIVE | D6F_VEL)) == (D6F_DRIVE | D6F_VEL)) return true;
    return false;
}
// bodies.rs:50-53 position_damped
static __device__ __forceinline__ bool d6_position_damped(const D6AxD &x) {
    return (x.f & D6F_ENABLED) && x.comp > 0.0f && x.damp > 0.0f && !(x.f & D6F_LIMIT) && (!(x.f & D6F_DRIVE) || !(x.f & D6F_VEL));
}
// bodies.rs:895-906 d6_hinge_axis (-1: none)
static __device__ int d6_hinge_axis(const Dev &d, uint32_t q) {
    const uint32_t h = d.jt_hinge[q];
    if (h != RP_NONE) return h < 3u ? (int)h : -1;
    int freeax = -1, n = 0;
    float c[3];
    bool en[3];
    for (int i = 0; i < 3; i++) {
        const D6AxD r = d6_ax(d, q, 3 + i);
        en[i] = (r.f & D6F_ENABLED) != 0;
        if (!(en[i] && !(r.f & D6F_LIMIT) && !(r.f & D6F_DRIVE))) { freeax = i; n++; }
        c[i] = en[i] ? fmaxf(jfin(r.comp), 0.0f) : 0.0f;
    }
    if (n == 1) return freeax;
    for (int i = 0; i < 3; i++) {
        if (!(en[i] && c[i] > 0.0f)) continue;
        bool all = true;
        for (int k = 0; k < 3; k++) if (k != i && !(c[i] >= 100.0f * c[k])) all = false;
        if (all) return i;
    }
    return -1;
}
static __device__ __forceinline__ float v3at(V3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
// bodies.rs:913-922 d6_row_angle (velocity rows' limit measure)
static __device__ float d6_row_angle(int hinge, Q4 rot_a, Q4 rot_b, int i) {
    Q4 q = qmul(qconj(rot_a), rot_b);
    if (q.w < 0.0f) q = mkq(-q.x, -q.y, -q.z, -q.w);
    const V3 v = mk3(q.x, q.y, q.z);
    if (hinge == i) return 2.0f * rp_atan2f(v3at(v, i), q.w);
    return dot3(qrot(rot_a, v * 2.0f), qrot(rot_a, unit_axis((uint32_t)i)));
}
static __device__ __forceinline__ void d6_put_row(const Dev &d, uint32_t q, uint32_t r, V3 u, V3 ra, V3 rb, float m, float target, float gamma,
                                                  float lo, float hi, uint32_t tag, bool angular) {
    float4 *o = d.d6_row + 4 * ((size_t)32 * q + r);
    o[0] = S3(u, m); o[1] = S3(ra, target); o[2] = S3(rb, gamma);
    o[3] = make_float4(lo, hi, __uint_as_float(tag | (angular ? 256u : 0u)), 0.0f);
    d.d6_racc[32 * q + r] = 0.0f;
}
// solver.rs:336-424 d6_rows (no armature) of joint q from the tick-start poses; returns the count
static __device__ uint32_t d6_rows(const Dev &d, uint32_t q, float dt) {
    const uint32_t a = d.jt_a[q], b = d.jt_b[q];
    const V3 pa = L3(d.pos[a]), pb = L3(d.pos[b]);
    const Q4 qa = LQ(d.rot[a]), qb = LQ(d.rot[b]);
    const Q4 rot_a = qmul(qa, LQ(d.jt_far[q])), rot_b = qmul(qb, LQ(d.jt_fbr[q]));
    const V3 xa = pa + qrot(qa, L3(d.jt_fap[q])), xb = pb + qrot(qb, L3(d.jt_fbp[q]));
    const float wa = d.inv_mass[a], wb = d.inv_mass[b];
    if (wa == 0.0f && wb == 0.0f) return 0;
    const M3 ia = load_m9(d.iiw + 9 * a), ib = load_m9(d.iiw + 9 * b);
    const V3 ra = xa - pa, rb = xb - pb, dd = xb - xa;
    const int hinge = d6_hinge_axis(d, q);
    uint32_t n = 0;
    for (int angular = 0; angular < 2; angular++)
        for (int i = 0; i < 3; i++) {
            const D6AxD row = d6_ax(d, q, 3 * angular + i);
            const V3 u = qrot(rot_a, unit_axis((uint32_t)i));
            float kk;
            if (angular) kk = dot3(u, mv(ia, u)) + dot3(u, mv(ib, u));
            else { const V3 ca = cross3(ra, u), cb = cross3(rb, u); kk = ((wa + wb) + dot3(ca, mv(ia, ca))) + dot3(cb, mv(ib, cb)); }
            const uint32_t tag = ((uint32_t)angular << 5) | ((uint32_t)i << 3);
            const float damp = d6_position_damped(row) ? 0.0f : row.damp;
            if ((damp > 0.0f || row.fric > 0.0f) && kk > 1.0e-9f) {
                if (damp > 0.0f) {
                    const float gamma = 1.0f / (damp * dt);
                    d6_put_row(d, q, n++, u, ra, rb, 1.0f / (kk + gamma), 0.0f, gamma, -INFINITY, INFINITY, tag | 4u, angular);
                }
                if (row.fric > 0.0f) {
                    const float cap = row.fric * dt;
                    d6_put_row(d, q, n++, u, ra, rb, 1.0f / kk, 0.0f, 0.0f, -cap, cap, tag | 5u, angular);
                }
            }
            if (!(row.f & D6F_ENABLED) || row.comp > 0.0f) continue;
            if ((row.f & D6F_DRIVE) && !(row.f & D6F_VEL)) continue;
            if (kk <= 1.0e-9f) continue;
            const float m = 1.0f / kk;
            if (row.f & D6F_DRIVE) {
                const float cap = fabsf(row.dmax);
                d6_put_row(d, q, n++, u, ra, rb, m, row.target, 0.0f, -cap, cap, tag | 1u, angular);
            }
            if (!(row.f & D6F_LIMIT)) {
                if (!(row.f & D6F_DRIVE)) d6_put_row(d, q, n++, u, ra, rb, m, 0.0f, 0.0f, -INFINITY, INFINITY, tag, angular);
            } else {
                const float x = angular ? d6_row_angle(hinge, rot_a, rot_b, i) : dot3(dd, u);
                d6_put_row(d, q, n++, u, ra, rb, m, fminf((row.lo - x) / dt, 0.0f), 0.0f, 0.0f, INFINITY, tag | 2u, angular);
                d6_put_row(d, q, n++, u, ra, rb, m, fmaxf((row.hi - x) / dt, 0.0f), 0.0f, -INFINITY, 0.0f, tag | 3u, angular);
            }
        }
    return n;
}
// apply_angular_impulse on one side (dynamics.rs:193-205)
static __device__ __forceinline__ void d6_ang_impulse(const Dev &d, uint32_t b, V3 tau) {
    if (d.inv_mass[b] > 0.0f && fin3(tau)) d.ang[b] = S3(fin_v(L3(d.ang[b]) + mv(load_m9(d.iiw + 9 * b), tau)));
}
// joint_row_apply / joint_row_vel (solver.rs:291-307) for row r of joint q
static __device__ __forceinline__ void d6_apply(const Dev &d, uint32_t a, uint32_t b, const float4 *o, float lam) {
    const V3 u = L3(o[0]);
    if (__float_as_uint(o[3].z) & 256u) {
        const V3 p = u * lam;
        if (d.inv_mass[a] > 0.0f) d6_ang_impulse(d, a, -p);
        if (d.inv_mass[b] > 0.0f) d6_ang_impulse(d, b, p);
    } else joint_apply(d, a, b, L3(o[1]), L3(o[2]), u, lam);
}
static __device__ __forceinline__ float d6_vel(const Dev &d, uint32_t a, uint32_t b, const float4 *o) {
    const V3 u = L3(o[0]);
    if (__float_as_uint(o[3].z) & 256u) return dot3(L3(d.ang[b]) - L3(d.ang[a]), u);
    return joint_vel(d, a, b, L3(o[1]), L3(o[2]), u);
}
static __device__ __forceinline__ float d6_clamp(float x, float lo, float hi) { if (x < lo) x = lo; if (x > hi) x = hi; return x; }
// solver.rs:700-715: last tick's rows when the joint has the same rows (tags), clamped to the bounds
static __device__ void d6_warm(const Dev &d, uint32_t q) {
    const uint32_t n = d.d6_nrows[q], wn = d.d6_wn[q];
    if (!n || !(wn >> 31) || (wn & 0xFFu) != n) return;
    for (uint32_t r = 0; r < n; r++)
        if (d.d6_wtag[32 * q + r] != (__float_as_uint(d.d6_row[4 * ((size_t)32 * q + r) + 3].z) & 255u)) return;
    const uint32_t a = d.jt_a[q], b = d.jt_b[q];
    for (uint32_t r = 0; r < n; r++) {
        const float4 *o = d.d6_row + 4 * ((size_t)32 * q + r);
        const float lam = d6_clamp(d.d6_wlam[32 * q + r], o[3].x, o[3].y);
        d.d6_racc[32 * q + r] = lam;
        if (lam != 0.0f) d6_apply(d, a, b, o, lam);
    }
}
// solver.rs:822-832: one pass over the joint's rows; residual bits |d lambda| (K + gamma)
static __device__ uint32_t d6_pgs(const Dev &d, uint32_t q, bool resid) {
    const uint32_t n = d.d6_nrows[q], a = d.jt_a[q], b = d.jt_b[q];
    uint32_t rbits = 0;
    for (uint32_t r = 0; r < n; r++) {
        const float4 *o = d.d6_row + 4 * ((size_t)32 * q + r);
        const float m = o[0].w, target = o[1].w, gamma = o[2].w;
        const float v = d6_vel(d, a, b, o);
        const float old = d.d6_racc[32 * q + r];
        float nw = gamma == 0.0f ? old + (target - v) * m : old + ((target - v) - gamma * old) * m;
        nw = d6_clamp(nw, o[3].x, o[3].y);
        d.d6_racc[32 * q + r] = nw;
        if (nw != old) d6_apply(d, a, b, o, nw - old);
        if (resid) rbits = max(rbits, __float_as_uint(fabsf(nw - old) / m));
    }
    return rbits;
}
// solver.rs:945-952: a joint with rows stores them
static __device__ __forceinline__ void d6_store(const Dev &d, uint32_t q) {
    const uint32_t n = d.d6_nrows[q];
    if (!n) return;
    d.d6_wn[q] = 0x80000000u | n;
    for (uint32_t r = 0; r < n; r++) {
        d.d6_wtag[32 * q + r] = __float_as_uint(d.d6_row[4 * ((size_t)32 * q + r) + 3].z) & 255u;
        d.d6_wlam[32 * q + r] = d.d6_racc[32 * q + r];
    }
}
// bodies.rs:1010-1030 row_error; false: no row
static __device__ __forceinline__ bool d6_row_error(const D6AxD &x, float coord, float &c) {
    bool has = true;
    c = jfin(coord - 0.0f);
    if (x.f & D6F_LIMIT) {
        const float lo = jfin(x.lo), hi = jfin(x.hi);
        if (coord < lo) c = jfin(coord - lo);
        else if (coord > hi) c = jfin(coord - hi);
        else has = false;
    }
    if (x.f & D6F_DRIVE) {
        if (!(x.f & D6F_VEL)) {
            float t = jfin(x.target);
            if (x.f & D6F_LIMIT) t = d6_clamp(t, jfin(x.lo), jfin(x.hi));
            c = jfin(coord - t);
            has = true;
        } else if (!(x.f & D6F_LIMIT)) has = false;
    }
    return has;
}
static __device__ __forceinline__ float d6_alpha(float c, float dt2) { c = fmaxf(jfin(c), 0.0f); return c <= 0.0f ? 0.0f : jfin(c / dt2); }
// the angle a rotational position row measures at frame rotations ra, rb (bodies.rs:786-795)
static __device__ float d6_pos_angle(int hinge, int i, Q4 ra, Q4 rb, V3 u) {
    Q4 q = qmul(qconj(ra), rb);
    if (q.w < 0.0f) q = mkq(-q.x, -q.y, -q.z, -q.w);
    const V3 v = mk3(q.x, q.y, q.z);
    if (hinge == i) return jfin(2.0f * rp_atan2f(v3at(v, i), q.w));
    if (hinge >= 0) {
        const V3 ah = normalize_fin(qrot(ra, unit_axis((uint32_t)hinge))), bh = normalize_fin(qrot(rb, unit_axis((uint32_t)hinge)));
        return jfin(dot3(fin_v(cross3(ah, bh)), u));
    }
    return jfin(dot3(fin_v(qrot(ra, fin_v(v * 2.0f))), u));
}
// bodies.rs:616-873 solve_d6 (no rotor rows), one pass: the translational rows, then the rotational
// ones, each measured from the current poses; SHAKE directions from the step's start poses unless a
// body turned 0.5 rad or more; a damped compliant row adds XPBD's damping term.
static __device__ void d6_pos(const Dev &d, uint32_t q, float dt) {
    const uint32_t a = d.jt_a[q], b = d.jt_b[q];
    const float wa = jfin(d.inv_mass[a]), wb = jfin(d.inv_mass[b]);
    const float dt2 = fmaxf(dt * dt, 1.0e-12f);
    const V3 sap = L3(d.opos[a]), sbp = L3(d.opos[b]);
    const Q4 sa = LQ(d.orot[a]), sb = LQ(d.orot[b]);
    const float cmin = __uint_as_float(RP_SHAKE_COS_BITS);
    const bool shake = fminf(fabsf(qmul(sa, qconj(LQ(d.rot[a]))).w), 1.0f) >= cmin && fminf(fabsf(qmul(sb, qconj(LQ(d.rot[b]))).w), 1.0f) >= cmin;
    const Q4 fra = LQ(d.jt_far[q]), frb = LQ(d.jt_fbr[q]);
    const V3 fap = L3(d.jt_fap[q]), fbp = L3(d.jt_fbp[q]);
    V3 la0 = splat3(0.0f), lb0 = splat3(0.0f);
    M3 ia0 = mzero(), ib0 = mzero();
    Q4 ra0s = qid();
    if (shake) {
        ra0s = qmul(sa, fra);
        la0 = fin_v(qrot(sa, fap)); lb0 = fin_v(qrot(sb, fbp));
        ia0 = inertia_world_of(d.inv_mass[a] == 0.0f ? mzero() : load_m9(d.iil + 9 * a), sa);
        ib0 = inertia_world_of(d.inv_mass[b] == 0.0f ? mzero() : load_m9(d.iil + 9 * b), sb);
    }
    for (int i = 0; i < 3; i++) {
        const D6AxD row = d6_ax(d, q, i);
        if (!(row.f & D6F_ENABLED)) continue;
        const V3 pa = L3(d.pos[a]), pb = L3(d.pos[b]);
        const Q4 qa = LQ(d.rot[a]), qb = LQ(d.rot[b]);
        const V3 u = normalize_fin(qrot(qmul(qa, fra), unit_axis((uint32_t)i)));
        if (zero3(u)) continue;
        const V3 xa = fin_v(pa + qrot(qa, fap)), xb = fin_v(pb + qrot(qb, fbp));
        const V3 dd = fin_v(xb - xa);
        V3 un, lva, lvb; M3 iaw, ibw;
        if (shake) { un = normalize_fin(qrot(ra0s, unit_axis((uint32_t)i))); lva = la0; lvb = lb0; iaw = ia0; ibw = ib0; }
        else { un = u; lva = xa - pa; lvb = xb - pb; iaw = inv_inertia_world(d, a, qa); ibw = inv_inertia_world(d, b, qb); }
        const float coord = jfin(dot3(dd, u));
        float c;
        if (!d6_row_error(row, coord, c)) continue;
        const V3 rau = fin_v(cross3(lva, un)), rbu = fin_v(cross3(lvb, un));
        const float wra = jfin(dot3(rau, mv(iaw, rau))), wrb = jfin(dot3(rbu, mv(ibw, rbu)));
        const float wsum = jfin(((wa + wb) + wra) + wrb);
        const float alpha = d6_alpha(row.comp, dt2);
        const float denom = jfin(wsum + alpha);
        if (denom <= 0.0f) continue;
        c = jfin(c + jfin(0.0f * dt));
        const float lold = jfin(d.d6_plam[6 * q + i]);
        float dl;
        if (d6_position_damped(row)) {
            const V3 u0 = normalize_fin(qrot(qmul(sa, fra), unit_axis((uint32_t)i)));
            const V3 d0 = fin_v((sbp + qrot(sb, fbp)) - (sap + qrot(sa, fap)));
            const float gamma = jfin(row.comp * row.damp / dt);
            const float dc = jfin(coord - dot3(d0, u0));
            dl = jfin(-((c + jfin(alpha * lold)) + jfin(gamma * dc)) / jfin((1.0f + gamma) * wsum + alpha));
        } else dl = jfin(-(c + jfin(alpha * lold)) / denom);
        d.d6_plam[6 * q + i] = jfin(lold + dl);
        joint_pos_delta(d, a, fin_v((-un) * jfin(dl * wa)));
        joint_pos_delta(d, b, fin_v(un * jfin(dl * wb)));
        joint_orient_delta(d, a, fin_v((-mv(iaw, rau)) * dl));
        joint_orient_delta(d, b, fin_v(mv(ibw, rbu) * dl));
    }
    const int hinge = d6_hinge_axis(d, q);
    for (int i = 0; i < 3; i++) {
        const D6AxD row = d6_ax(d, q, 3 + i);
        if (!(row.f & D6F_ENABLED)) continue;
        const Q4 ra = qmul(LQ(d.rot[a]), fra), rb = qmul(LQ(d.rot[b]), frb);
        const V3 u = normalize_fin(qrot(ra, unit_axis((uint32_t)i)));
        if (zero3(u)) continue;
        const float angle = d6_pos_angle(hinge, i, ra, rb, u);
        V3 un; M3 iaw, ibw;
        if (shake) { un = normalize_fin(qrot(ra0s, unit_axis((uint32_t)i))); iaw = ia0; ibw = ib0; }
        else { un = u; iaw = inv_inertia_world(d, a, LQ(d.rot[a])); ibw = inv_inertia_world(d, b, LQ(d.rot[b])); }
        float c;
        if (!d6_row_error(row, angle, c)) continue;
        const float wsum = jfin(dot3(un, mv(iaw, un)) + dot3(un, mv(ibw, un)));
        const float alpha = d6_alpha(row.comp, dt2);
        const float denom = jfin(wsum + alpha);
        if (denom <= 0.0f) continue;
        c = jfin(c + jfin(0.0f * dt));
        const float lold = jfin(d.d6_plam[6 * q + 3 + i]);
        float dl;
        if (d6_position_damped(row)) {
            const Q4 r0a = qmul(sa, fra), r0b = qmul(sb, frb);
            const V3 u0 = normalize_fin(qrot(r0a, unit_axis((uint32_t)i)));
            const float angle0 = d6_pos_angle(hinge, i, r0a, r0b, u0);
            const float gamma = jfin(row.comp * row.damp / dt);
            const float dc = jfin(angle - angle0);
            dl = jfin(-((c + jfin(alpha * lold)) + jfin(gamma * dc)) / jfin((1.0f + gamma) * wsum + alpha));
        } else dl = jfin(-(c + jfin(alpha * lold)) / denom);
        d.d6_plam[6 * q + 3 + i] = jfin(lold + dl);
        joint_orient_delta(d, a, fin_v((-mv(iaw, un)) * dl));
        joint_orient_delta(d, b, fin_v(mv(ibw, un) * dl));
    }
}
// solver.rs:795-820: a rigid distance joint's legacy row, run with the joint rows in every velocity
// sweep: a Baumgarte velocity row along the centre line, neither warm-started nor accumulated.
// Returns the residual's bits, |j| (wa + wb).
static __device__ uint32_t dist_vel(const Dev &d, uint32_t q, float dt) {
    if (dist_compliance(d, q) > 0.0f) return 0u;  // a compliant joint acts at position level only
    const uint32_t a = d.jt_a[q], b = d.jt_b[q];
    const V3 dir = L3(d.pos[b]) - L3(d.pos[a]);
    const float len = len3(dir);
    if (len < 1.0e-6f) return 0u;
    const V3 n = dir / len;
    const float ia = d.inv_mass[a], ib = d.inv_mass[b];
    const float isum = ia + ib;
    if (isum < 1.0e-9f) return 0u;
    const float vrel = dot3(L3(d.lin[b]) - L3(d.lin[a]), n);
    const float bias = (len - dist_rest(d, q)) * (0.1f / dt);
    const float jm = -(vrel + bias) / isum;
    const V3 pa = (-n) * jm, pb = n * jm;
    if (ia > 0.0f && fin3(pa)) d.lin[a] = S3(fin_v(L3(d.lin[a]) + pa * ia));
    if (ib > 0.0f && fin3(pb)) d.lin[b] = S3(fin_v(L3(d.lin[b]) + pb * ib));
    return __float_as_uint(fabsf(jm) * isum);
}
// bodies.rs:501-548 solve_distance_row: the XPBD row along the step-start centre line (SHAKE)
// unless the rod turned by 0.5 rad or more, its multiplier accumulated over the step's passes in
// d6_plam[6 q]
static __device__ void dist_pos(const Dev &d, uint32_t q, float dt) {
    const uint32_t a = d.jt_a[q], b = d.jt_b[q];
    const float wa = d.inv_mass[a], wb = d.inv_mass[b];
    const float wsum = jfin(wa + wb);
    if (wsum <= 0.0f) return;
    const V3 dd = L3(d.pos[b]) - L3(d.pos[a]);
    const float l2 = len2(dd);
    if (!isfinite(l2) || l2 <= 1.0e-20f) return;
    const float len = sqrtf(l2);
    if (len <= 1.0e-6f) return;
    const V3 ncur = normalize_fin(dd / len);
    if (zero3(ncur)) return;
    V3 n = ncur;
    const V3 n0 = normalize_fin(L3(d.opos[b]) - L3(d.opos[a]));
    if (!zero3(n0) && dot3(n0, ncur) > __uint_as_float(0x3f60a940u)) n = n0;  // cos 0.5 as f32
    const float c = jfin(len - jfin(dist_rest(d, q)));
    const float dt2 = fmaxf(dt * dt, 1.0e-12f);
    const float comp = dist_compliance(d, q);
    const float alpha = comp <= 0.0f ? 0.0f : jfin(fmaxf(jfin(comp), 0.0f) / dt2);
    const float denom = jfin(wsum + alpha);
    if (denom <= 0.0f) return;
    const float lold = jfin(d.d6_plam[6 * q]);
    const float lam = jfin(-(c + jfin(alpha * lold)) / (denom * dot3(n, ncur)));
    if (fabsf(lam) <= 1.0e-9f) return;
    d.d6_plam[6 * q] = jfin(lold + lam);
    joint_pos_delta(d, a, fin_v((-n) * jfin(lam * wa)));
    joint_pos_delta(d, b, fin_v(n * jfin(lam * wb)));
}
// Joints::begin_step (bodies.rs:308-318): the step's position multipliers start at zero
__global__ void k_d6_begin(Dev d) {
    if (aborted(d)) return;
    GRID_STRIDE(i, 6 * d.nj) d.d6_plam[i] = 0.0f;
}
