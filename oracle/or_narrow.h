/* Oracle (test infrastructure only): narrowphase restatement of
 * crates/riftphys-collision/src/narrowphase.rs (collide_shapes :73-161 and every case it
 * dispatches), in the reference's evaluation order, plus the build-defined contact feature ids
 * (kind<<24 | which<<16 | k, bit 23 = shapes were swapped).  Feature ids do not exist in the
 * reference; they are deterministic functions of the reference computation (which SAT axis won,
 * which clipped candidate `reduce_to_4` picked, which capsule test point was deepest).
 */
#ifndef OR_NARROW_H
#define OR_NARROW_H
#include "or_math.h"

enum { K_SS = 1, K_BB_FACE = 2, K_BB_EDGE = 3, K_BS_IN = 4, K_BS_OUT = 5, K_BT = 6, K_ST = 7 };
#define FEAT(kind, which, k) (((uint32_t)(kind) << 24) | ((uint32_t)(which) << 16) | (uint32_t)(k))
#define FEAT_FLIP (1u << 23)

typedef struct { V3 pos, n; float pen; uint32_t feat; } CPt;
typedef struct { V3 normal; CPt p[4]; int len; } Manifold;
typedef struct { V3 a, b, c; } Tri;

/* narrowphase.rs:35-54 */
static inline float nfin(float x) { return isfinite(x) ? x : 0.0f; }
static inline V3 nfin_v(V3 v) { return v3(nfin(v.x), nfin(v.y), nfin(v.z)); }
static inline V3 norm_or(V3 v, V3 fb) {
    v = nfin_v(v);
    float l2 = vlen2(v);
    if (l2 > 1.0e-12f) return vmul(v, 1.0f / sqrtf(l2));
    return fb;
}
static inline void mpush(Manifold *m, V3 pos, V3 n, float pen, uint32_t feat) {
    if (m->len < 4) { m->p[m->len].pos = pos; m->p[m->len].n = n; m->p[m->len].pen = pen; m->p[m->len].feat = feat; m->len++; }
}
/* narrowphase.rs:62-68 */
static inline void flip_manifold(Manifold *m) {
    m->normal = vneg(m->normal);
    for (int i = 0; i < m->len; i++) { m->p[i].n = m->normal; m->p[i].feat |= FEAT_FLIP; }
}

/* crates/riftphys-core/src/det.rs:21-25 */
static inline int64_t sort_key6(float x) {
    if (isnan(x)) return INT64_MAX;
    double k = round((double)x * 1.0e6);
    if (k >= (double)INT64_MAX) return INT64_MAX;
    if (k <= (double)INT64_MIN) return INT64_MIN;
    return (int64_t)k;
}

typedef struct { float depth; V3 p; } Cand;

/* narrowphase.rs:312-338 reduce_to_4: argmax by sort_key6, lowest index on ties */
static int reduce_to_4(const Cand *c, int n, V3 nrm, int out[4]) {
    if (n <= 4) { for (int i = 0; i < n; i++) out[i] = i; return n; }
    int i0 = 0; int64_t bk = INT64_MIN;
    for (int i = 0; i < n; i++) { int64_t k = sort_key6(c[i].depth); if (k > bk) { bk = k; i0 = i; } }
    V3 p0 = c[i0].p;
    int cnt = 0;
    out[cnt++] = i0;
    int i1 = 0; float d1 = -INFINITY; bk = INT64_MIN;
    for (int i = 0; i < n; i++) { float v = vlen2(vsub(c[i].p, p0)); int64_t k = sort_key6(v); if (k > bk) { bk = k; i1 = i; d1 = v; } }
    if (!(d1 > 1.0e-12f)) return cnt;
    out[cnt++] = i1;
    V3 e = vsub(c[i1].p, p0);
    int i2 = 0; float a2 = -INFINITY; bk = INT64_MIN;
    for (int i = 0; i < n; i++) { float v = fabsf(vdot(vcross(vsub(c[i].p, p0), e), nrm)); int64_t k = sort_key6(v); if (k > bk) { bk = k; i2 = i; a2 = v; } }
    if (!(a2 > 1.0e-12f)) return cnt;
    out[cnt++] = i2;
    float side2 = vdot(vcross(vsub(c[i2].p, p0), e), nrm);
    float s2 = isnan(side2) ? side2 : (signbit(side2) ? -1.0f : 1.0f);
    int i3 = 0; float a3 = -INFINITY; bk = INT64_MIN;
    for (int i = 0; i < n; i++) { float v = -vdot(vcross(vsub(c[i].p, p0), e), nrm) * s2; int64_t k = sort_key6(v); if (k > bk) { bk = k; i3 = i; a3 = v; } }
    if (a3 > 1.0e-12f) out[cnt++] = i3;
    return cnt;
}

/* narrowphase.rs:289-301 Sutherland-Hodgman, keeps p.n <= d */
static int clip_poly(const V3 *in, int n, V3 nrm, float d, V3 *out) {
    int m = 0;
    for (int i = 0; i < n; i++) {
        V3 p = in[i], q = in[(i + 1) % n];
        float dp = vdot(p, nrm) - d, dq = vdot(q, nrm) - d;
        if (dp <= 0.0f) out[m++] = p;
        if ((dp <= 0.0f) != (dq <= 0.0f)) { float t = dp / (dp - dq); out[m++] = vadd(p, vmul(vsub(q, p), t)); }
    }
    return m;
}

static inline float box_radius(const V3 ax[3], const float e[3], V3 l) {
    return (e[0] * fabsf(vdot(ax[0], l)) + e[1] * fabsf(vdot(ax[1], l))) + e[2] * fabsf(vdot(ax[2], l));
}

/* narrowphase.rs:261-286 */
static int clip_incident_face(V3 rc, const float re[3], const V3 rax[3], int k, V3 nr, V3 ic, const float ie[3],
                              const V3 iax[3], V3 *poly_out) {
    int f = 0; float fd = -INFINITY;
    for (int i = 0; i < 3; i++) { float d = fabsf(vdot(iax[i], nr)); if (d > fd) { fd = d; f = i; } }
    float s = vdot(iax[f], nr) > 0.0f ? -1.0f : 1.0f;
    int u = (f + 1) % 3, v = (f + 2) % 3;
    V3 c = vadd(ic, vmul(iax[f], s * ie[f]));
    V3 du = vmul(iax[u], ie[u]), dv = vmul(iax[v], ie[v]);
    V3 bufA[16], bufB[16];
    bufA[0] = vadd(vadd(c, du), dv);
    bufA[1] = vadd(vsub(c, du), dv);
    bufA[2] = vsub(vsub(c, du), dv);
    bufA[3] = vsub(vadd(c, du), dv);
    int n = 4;
    int js[2] = {(k + 1) % 3, (k + 2) % 3};
    for (int t = 0; t < 2; t++) {
        V3 a = rax[js[t]];
        float mid = vdot(rc, a);
        n = clip_poly(bufA, n, a, mid + re[js[t]], bufB);
        if (n == 0) return 0;
        n = clip_poly(bufB, n, vneg(a), -(mid - re[js[t]]), bufA);
        if (n == 0) return 0;
    }
    memcpy(poly_out, bufA, sizeof(V3) * n);
    return n;
}

/* narrowphase.rs:173-256 */
static int collide_box_box(V3 pa, Q4 qa, V3 ea, V3 pb, Q4 qb, V3 eb, float cd, Manifold *m) {
    V3 aax[3] = {qrot(qa, v3(1, 0, 0)), qrot(qa, v3(0, 1, 0)), qrot(qa, v3(0, 0, 1))};
    V3 bax[3] = {qrot(qb, v3(1, 0, 0)), qrot(qb, v3(0, 1, 0)), qrot(qb, v3(0, 0, 1))};
    float eaa[3] = {ea.x, ea.y, ea.z}, eba[3] = {eb.x, eb.y, eb.z};
    V3 t = vsub(pb, pa);
    int have = 0; float bo = 0; V3 bl = v3s(0); int bw = 0;
    for (int it = 0; it < 15; it++) {
        V3 l; int which;
        if (it < 3) { l = aax[it]; which = it; }
        else if (it < 6) { l = bax[it - 3]; which = it; }
        else { l = vcross(aax[(it - 6) / 3], bax[(it - 6) % 3]); which = 6; }
        float len2 = vlen2(l);
        if (len2 < 1.0e-10f) continue;
        l = vdivs(l, sqrtf(len2));
        float overlap = (box_radius(aax, eaa, l) + box_radius(bax, eba, l)) - fabsf(vdot(t, l));
        if (overlap < -cd) return 0;
        int is_face = which < 6;
        int better;
        if (!have) better = 1;
        else {
            int face = bw < 6;
            if (is_face && face && ((which >= 3) != (bw >= 3))) better = overlap < bo - (0.02f * fabsf(bo) + 1.0e-3f);
            else if (is_face == face) better = overlap < bo;
            else if (is_face) better = overlap < bo * 1.05f + 1.0e-4f;
            else better = overlap < bo * 0.95f - 1.0e-4f;
        }
        if (better) { have = 1; bo = overlap; bl = l; bw = which; }
    }
    if (!have) return 0;
    V3 n = vdot(t, bl) < 0.0f ? vneg(bl) : bl;
    V3 normal = norm_or(n, v3(0, 1, 0));
    m->normal = normal; m->len = 0;
    if (bw == 6) {
        V3 sa = vadd(pa, vmul(normal, box_radius(aax, eaa, normal)));
        V3 sb = vsub(pb, vmul(normal, box_radius(bax, eba, normal)));
        mpush(m, nfin_v(vmul(vadd(sa, sb), 0.5f)), normal, nfin(bo + cd), FEAT(K_BB_EDGE, 6, 0));
        return 1;
    }
    V3 rc, ic, nr; const float *re, *ie; const V3 *rax, *iax; int k;
    if (bw < 3) { rc = pa; re = eaa; rax = aax; ic = pb; ie = eba; iax = bax; k = bw; nr = normal; }
    else { rc = pb; re = eba; rax = bax; ic = pa; ie = eaa; iax = aax; k = bw - 3; nr = vneg(normal); }
    Cand cands[16]; int nc = 0;
    V3 poly[16];
    int np = clip_incident_face(rc, re, rax, k, nr, ic, ie, iax, poly);
    if (np > 0) {
        float plane = vdot(rc, nr) + re[k];
        for (int i = 0; i < np; i++) {
            float depth = plane - vdot(poly[i], nr);
            if (depth > -cd) { cands[nc].depth = depth; cands[nc].p = vadd(poly[i], vmul(nr, 0.5f * depth)); nc++; }
        }
    }
    if (nc == 0) {
        V3 sa = vadd(pa, vmul(normal, box_radius(aax, eaa, normal)));
        V3 sb = vsub(pb, vmul(normal, box_radius(bax, eba, normal)));
        mpush(m, nfin_v(vmul(vadd(sa, sb), 0.5f)), normal, nfin(bo + cd), FEAT(K_BB_EDGE, bw, 0xFFFF));
        return 1;
    }
    int idx[4]; int ni = reduce_to_4(cands, nc, normal, idx);
    for (int i = 0; i < ni; i++)
        mpush(m, nfin_v(cands[idx[i]].p), normal, nfin(cands[idx[i]].depth + cd), FEAT(K_BB_FACE, bw, idx[i]));
    return 1;
}

/* narrowphase.rs:349-409 */
static int collide_box_triangle(V3 bp, Q4 bq, V3 be, Tri tri, float cd, Manifold *m) {
    V3 ta = tri.a, tb = tri.b, tc = tri.c; /* identity transform: pos 0 + rot*I */
    ta = vadd(v3s(0), qrot(qident(), ta)); tb = vadd(v3s(0), qrot(qident(), tb)); tc = vadd(v3s(0), qrot(qident(), tc));
    V3 raw = vcross(vsub(tb, ta), vsub(tc, ta));
    if (!(vlen2(raw) > 1.0e-12f) || !vfinite(raw)) return 0;
    V3 n = vnormalize(raw);
    V3 ax[3] = {qrot(bq, v3(1, 0, 0)), qrot(bq, v3(0, 1, 0)), qrot(bq, v3(0, 0, 1))};
    float e[3] = {be.x, be.y, be.z};
    float r = box_radius(ax, e, n);
    float centre = vdot(vsub(bp, ta), n);
    if (centre - r > cd) return 0;
    int f = 0; float fd = -INFINITY;
    for (int i = 0; i < 3; i++) { float d = fabsf(vdot(ax[i], n)); if (d > fd) { fd = d; f = i; } }
    float sgn = vdot(ax[f], n) > 0.0f ? -1.0f : 1.0f;
    int u = (f + 1) % 3, v = (f + 2) % 3;
    V3 c = vadd(bp, vmul(ax[f], sgn * e[f]));
    V3 du = vmul(ax[u], e[u]), dv = vmul(ax[v], e[v]);
    V3 bufA[16], bufB[16];
    bufA[0] = vadd(vadd(c, du), dv);
    bufA[1] = vadd(vsub(c, du), dv);
    bufA[2] = vsub(vsub(c, du), dv);
    bufA[3] = vsub(vadd(c, du), dv);
    int np = 4;
    V3 ep[3][2] = {{ta, tb}, {tb, tc}, {tc, ta}};
    for (int s = 0; s < 3; s++) {
        V3 out = vcross(vsub(ep[s][1], ep[s][0]), n);
        if (!(vlen2(out) > 1.0e-12f)) return 0;
        out = vnormalize(out);
        np = clip_poly(bufA, np, out, vdot(ep[s][0], out), bufB);
        if (np == 0) return 0;
        memcpy(bufA, bufB, sizeof(V3) * np);
    }
    Cand cands[16]; int nc = 0;
    for (int i = 0; i < np; i++) {
        float dist = vdot(vsub(bufA[i], ta), n);
        if (dist <= cd) { cands[nc].depth = -dist; cands[nc].p = vsub(bufA[i], vmul(n, 0.5f * dist)); nc++; }
    }
    if (nc == 0) return 0;
    V3 normal = vneg(n);
    m->normal = normal; m->len = 0;
    int idx[4]; int ni = reduce_to_4(cands, nc, normal, idx);
    for (int i = 0; i < ni; i++)
        mpush(m, nfin_v(cands[idx[i]].p), normal, nfin(cd + cands[idx[i]].depth), FEAT(K_BT, f, idx[i]));
    return 1;
}

/* narrowphase.rs:412-442 */
static int collide_sphere_sphere(V3 ca, float ra, V3 cb, float rb, float cd, uint32_t feat, Manifold *m) {
    V3 delta = vsub(cb, ca);
    float dist2 = vlen2(delta);
    ra = nfin(fmaxf(ra, 0.0f)); rb = nfin(fmaxf(rb, 0.0f));
    float allowed = nfin((ra + rb) + cd);
    if (!isfinite(dist2) || dist2 > allowed * allowed) return 0;
    float dist = sqrtf(dist2);
    V3 raw = dist > 1.0e-6f ? vdivs(delta, dist) : v3(1, 0, 0);
    V3 normal = norm_or(raw, v3(1, 0, 0));
    float pen = nfin(fmaxf(allowed - dist, 0.0f));
    V3 p_a = vadd(ca, vmul(normal, ra)), p_b = vsub(cb, vmul(normal, rb));
    m->normal = normal; m->len = 0;
    mpush(m, nfin_v(vmul(vadd(p_a, p_b), 0.5f)), normal, pen, feat);
    return 1;
}

/* narrowphase.rs:444-460 */
static inline void capsule_seg(V3 p, Q4 q, float half, V3 *a, V3 *b) {
    V3 axis = qrot(q, v3(0, 1, 0));
    float h = fabsf(half);
    *a = vsub(p, vmul(axis, h)); *b = vadd(p, vmul(axis, h));
}
static inline V3 closest_point_segment(V3 p, V3 a, V3 b) {
    V3 ab = vsub(b, a);
    float l2 = vlen2(ab);
    if (l2 <= 1.0e-12f) return a;
    float t = vdot(vsub(p, a), ab) / l2;
    if (t < 0.0f) t = 0.0f;
    if (t > 1.0f) t = 1.0f;
    return vadd(a, vmul(ab, t));
}
/* narrowphase.rs:462-509 */
static void closest_seg_seg(V3 p0, V3 p1, V3 q0, V3 q1, V3 *pa, V3 *pb) {
    V3 u = vsub(p1, p0), v = vsub(q1, q0), w = vsub(p0, q0);
    float a = vdot(u, u), b = vdot(u, v), c = vdot(v, v), d = vdot(u, w), e = vdot(v, w);
    float d0 = a * c - b * b;
    float sn, tn, sd = d0, td = d0;
    const float EPS = 1.0e-6f;
    if (d0 < EPS) { sn = 0.0f; sd = 1.0f; tn = e; td = c; }
    else {
        sn = b * e - c * d; tn = a * e - b * d;
        if (sn < 0.0f) { sn = 0.0f; tn = e; td = c; }
        else if (sn > sd) { sn = sd; tn = e + b; td = c; }
    }
    if (tn < 0.0f) {
        tn = 0.0f;
        if (-d < 0.0f) sn = 0.0f; else if (-d > a) sn = sd; else { sn = -d; sd = a; }
    } else if (tn > td) {
        tn = td;
        if ((-d + b) < 0.0f) sn = 0.0f; else if ((-d + b) > a) sn = sd; else { sn = -d + b; sd = a; }
    }
    float sc = fabsf(sn) < EPS ? 0.0f : sn / sd;
    float tc = fabsf(tn) < EPS ? 0.0f : tn / td;
    *pa = vadd(p0, vmul(u, sc)); *pb = vadd(q0, vmul(v, tc));
}

/* narrowphase.rs:540-593 ; `tp` is the box-capsule test point (0 for a plain sphere) */
static int collide_box_sphere(V3 bp, Q4 bq, V3 be, V3 sc, float sr, float cd, int tp, Manifold *m) {
    V3 pl = nfin_v(qrot(qconj(bq), vsub(sc, bp)));
    V3 cl = vclamp(pl, vneg(be), be);
    float r = nfin(fmaxf(sr, 0.0f));
    float allowed = nfin(r + cd);
    m->len = 0;
    if (veq(cl, pl)) {
        V3 dtf = vsub(be, vabs(pl));
        V3 axis; float depth; int which;
        if (dtf.x < dtf.y && dtf.x < dtf.z) { axis = vmul(v3(1, 0, 0), pl.x > 0.0f ? 1.0f : -1.0f); depth = dtf.x; which = 0; }
        else if (dtf.y < dtf.z) { axis = vmul(v3(0, 1, 0), pl.y > 0.0f ? 1.0f : -1.0f); depth = dtf.y; which = 1; }
        else { axis = vmul(v3(0, 0, 1), pl.z > 0.0f ? 1.0f : -1.0f); depth = dtf.z; which = 2; }
        V3 normal = norm_or(qrot(bq, axis), qrot(bq, v3(0, 1, 0)));
        V3 face = nfin_v(vadd(sc, vmul(normal, depth)));
        float pen = nfin(allowed + fmaxf(depth, 0.0f));
        V3 pos = nfin_v(vmul(vadd(face, vsub(sc, vmul(normal, r))), 0.5f));
        m->normal = normal;
        mpush(m, pos, normal, pen, FEAT(K_BS_IN, tp, which));
        return 1;
    }
    V3 cw = nfin_v(vadd(bp, qrot(bq, cl)));
    V3 delta = vsub(sc, cw);
    float dist = sqrtf(vlen2(delta));
    if (dist > allowed) return 0;
    V3 raw = dist > 1.0e-6f ? vdivs(delta, dist) : qrot(bq, v3(0, 1, 0));
    V3 normal = norm_or(raw, qrot(bq, v3(0, 1, 0)));
    float pen = nfin(allowed - dist);
    V3 pos = nfin_v(vmul(vadd(cw, vsub(sc, vmul(normal, r))), 0.5f));
    m->normal = normal;
    mpush(m, pos, normal, pen, FEAT(K_BS_OUT, tp, 0));
    return 1;
}

/* narrowphase.rs:595-622 */
static int collide_box_capsule(V3 bp, Q4 bq, V3 be, V3 cp, Q4 cq, float cr, float chh, float cd, Manifold *m) {
    V3 aw, bw;
    capsule_seg(cp, cq, chh, &aw, &bw);
    Q4 inv = qconj(bq);
    V3 al = nfin_v(qrot(inv, vsub(aw, bp)));
    V3 bl = nfin_v(qrot(inv, vsub(bw, bp)));
    V3 ml = closest_point_segment(v3s(0), al, bl);
    V3 tps[3] = {al, bl, ml};
    int have = 0;
    for (int i = 0; i < 3; i++) {
        V3 pw = vadd(bp, qrot(bq, tps[i]));
        Manifold t;
        if (collide_box_sphere(bp, bq, be, pw, cr, cd, i, &t)) {
            if (!have || t.p[0].pen > m->p[0].pen) { *m = t; have = 1; }
        }
    }
    return have;
}

/* narrowphase.rs:627-669 (Ericson, Real-Time Collision Detection 5.1.5) */
static V3 closest_point_triangle(V3 p, V3 a, V3 b, V3 c) {
    V3 ab = vsub(b, a), ac = vsub(c, a), ap = vsub(p, a);
    float d1 = vdot(ab, ap), d2 = vdot(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f) return a;
    V3 bp = vsub(p, b);
    float d3 = vdot(ab, bp), d4 = vdot(ac, bp);
    if (d3 >= 0.0f && d4 <= d3) return b;
    float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) { float v = d1 / (d1 - d3); return vadd(a, v3(v * ab.x, v * ab.y, v * ab.z)); }
    V3 cp = vsub(p, c);
    float d5 = vdot(ab, cp), d6 = vdot(ac, cp);
    if (d6 >= 0.0f && d5 <= d6) return c;
    float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) { float w = d2 / (d2 - d6); return vadd(a, v3(w * ac.x, w * ac.y, w * ac.z)); }
    float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        V3 cb = vsub(c, b);
        return vadd(b, v3(w * cb.x, w * cb.y, w * cb.z));
    }
    float denom = 1.0f / ((va + vb) + vc);
    float v = vb * denom, w = vc * denom;
    return vadd(vadd(a, vmul(ab, v)), vmul(ac, w));
}

/* narrowphase.rs:671-707 */
static int collide_sphere_triangle(V3 sc, float sr, Tri tri, float cd, int tp, Manifold *m) {
    V3 a = vadd(v3s(0), qrot(qident(), tri.a)), b = vadd(v3s(0), qrot(qident(), tri.b)), c = vadd(v3s(0), qrot(qident(), tri.c));
    V3 cl = closest_point_triangle(sc, a, b, c);
    V3 delta = vsub(sc, cl);
    float d2 = vlen2(delta);
    float r = nfin(fmaxf(sr, 0.0f));
    float allowed = nfin(r + cd);
    if (d2 > allowed * allowed) return 0;
    float dist = sqrtf(d2);
    V3 tn = vnormalize_or_zero(vcross(vsub(b, a), vsub(c, a)));
    V3 raw = dist > 1.0e-6f ? vdivs(delta, dist) : tn;
    V3 normal = norm_or(raw, tn);
    m->normal = normal; m->len = 0;
    mpush(m, nfin_v(cl), normal, nfin(allowed - dist), FEAT(K_ST, tp, 0));
    return 1;
}

/* narrowphase.rs:709-733 */
static int collide_capsule_triangle(V3 cp, Q4 cq, float cr, float chh, Tri tri, float cd, Manifold *m) {
    V3 aw, bw;
    capsule_seg(cp, cq, chh, &aw, &bw);
    V3 mid = vmul(vadd(aw, bw), 0.5f);
    V3 tps[3] = {aw, bw, mid};
    int have = 0;
    for (int i = 0; i < 3; i++) {
        Manifold t;
        if (collide_sphere_triangle(tps[i], cr, tri, cd, i, &t)) {
            if (!have || t.p[0].pen > m->p[0].pen) { *m = t; have = 1; }
        }
    }
    return have;
}

/* shape: 0 sphere {r}, 1 box {hx,hy,hz}, 2 capsule {r,hh} */
typedef struct { int type; float d[3]; } ShapeD;

/* narrowphase.rs:73-161 for body-body pairs */
static int collide_shapes(ShapeD sa, V3 pa, Q4 qa, ShapeD sb, V3 pb, Q4 qb, float cd, Manifold *m) {
    if (!vfinite(pa) || !vfinite(pb)) return 0;
    cd = nfin(fmaxf(cd, 0.0f));
    m->len = 0;
    int ta = sa.type, tb = sb.type;
    if (ta == 0 && tb == 0) return collide_sphere_sphere(pa, sa.d[0], pb, sb.d[0], cd, FEAT(K_SS, 0, 0), m);
    if (ta == 2 && tb == 0) { V3 a, b; capsule_seg(pa, qa, sa.d[1], &a, &b); V3 q = closest_point_segment(pb, a, b); return collide_sphere_sphere(q, sa.d[0], pb, sb.d[0], cd, FEAT(K_SS, 1, 0), m); }
    if (ta == 0 && tb == 2) { V3 a, b; capsule_seg(pb, qb, sb.d[1], &a, &b); V3 q = closest_point_segment(pa, a, b); if (!collide_sphere_sphere(q, sb.d[0], pa, sa.d[0], cd, FEAT(K_SS, 1, 0), m)) return 0; flip_manifold(m); return 1; }
    if (ta == 2 && tb == 2) { V3 a0, a1, b0, b1, x, y; capsule_seg(pa, qa, sa.d[1], &a0, &a1); capsule_seg(pb, qb, sb.d[1], &b0, &b1); closest_seg_seg(a0, a1, b0, b1, &x, &y); return collide_sphere_sphere(x, sa.d[0], y, sb.d[0], cd, FEAT(K_SS, 2, 0), m); }
    if (ta == 1 && tb == 0) return collide_box_sphere(pa, qa, v3(sa.d[0], sa.d[1], sa.d[2]), pb, sb.d[0], cd, 0, m);
    if (ta == 0 && tb == 1) { if (!collide_box_sphere(pb, qb, v3(sb.d[0], sb.d[1], sb.d[2]), pa, sa.d[0], cd, 0, m)) return 0; flip_manifold(m); return 1; }
    if (ta == 1 && tb == 2) return collide_box_capsule(pa, qa, v3(sa.d[0], sa.d[1], sa.d[2]), pb, qb, sb.d[0], sb.d[1], cd, m);
    if (ta == 2 && tb == 1) { if (!collide_box_capsule(pb, qb, v3(sb.d[0], sb.d[1], sb.d[2]), pa, qa, sa.d[0], sa.d[1], cd, m)) return 0; flip_manifold(m); return 1; }
    if (ta == 1 && tb == 1) return collide_box_box(pa, qa, v3(sa.d[0], sa.d[1], sa.d[2]), pb, qb, v3(sb.d[0], sb.d[1], sb.d[2]), cd, m);
    return 0;
}

/* narrowphase.rs:121-151 for shape vs terrain triangle (shape first) */
static int collide_shape_tri(ShapeD s, V3 p, Q4 q, Tri tri, float cd, Manifold *m) {
    if (!vfinite(p)) return 0;
    cd = nfin(fmaxf(cd, 0.0f));
    m->len = 0;
    if (s.type == 0) return collide_sphere_triangle(p, s.d[0], tri, cd, 0, m);
    if (s.type == 2) return collide_capsule_triangle(p, q, s.d[0], s.d[1], tri, cd, m);
    if (s.type == 1) return collide_box_triangle(p, q, v3(s.d[0], s.d[1], s.d[2]), tri, cd, m);
    return 0;
}

#endif
