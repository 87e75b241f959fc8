// Device narrowphase: one lane evaluates one shape pair (body pair or body vs one terrain
// triangle).  Behaviour follows crates/riftphys-collision/src/narrowphase.rs (collide_shapes
// :73-161 and its cases), in the reference's evaluation order; each function cites its lines.
// Contact feature ids (not present in the reference) are emitted as kind<<24 | which<<16 | k,
// bit 23 set when the shapes were swapped: which SAT axis / capsule test point / incident-face
// axis won, and which clipped candidate reduce_to_4 kept.
#pragma once
#include "rp_math.h"

enum { RK_SS = 1, RK_BB_FACE = 2, RK_BB_EDGE = 3, RK_BS_IN = 4, RK_BS_OUT = 5, RK_BT = 6, RK_ST = 7 };
#define RP_FEAT(kind, which, k) (((uint32_t)(kind) << 24) | ((uint32_t)(which) << 16) | (uint32_t)(k))
#define RP_FEAT_FLIP (1u << 23)

struct CPt { V3 pos, n; float pen; uint32_t feat; };
struct Manifold { CPt p[4]; int len; };
struct Cand { float depth; V3 p; };
struct Shape { int type; float d0, d1, d2; };  // 0 sphere (r), 1 box (hx,hy,hz), 2 capsule (r,hh)

RP_HD V3 norm_or(V3 v, V3 fb) {  // narrowphase.rs:45-54
    v = fin_v(v);
    float l2 = len2(v);
    return l2 > 1.0e-12f ? v * (1.0f / sqrtf(l2)) : fb;
}
RP_HD void mpush(Manifold &m, V3 pos, V3 n, float pen, uint32_t feat) {
    if (m.len < 4) { m.p[m.len].pos = pos; m.p[m.len].n = n; m.p[m.len].pen = pen; m.p[m.len].feat = feat; m.len++; }
}
RP_HD void mflip(Manifold &m) {  // narrowphase.rs:62-68
    for (int i = 0; i < m.len; i++) { m.p[i].n = -m.p[i].n; m.p[i].feat |= RP_FEAT_FLIP; }
}

// narrowphase.rs:312-338: deepest, farthest from it, farthest from that line, farthest on the
// other side; keys in 1e-6 units (sort_key6), lowest index wins ties.
RP_HD int reduce_to_4(const Cand *c, int n, V3 nrm, int out[4]) {
    if (n <= 4) { for (int i = 0; i < n; i++) out[i] = i; return n; }
    int i0 = 0; int64_t bk = INT64_MIN;
    for (int i = 0; i < n; i++) { int64_t k = sort_key6(c[i].depth); if (k > bk) { bk = k; i0 = i; } }
    V3 p0 = c[i0].p;
    int cnt = 0; out[cnt++] = i0;
    int i1 = 0; float d1 = -INFINITY; bk = INT64_MIN;
    for (int i = 0; i < n; i++) { float v = len2(c[i].p - p0); int64_t k = sort_key6(v); if (k > bk) { bk = k; i1 = i; d1 = v; } }
    if (!(d1 > 1.0e-12f)) return cnt;
    out[cnt++] = i1;
    V3 e = c[i1].p - p0;
    int i2 = 0; float a2 = -INFINITY; bk = INT64_MIN;
    for (int i = 0; i < n; i++) { float v = fabsf(dot3(cross3(c[i].p - p0, e), nrm)); int64_t k = sort_key6(v); if (k > bk) { bk = k; i2 = i; a2 = v; } }
    if (!(a2 > 1.0e-12f)) return cnt;
    out[cnt++] = i2;
    float side2 = dot3(cross3(c[i2].p - p0, e), nrm);
    float s2 = (side2 != side2) ? side2 : (signbit(side2) ? -1.0f : 1.0f);
    int i3 = 0; float a3 = -INFINITY; bk = INT64_MIN;
    for (int i = 0; i < n; i++) { float v = -dot3(cross3(c[i].p - p0, e), nrm) * s2; int64_t k = sort_key6(v); if (k > bk) { bk = k; i3 = i; a3 = v; } }
    if (a3 > 1.0e-12f) out[cnt++] = i3;
    return cnt;
}

// narrowphase.rs:289-301 Sutherland-Hodgman (keeps p.n <= d)
RP_HD int clip_poly(const V3 *in, int n, V3 nrm, float d, V3 *out) {
    int m = 0;
    for (int i = 0; i < n; i++) {
        V3 p = in[i], q = in[(i + 1 == n) ? 0 : i + 1];
        float dp = dot3(p, nrm) - d, dq = dot3(q, nrm) - d;
        if (dp <= 0.0f) out[m++] = p;
        if ((dp <= 0.0f) != (dq <= 0.0f)) { float t = dp / (dp - dq); out[m++] = p + (q - p) * t; }
    }
    return m;
}

RP_HD float box_radius(const V3 *ax, const float *e, V3 l) {
    return (e[0] * fabsf(dot3(ax[0], l)) + e[1] * fabsf(dot3(ax[1], l))) + e[2] * fabsf(dot3(ax[2], l));
}

// Incident face quad: the face of the box (axes ax, extents e, centre c0) most anti-parallel to
// nr, lowest index on ties; returns its axis.
RP_HD int incident_quad(V3 c0, const V3 *ax, const float *e, V3 nr, V3 *poly) {
    int f = 0; float fd = -INFINITY;
    for (int i = 0; i < 3; i++) { float d = fabsf(dot3(ax[i], nr)); if (d > fd) { fd = d; f = i; } }
    float s = dot3(ax[f], nr) > 0.0f ? -1.0f : 1.0f;
    int u = f + 1 > 2 ? f - 2 : f + 1, v = f + 2 > 2 ? f - 1 : f + 2;
    V3 c = c0 + ax[f] * (s * e[f]);
    V3 du = ax[u] * e[u], dv = ax[v] * e[v];
    poly[0] = (c + du) + dv;
    poly[1] = (c - du) + dv;
    poly[2] = (c - du) - dv;
    poly[3] = (c + du) - dv;
    return f;
}

// narrowphase.rs:173-256 box-box SAT + face clipping
RP_HD bool collide_box_box(V3 pa, Q4 qa, const float *ea, V3 pb, Q4 qb, const float *eb, float cd, Manifold &m) {
    V3 aax[3] = {qrot(qa, mk3(1, 0, 0)), qrot(qa, mk3(0, 1, 0)), qrot(qa, mk3(0, 0, 1))};
    V3 bax[3] = {qrot(qb, mk3(1, 0, 0)), qrot(qb, mk3(0, 1, 0)), qrot(qb, mk3(0, 0, 1))};
    V3 t = pb - pa;
    bool have = false; float bo = 0.0f; V3 bl = splat3(0.0f); int bw = 0;
    for (int it = 0; it < 15; it++) {
        V3 l; int which;
        if (it < 3) { l = aax[it]; which = it; }
        else if (it < 6) { l = bax[it - 3]; which = it; }
        else { l = cross3(aax[(it - 6) / 3], bax[(it - 6) % 3]); which = 6; }
        float l2 = len2(l);
        if (l2 < 1.0e-10f) continue;
        l = l / sqrtf(l2);
        float ov = (box_radius(aax, ea, l) + box_radius(bax, eb, l)) - fabsf(dot3(t, l));
        if (ov < -cd) return false;
        bool isf = which < 6, better;
        if (!have) better = true;
        else {
            bool face = bw < 6;
            if (isf && face && ((which >= 3) != (bw >= 3))) better = ov < bo - (0.02f * fabsf(bo) + 1.0e-3f);
            else if (isf == face) better = ov < bo;
            else if (isf) better = ov < bo * 1.05f + 1.0e-4f;
            else better = ov < bo * 0.95f - 1.0e-4f;
        }
        if (better) { have = true; bo = ov; bl = l; bw = which; }
    }
    if (!have) return false;
    V3 normal = norm_or(dot3(t, bl) < 0.0f ? -bl : bl, mk3(0, 1, 0));
    m.len = 0;
    if (bw < 6) {
        const V3 *rax = bw < 3 ? aax : bax;
        const V3 *iax = bw < 3 ? bax : aax;
        const float *re = bw < 3 ? ea : eb;
        const float *ie = bw < 3 ? eb : ea;
        V3 rc = bw < 3 ? pa : pb, ic = bw < 3 ? pb : pa;
        int k = bw < 3 ? bw : bw - 3;
        V3 nr = bw < 3 ? normal : -normal;
        V3 A[16], B[16];
        incident_quad(ic, iax, ie, nr, A);
        int n = 4;
        for (int tt = 0; tt < 2 && n > 0; tt++) {
            int j = (k + 1 + tt) % 3;
            V3 a = rax[j];
            float mid = dot3(rc, a);
            n = clip_poly(A, n, a, mid + re[j], B);
            if (n == 0) break;
            n = clip_poly(B, n, -a, -(mid - re[j]), A);
        }
        Cand cands[16]; int nc = 0;
        if (n > 0) {
            float plane = dot3(rc, nr) + re[k];
            for (int i = 0; i < n; i++) {
                float depth = plane - dot3(A[i], nr);
                if (depth > -cd) { cands[nc].depth = depth; cands[nc].p = A[i] + nr * (0.5f * depth); nc++; }
            }
        }
        if (nc > 0) {
            int idx[4]; int ni = reduce_to_4(cands, nc, normal, idx);
            for (int i = 0; i < ni; i++)
                mpush(m, fin_v(cands[idx[i]].p), normal, finf(cands[idx[i]].depth + cd), RP_FEAT(RK_BB_FACE, bw, idx[i]));
            return true;
        }
    }
    V3 sa = pa + normal * box_radius(aax, ea, normal);
    V3 sb = pb - normal * box_radius(bax, eb, normal);
    mpush(m, fin_v((sa + sb) * 0.5f), normal, finf(bo + cd), RP_FEAT(RK_BB_EDGE, bw, bw == 6 ? 0u : 0xFFFFu));
    return true;
}

// narrowphase.rs:349-409 box vs one-sided triangle (triangle given in world space)
RP_HD bool collide_box_tri(V3 bp, Q4 bq, const float *e, V3 ta, V3 tb, V3 tc, float cd, Manifold &m) {
    V3 raw = cross3(tb - ta, tc - ta);
    if (!(len2(raw) > 1.0e-12f) || !fin3(raw)) return false;
    V3 n = normalize3(raw);
    V3 ax[3] = {qrot(bq, mk3(1, 0, 0)), qrot(bq, mk3(0, 1, 0)), qrot(bq, mk3(0, 0, 1))};
    float r = box_radius(ax, e, n);
    float centre = dot3(bp - ta, n);
    if (centre - r > cd) return false;
    V3 A[16], B[16];
    int f = incident_quad(bp, ax, e, n, A);
    int np = 4;
    V3 ep0[3] = {ta, tb, tc}, ep1[3] = {tb, tc, ta};
    for (int s = 0; s < 3; s++) {
        V3 out = cross3(ep1[s] - ep0[s], n);
        if (!(len2(out) > 1.0e-12f)) return false;
        out = normalize3(out);
        np = clip_poly(A, np, out, dot3(ep0[s], out), B);
        if (np == 0) return false;
        for (int i = 0; i < np; i++) A[i] = B[i];
    }
    Cand cands[16]; int nc = 0;
    for (int i = 0; i < np; i++) {
        float dist = dot3(A[i] - ta, n);
        if (dist <= cd) { cands[nc].depth = -dist; cands[nc].p = A[i] - n * (0.5f * dist); nc++; }
    }
    if (nc == 0) return false;
    V3 normal = -n;
    m.len = 0;
    int idx[4]; int ni = reduce_to_4(cands, nc, normal, idx);
    for (int i = 0; i < ni; i++)
        mpush(m, fin_v(cands[idx[i]].p), normal, finf(cd + cands[idx[i]].depth), RP_FEAT(RK_BT, f, idx[i]));
    return true;
}

// narrowphase.rs:412-442
RP_HD bool collide_sphere_sphere(V3 ca, float ra, V3 cb, float rb, float cd, uint32_t feat, Manifold &m) {
    V3 delta = cb - ca;
    float d2 = len2(delta);
    ra = finf(fmaxf(ra, 0.0f)); rb = finf(fmaxf(rb, 0.0f));
    float allowed = finf((ra + rb) + cd);
    if (!isfinite(d2) || d2 > allowed * allowed) return false;
    float dist = sqrtf(d2);
    V3 normal = norm_or(dist > 1.0e-6f ? delta / dist : mk3(1, 0, 0), mk3(1, 0, 0));
    float pen = finf(fmaxf(allowed - dist, 0.0f));
    V3 p_a = ca + normal * ra, p_b = cb - normal * rb;
    m.len = 0;
    mpush(m, fin_v((p_a + p_b) * 0.5f), normal, pen, feat);
    return true;
}

RP_HD void capsule_seg(V3 p, Q4 q, float half, V3 &a, V3 &b) {  // narrowphase.rs:444-449
    V3 axis = qrot(q, mk3(0, 1, 0));
    float h = fabsf(half);
    a = p - axis * h; b = p + axis * h;
}
RP_HD V3 closest_on_seg(V3 p, V3 a, V3 b) {  // narrowphase.rs:451-460
    V3 ab = b - a;
    float l2 = len2(ab);
    if (l2 <= 1.0e-12f) return a;
    float t = clampf_rs(dot3(p - a, ab) / l2, 0.0f, 1.0f);
    return a + ab * t;
}
RP_HD void seg_seg(V3 p0, V3 p1, V3 q0, V3 q1, V3 &pa, V3 &pb) {  // narrowphase.rs:462-509
    V3 u = p1 - p0, v = q1 - q0, w = p0 - q0;
    float a = dot3(u, u), b = dot3(u, v), c = dot3(v, v), d = dot3(u, w), e = dot3(v, w);
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
    pa = p0 + u * sc; pb = q0 + v * tc;
}

// narrowphase.rs:540-593 (tp = box-capsule test point, 0 for a sphere)
RP_HD bool collide_box_sphere(V3 bp, Q4 bq, const float *e3, V3 sc, float sr, float cd, int tp, Manifold &m) {
    V3 be = mk3(e3[0], e3[1], e3[2]);
    V3 pl = fin_v(qrot(qconj(bq), sc - bp));
    V3 cl = clamp3(pl, -be, be);
    float r = finf(fmaxf(sr, 0.0f));
    float allowed = finf(r + cd);
    m.len = 0;
    if (eq3(cl, pl)) {
        V3 dtf = be - v3abs(pl);
        V3 axis; float depth; int which;
        if (dtf.x < dtf.y && dtf.x < dtf.z) { axis = mk3(1, 0, 0) * (pl.x > 0.0f ? 1.0f : -1.0f); depth = dtf.x; which = 0; }
        else if (dtf.y < dtf.z) { axis = mk3(0, 1, 0) * (pl.y > 0.0f ? 1.0f : -1.0f); depth = dtf.y; which = 1; }
        else { axis = mk3(0, 0, 1) * (pl.z > 0.0f ? 1.0f : -1.0f); depth = dtf.z; which = 2; }
        V3 normal = norm_or(qrot(bq, axis), qrot(bq, mk3(0, 1, 0)));
        V3 face = fin_v(sc + normal * depth);
        float pen = finf(allowed + fmaxf(depth, 0.0f));
        mpush(m, fin_v((face + (sc - normal * r)) * 0.5f), normal, pen, RP_FEAT(RK_BS_IN, tp, which));
        return true;
    }
    V3 cw = fin_v(bp + qrot(bq, cl));
    V3 delta = sc - cw;
    float dist = sqrtf(len2(delta));
    if (dist > allowed) return false;
    V3 normal = norm_or(dist > 1.0e-6f ? delta / dist : qrot(bq, mk3(0, 1, 0)), qrot(bq, mk3(0, 1, 0)));
    mpush(m, fin_v((cw + (sc - normal * r)) * 0.5f), normal, finf(allowed - dist), RP_FEAT(RK_BS_OUT, tp, 0));
    return true;
}

// narrowphase.rs:595-622
RP_HD bool collide_box_capsule(V3 bp, Q4 bq, const float *e3, V3 cp, Q4 cq, float cr, float chh, float cd, Manifold &m) {
    V3 aw, bw; capsule_seg(cp, cq, chh, aw, bw);
    Q4 inv = qconj(bq);
    V3 al = fin_v(qrot(inv, aw - bp)), bl = fin_v(qrot(inv, bw - bp));
    V3 tps[3] = {al, bl, closest_on_seg(splat3(0.0f), al, bl)};
    bool have = false;
    for (int i = 0; i < 3; i++) {
        Manifold t;
        if (collide_box_sphere(bp, bq, e3, bp + qrot(bq, tps[i]), cr, cd, i, t))
            if (!have || t.p[0].pen > m.p[0].pen) { m = t; have = true; }
    }
    return have;
}

// narrowphase.rs:627-669 (Ericson 5.1.5)
RP_HD V3 closest_on_tri(V3 p, V3 a, V3 b, V3 c) {
    V3 ab = b - a, ac = c - a, ap = p - a;
    float d1 = dot3(ab, ap), d2 = dot3(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f) return a;
    V3 bp = p - b;
    float d3 = dot3(ab, bp), d4 = dot3(ac, bp);
    if (d3 >= 0.0f && d4 <= d3) return b;
    float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) { float v = d1 / (d1 - d3); return a + mk3(v * ab.x, v * ab.y, v * ab.z); }
    V3 cp = p - c;
    float d5 = dot3(ab, cp), d6 = dot3(ac, cp);
    if (d6 >= 0.0f && d5 <= d6) return c;
    float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) { float w = d2 / (d2 - d6); return a + mk3(w * ac.x, w * ac.y, w * ac.z); }
    float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        V3 cb = c - b;
        return b + mk3(w * cb.x, w * cb.y, w * cb.z);
    }
    float denom = 1.0f / ((va + vb) + vc);
    float v = vb * denom, w = vc * denom;
    return (a + ab * v) + ac * w;
}

// narrowphase.rs:671-707
RP_HD bool collide_sphere_tri(V3 sc, float sr, V3 a, V3 b, V3 c, float cd, int tp, Manifold &m) {
    V3 cl = closest_on_tri(sc, a, b, c);
    V3 delta = sc - cl;
    float d2 = len2(delta);
    float r = finf(fmaxf(sr, 0.0f));
    float allowed = finf(r + cd);
    if (d2 > allowed * allowed) return false;
    float dist = sqrtf(d2);
    V3 tn = normalize_or_zero3(cross3(b - a, c - a));
    V3 normal = norm_or(dist > 1.0e-6f ? delta / dist : tn, tn);
    m.len = 0;
    mpush(m, fin_v(cl), normal, finf(allowed - dist), RP_FEAT(RK_ST, tp, 0));
    return true;
}

// narrowphase.rs:709-733
RP_HD bool collide_capsule_tri(V3 cp, Q4 cq, float cr, float chh, V3 a, V3 b, V3 c, float cd, Manifold &m) {
    V3 aw, bw; capsule_seg(cp, cq, chh, aw, bw);
    V3 tps[3] = {aw, bw, (aw + bw) * 0.5f};
    bool have = false;
    for (int i = 0; i < 3; i++) {
        Manifold t;
        if (collide_sphere_tri(tps[i], cr, a, b, c, cd, i, t))
            if (!have || t.p[0].pen > m.p[0].pen) { m = t; have = true; }
    }
    return have;
}

// narrowphase.rs:73-161 for two bodies
RP_HD bool collide_pair(Shape sa, V3 pa, Q4 qa, Shape sb, V3 pb, Q4 qb, float cd, Manifold &m) {
    m.len = 0;
    if (!fin3(pa) || !fin3(pb)) return false;
    cd = finf(fmaxf(cd, 0.0f));
    float ea[3] = {sa.d0, sa.d1, sa.d2}, eb[3] = {sb.d0, sb.d1, sb.d2};
    int ta = sa.type, tb = sb.type;
    if (ta == 0 && tb == 0) return collide_sphere_sphere(pa, sa.d0, pb, sb.d0, cd, RP_FEAT(RK_SS, 0, 0), m);
    if (ta == 2 && tb == 0) { V3 a, b; capsule_seg(pa, qa, sa.d1, a, b); return collide_sphere_sphere(closest_on_seg(pb, a, b), sa.d0, pb, sb.d0, cd, RP_FEAT(RK_SS, 1, 0), m); }
    if (ta == 0 && tb == 2) { V3 a, b; capsule_seg(pb, qb, sb.d1, a, b); if (!collide_sphere_sphere(closest_on_seg(pa, a, b), sb.d0, pa, sa.d0, cd, RP_FEAT(RK_SS, 1, 0), m)) return false; mflip(m); return true; }
    if (ta == 2 && tb == 2) { V3 a0, a1, b0, b1, x, y; capsule_seg(pa, qa, sa.d1, a0, a1); capsule_seg(pb, qb, sb.d1, b0, b1); seg_seg(a0, a1, b0, b1, x, y); return collide_sphere_sphere(x, sa.d0, y, sb.d0, cd, RP_FEAT(RK_SS, 2, 0), m); }
    if (ta == 1 && tb == 0) return collide_box_sphere(pa, qa, ea, pb, sb.d0, cd, 0, m);
    if (ta == 0 && tb == 1) { if (!collide_box_sphere(pb, qb, eb, pa, sa.d0, cd, 0, m)) return false; mflip(m); return true; }
    if (ta == 1 && tb == 2) return collide_box_capsule(pa, qa, ea, pb, qb, sb.d0, sb.d1, cd, m);
    if (ta == 2 && tb == 1) { if (!collide_box_capsule(pb, qb, eb, pa, qa, sa.d0, sa.d1, cd, m)) return false; mflip(m); return true; }
    if (ta == 1 && tb == 1) return collide_box_box(pa, qa, ea, pb, qb, eb, cd, m);
    return false;
}

// narrowphase.rs:121-151: a body shape against a terrain triangle at the identity transform
RP_HD bool collide_terrain(Shape s, V3 p, Q4 q, V3 a, V3 b, V3 c, float cd, Manifold &m) {
    m.len = 0;
    if (!fin3(p)) return false;
    cd = finf(fmaxf(cd, 0.0f));
    // The triangle's vertices go through Isometry::IDENTITY (pos 0 + rot * v): applied as written.
    Q4 I = qid();
    a = splat3(0.0f) + qrot(I, a); b = splat3(0.0f) + qrot(I, b); c = splat3(0.0f) + qrot(I, c);
    float e[3] = {s.d0, s.d1, s.d2};
    if (s.type == 0) return collide_sphere_tri(p, s.d0, a, b, c, cd, 0, m);
    if (s.type == 2) return collide_capsule_tri(p, q, s.d0, s.d1, a, b, c, cd, m);
    if (s.type == 1) return collide_box_tri(p, q, e, a, b, c, cd, m);
    return false;
}
