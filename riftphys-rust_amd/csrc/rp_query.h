// Spatial queries on the device (world/src/world/queries.rs:50-261): World::raycast against the
// colliders' bounding boxes and the terrain's triangles, World::sweep_capsule (overlap at the
// start against the true shapes, then the capsule's tip and base spheres swept against boxes).
//
// One workgroup per query, over the current device state of one world of the scene.  The
// colliders (in collider order) and the terrain cells under the query's box (gz outer, gx inner,
// two triangles each) are tested 256 at a time; the reference's tie rules depend on scan order
// (a later hit replaces the best only when earlier by more than 1e-9), so each chunk's hits are
// folded in order by lane 0.  First-hit searches (the start overlap) take the lowest index.
//
// Included by rp_step.hip after the narrowphase and terrain helpers.
#pragma once

#define RPQ_T 256u

// world/src/world/geom.rs:40-93 normal_from_aabb_point
static __device__ V3 rq_normal_from_aabb_point(V3 mn, V3 mx, V3 p) {
    const float eps = 1.0e-4f;
    if (fabsf(p.x - mn.x) < eps) return mk3(-1, 0, 0);
    if (fabsf(p.x - mx.x) < eps) return mk3(1, 0, 0);
    if (fabsf(p.y - mn.y) < eps) return mk3(0, -1, 0);
    if (fabsf(p.y - mx.y) < eps) return mk3(0, 1, 0);
    if (fabsf(p.z - mn.z) < eps) return mk3(0, 0, -1);
    if (fabsf(p.z - mx.z) < eps) return mk3(0, 0, 1);
    const V3 n = p - (mn + mx) * 0.5f;
    const float l2 = len2(n);
    if (l2 > 1.0e-20f) return n / sqrtf(l2);
    return mk3(0, 1, 0);
}
// geom.rs:64-93 ray_aabb: slabs with the ray parameter in [0, max_dist]
static __device__ bool rq_ray_aabb(V3 o, V3 dv, V3 mn, V3 mx, float max_dist, float &t_out, V3 &p_out, V3 &n_out) {
    const float eps = 1.0e-8f;
    float tmin = 0.0f, tmax = max_dist;
    const float oo[3] = {o.x, o.y, o.z}, dd[3] = {dv.x, dv.y, dv.z};
    const float lo[3] = {mn.x, mn.y, mn.z}, hi[3] = {mx.x, mx.y, mx.z};
    for (int a = 0; a < 3; a++) {
        if (fabsf(dd[a]) < eps) {
            if (oo[a] < lo[a] || oo[a] > hi[a]) return false;
        } else {
            const float inv = 1.0f / dd[a];
            float t1 = (lo[a] - oo[a]) * inv, t2 = (hi[a] - oo[a]) * inv;
            if (t1 > t2) { const float x = t1; t1 = t2; t2 = x; }
            if (t1 > tmin) tmin = t1;
            if (t2 < tmax) tmax = t2;
            if (tmin > tmax) return false;
        }
    }
    if (tmin < 0.0f || tmin > max_dist) return false;
    t_out = tmin;
    p_out = o + dv * tmin;
    n_out = rq_normal_from_aabb_point(mn, mx, p_out);
    return true;
}
// geom/src/triangle.rs:215-232 Triangle::intersect_ray(ray, 0, t_max) through plane.rs:15-86 and
// the barycentric inside test (triangle.rs:94-143)
static __device__ bool rq_tri_ray(V3 ta, V3 tb, V3 tc, V3 o, V3 dv, float t_max, float &t_out) {
    if (0.0f > t_max) return false;
    const V3 nu = cross3(tb - ta, tc - ta);
    const float a2 = dot3(nu, nu);
    if (!isfinite(a2) || a2 <= 1.0e-16f) return false;
    if (!(fin3(ta) && fin3(tb) && fin3(tc))) return false;
    const float l2 = len2(nu);
    if (!isfinite(l2) || l2 <= 1.0e-6f * 1.0e-6f) return false;
    const V3 pn = nu * (1.0f / sqrtf(l2));
    const float pd = -dot3(pn, ta);
    if (!(fin3(pn) && isfinite(pd)) || !(fin3(o) && fin3(dv))) return false;
    const float denom = dot3(pn, dv);
    if (!isfinite(denom) || fabsf(denom) < 1.0e-6f) return false;
    const float numer = -(dot3(pn, o) + pd);
    if (!isfinite(numer)) return false;
    const float t = numer / denom;
    if (!isfinite(t) || !(t >= 0.0f && t <= t_max)) return false;
    const V3 p = o + dv * t;
    if (!fin3(p)) return false;
    const V3 v0 = tb - ta, v1 = tc - ta, v2 = p - ta;
    const float d00 = dot3(v0, v0), d01 = dot3(v0, v1), d11 = dot3(v1, v1), d20 = dot3(v2, v0), d21 = dot3(v2, v1);
    const float den = d00 * d11 - d01 * d01;
    const float scale = fmaxf(fabsf(d00 * d11), 1.0f);
    if (!isfinite(den) || fabsf(den) <= 1.0e-12f * scale) return false;
    const float id = 1.0f / den;
    const float bv = (d11 * d20 - d01 * d21) * id, bw = (d00 * d21 - d01 * d20) * id;
    const float bu = 1.0f - bv - bw;
    if (!(isfinite(bu) && isfinite(bv) && isfinite(bw))) return false;
    if (!(bu >= -1.0e-5f && bv >= -1.0e-5f && bw >= -1.0e-5f)) return false;
    t_out = t;
    return true;
}
static __device__ V3 rq_tri_normal(V3 ta, V3 tb, V3 tc) {
    const V3 n = cross3(tb - ta, tc - ta);
    if (!fin3(n)) return splat3(0.0f);
    return normalize_or_zero3(n);
}
// query/src/lib.rs ray_aabb_slab_unit over the box grown by r (sweep_sphere_aabb, dt = 1)
static __device__ bool rq_sweep_sphere(V3 p0, V3 dv, float r, V3 bmn, V3 bmx, float &t_out, V3 &n_out) {
    const V3 mn = bmn - splat3(r), mx = bmx + splat3(r);
    const V3 inv = mk3(p0.x == p0.x && fabsf(dv.x) > 1.0e-9f ? 1.0f / dv.x : 1.0e9f,
                       p0.y == p0.y && fabsf(dv.y) > 1.0e-9f ? 1.0f / dv.y : 1.0e9f,
                       p0.z == p0.z && fabsf(dv.z) > 1.0e-9f ? 1.0f / dv.z : 1.0e9f);
    const V3 t1 = mk3((mn.x - p0.x) * inv.x, (mn.y - p0.y) * inv.y, (mn.z - p0.z) * inv.z);
    const V3 t2 = mk3((mx.x - p0.x) * inv.x, (mx.y - p0.y) * inv.y, (mx.z - p0.z) * inv.z);
    const V3 tmin = v3min(t1, t2), tmax = v3max(t1, t2);
    float te = tmin.x;
    V3 n = mk3(t1.x > t2.x ? 1.0f : -1.0f, 0.0f, 0.0f);
    if (tmin.y > te) { te = tmin.y; n = mk3(0.0f, t1.y > t2.y ? 1.0f : -1.0f, 0.0f); }
    if (tmin.z > te) { te = tmin.z; n = mk3(0.0f, 0.0f, t1.z > t2.z ? 1.0f : -1.0f); }
    const float tx = fminf(fminf(tmax.x, tmax.y), tmax.z);
    if (te <= tx && tx >= 0.0f && te >= 0.0f && te <= 1.0f) { t_out = te; n_out = n; return true; }
    return false;
}
// sweep_two_spheres_aabbs over one box: the tip's hit, or the base's when earlier by > 1e-9
static __device__ bool rq_sweep_two(V3 tip0, V3 base0, V3 v, float r, V3 mn, V3 mx, float &t, V3 &n, int &kind) {
    float tt;
    V3 nn;
    bool have = false;
    if (rq_sweep_sphere(tip0, v, r, mn, mx, tt, nn)) { have = true; t = tt; n = nn; kind = 0; }
    if (rq_sweep_sphere(base0, v, r, mn, mx, tt, nn) && (!have || tt < t - 1.0e-9f)) { have = true; t = tt; n = nn; kind = 1; }
    return have;
}

struct RqBest { uint32_t have, idx; float t; V3 p, n; };
// lane 0 folds this chunk's candidates (in index order) into the best: a candidate replaces it
// when earlier by more than 1e-9 (the index tie rule never fires, indices ascend)
static __device__ void rq_fold(RqBest &b, const uint32_t *flag, const float *t, const float4 *p, const float4 *n, uint32_t cnt,
                               uint32_t idx0, bool terrain) {
    for (uint32_t k = 0; k < cnt; k++) {
        if (!flag[k]) continue;
        if (!b.have || t[k] < b.t - 1.0e-9f) {
            b.have = 1; b.t = t[k]; b.idx = terrain ? RP_TERRAIN : idx0 + k; b.p = L3(p[k]); b.n = L3(n[k]);
        }
    }
}

__global__ void __launch_bounds__(RPQ_T) k_query(Dev d, const RpQuery *qs, RpQueryHit *out) {
    __shared__ uint32_t flag[2 * RPQ_T];
    __shared__ float tt[2 * RPQ_T];
    __shared__ float4 pp[2 * RPQ_T], nn[2 * RPQ_T];
    __shared__ RqBest best;
    __shared__ uint32_t first;
    const RpQuery q = qs[blockIdx.x];
    const uint32_t lane = threadIdx.x, w = q.world;
    const uint32_t c0 = d.w_col_off[w], c1 = d.w_col_off[w + 1];
    const HfDesc hf = d.hf[w];
    if (lane == 0) { best.have = 0; best.idx = RP_NONE; best.t = 0.0f; first = RP_NONE; }
    __syncthreads();
    const V3 A = mk3(q.a[0], q.a[1], q.a[2]), B3 = mk3(q.b[0], q.b[1], q.b[2]);
    if (q.kind == 0) {
        // World::raycast (queries.rs:50-116): colliders' boxes, then the terrain's triangles
        const float max_d = q.max_d;
        for (uint32_t base = c0; base < c1; base += RPQ_T) {
            const uint32_t c = base + lane;
            flag[lane] = 0;
            if (c < c1) {
                const uint32_t b = d.col_body[c];
                if (b != q.ignore) {
                    V3 mn, mx, p, n;
                    float t;
                    aabb_of(d.btype[b], d.bdims[b], L3(d.pos[b]), LQ(d.rot[b]), mn, mx);
                    if (rq_ray_aabb(A, B3, mn, mx, max_d, t, p, n)) { flag[lane] = 1; tt[lane] = t; pp[lane] = S3(p); nn[lane] = S3(n); }
                }
            }
            __syncthreads();
            if (lane == 0) rq_fold(best, flag, tt, pp, nn, min(RPQ_T, c1 - base), base, false);
            __syncthreads();
        }
        const V3 e = A + B3 * max_d;
        const V3 qa = v3min(A, e), qb = v3max(A, e);
        V3 mn = v3min(mk3(qf32(qa.x), qf32(qa.y), qf32(qa.z)), mk3(qf32(qb.x), qf32(qb.y), qf32(qb.z)));
        V3 mx = v3max(mk3(qf32(qa.x), qf32(qa.y), qf32(qa.z)), mk3(qf32(qb.x), qf32(qb.y), qf32(qb.z)));
        int gx0, gx1, gz0, gz1;
        if (hf_range(hf, mn, mx, gx0, gx1, gz0, gz1)) {
            const uint32_t ncx = (uint32_t)(gx1 - gx0 + 1), ncell = ncx * (uint32_t)(gz1 - gz0 + 1);
            for (uint32_t base = 0; base < ncell; base += RPQ_T) {
                const uint32_t e2 = base + lane;
                flag[2 * lane] = flag[2 * lane + 1] = 0;
                if (e2 < ncell) {
                    const int gx = gx0 + (int)(e2 % ncx), gz = gz0 + (int)(e2 / ncx);
                    for (int k = 0; k < 2; k++) {
                        V3 ta, tb, tc;
                        hf_tri(d, hf, gx, gz, k, ta, tb, tc);
                        float t;
                        if (rq_tri_ray(ta, tb, tc, A, B3, max_d, t)) {
                            flag[2 * lane + k] = 1; tt[2 * lane + k] = t;
                            pp[2 * lane + k] = S3(A + B3 * t); nn[2 * lane + k] = S3(rq_tri_normal(ta, tb, tc));
                        }
                    }
                }
                __syncthreads();
                if (lane == 0) rq_fold(best, flag, tt, pp, nn, 2 * min(RPQ_T, ncell - base), 0, true);
                __syncthreads();
            }
        }
        if (lane == 0) {
            RpQueryHit h = {};
            if (best.have) {
                h.hit = 1;
                h.body = best.idx == RP_TERRAIN ? RP_TERRAIN : d.col_body[best.idx];
                h.fraction = clampf_rs(best.t / max_d, 0.0f, 1.0f);
                h.point[0] = best.p.x; h.point[1] = best.p.y; h.point[2] = best.p.z;
                h.normal[0] = best.n.x; h.normal[1] = best.n.y; h.normal[2] = best.n.z;
            }
            out[blockIdx.x] = h;
        }
        return;
    }
    // World::sweep_capsule (queries.rs:118-261)
    const float r = q.radius, hh = q.half_height;
    const V3 v = B3 - A;
    const float dist_sq = len2(v);
    Shape cap;
    cap.type = 2; cap.d0 = r; cap.d1 = hh; cap.d2 = 0.0f;
    const Q4 qi = qid();
    // 1. overlap at the start: the first collider in order, then the first terrain triangle
    for (uint32_t base = c0; base < c1 && first == RP_NONE; base += RPQ_T) {
        const uint32_t c = base + lane;
        Manifold m;
        bool hit = false;
        if (c < c1) {
            const uint32_t b = d.col_body[c];
            if (b != q.ignore) hit = collide_pair(cap, A, qi, shape_of(d, b), L3(d.pos[b]), LQ(d.rot[b]), 0.0f, m) && m.len > 0;
        }
        if (hit) atomicMin(&first, c);
        __syncthreads();
        if (hit && first == c) {
            RpQueryHit h = {};
            h.hit = 1; h.body = d.col_body[c]; h.started = 1; h.fraction = 0.0f;
            h.point[0] = m.p[0].pos.x; h.point[1] = m.p[0].pos.y; h.point[2] = m.p[0].pos.z;
            h.normal[0] = m.p[0].n.x; h.normal[1] = m.p[0].n.y; h.normal[2] = m.p[0].n.z;
            out[blockIdx.x] = h;
        }
        __syncthreads();
    }
    if (first != RP_NONE) return;
    int gx0, gx1, gz0, gz1;
    {
        V3 mn, mx;
        aabb_of(2u, make_float4(r, hh, 0.0f, 0.0f), A, qi, mn, mx);
        if (hf_range(hf, mn, mx, gx0, gx1, gz0, gz1)) {
            const uint32_t ncx = (uint32_t)(gx1 - gx0 + 1), ncell = ncx * (uint32_t)(gz1 - gz0 + 1);
            for (uint32_t base = 0; base < ncell && first == RP_NONE; base += RPQ_T) {
                const uint32_t e2 = base + lane;
                Manifold m;
                int k_hit = -1;
                if (e2 < ncell) {
                    const int gx = gx0 + (int)(e2 % ncx), gz = gz0 + (int)(e2 / ncx);
                    for (int k = 0; k < 2 && k_hit < 0; k++) {
                        V3 ta, tb, tc;
                        hf_tri(d, hf, gx, gz, k, ta, tb, tc);
                        if (collide_terrain(cap, A, qi, ta, tb, tc, 0.0f, m) && m.len > 0) k_hit = k;
                    }
                }
                if (k_hit >= 0) atomicMin(&first, 2 * e2 + (uint32_t)k_hit);
                __syncthreads();
                if (k_hit >= 0 && first == 2 * e2 + (uint32_t)k_hit) {
                    RpQueryHit h = {};
                    h.hit = 1; h.body = RP_TERRAIN; h.started = 1; h.fraction = 0.0f;
                    h.point[0] = m.p[0].pos.x; h.point[1] = m.p[0].pos.y; h.point[2] = m.p[0].pos.z;
                    h.normal[0] = m.p[0].n.x; h.normal[1] = m.p[0].n.y; h.normal[2] = m.p[0].n.z;
                    out[blockIdx.x] = h;
                }
                __syncthreads();
            }
        }
    }
    if (first != RP_NONE) return;
    if (dist_sq < 1.0e-12f) {
        if (lane == 0) { RpQueryHit h = {}; out[blockIdx.x] = h; }
        return;
    }
    // 2. the tip and base spheres swept against the colliders' boxes, then the triangles' boxes
    const V3 tip0 = A + mk3(0.0f, hh, 0.0f), base0 = A + mk3(0.0f, -hh, 0.0f);
    for (uint32_t base = c0; base < c1; base += RPQ_T) {
        const uint32_t c = base + lane;
        flag[lane] = 0;
        if (c < c1) {
            const uint32_t b = d.col_body[c];
            if (b != q.ignore) {
                V3 mn, mx, n;
                float t;
                int kind;
                aabb_of(d.btype[b], d.bdims[b], L3(d.pos[b]), LQ(d.rot[b]), mn, mx);
                if (rq_sweep_two(tip0, base0, v, r, mn, mx, t, n, kind)) {
                    t = clampf_rs(t, 0.0f, 1.0f);
                    flag[lane] = 1; tt[lane] = t; nn[lane] = S3(n);
                    pp[lane] = S3(((kind == 0 ? tip0 : base0) + v * t) - n * r);
                }
            }
        }
        __syncthreads();
        if (lane == 0) rq_fold(best, flag, tt, pp, nn, min(RPQ_T, c1 - base), base, false);
        __syncthreads();
    }
    {
        const V3 ext = splat3(r + hh);
        const V3 a0 = v3min(A, B3) - ext, a1 = v3max(A, B3) + ext;
        const V3 mn = v3min(mk3(qf32(a0.x), qf32(a0.y), qf32(a0.z)), mk3(qf32(a1.x), qf32(a1.y), qf32(a1.z)));
        const V3 mx = v3max(mk3(qf32(a0.x), qf32(a0.y), qf32(a0.z)), mk3(qf32(a1.x), qf32(a1.y), qf32(a1.z)));
        if (hf_range(hf, mn, mx, gx0, gx1, gz0, gz1)) {
            const uint32_t ncx = (uint32_t)(gx1 - gx0 + 1), ncell = ncx * (uint32_t)(gz1 - gz0 + 1);
            for (uint32_t base = 0; base < ncell; base += RPQ_T) {
                const uint32_t e2 = base + lane;
                flag[2 * lane] = flag[2 * lane + 1] = 0;
                if (e2 < ncell) {
                    const int gx = gx0 + (int)(e2 % ncx), gz = gz0 + (int)(e2 / ncx);
                    for (int k = 0; k < 2; k++) {
                        V3 ta, tb, tc;
                        hf_tri(d, hf, gx, gz, k, ta, tb, tc);
                        // aabb_of(Triangle, identity): the vertices through the identity pose, quantised
                        const V3 wa = splat3(0.0f) + qrot(qi, ta), wb = splat3(0.0f) + qrot(qi, tb), wc = splat3(0.0f) + qrot(qi, tc);
                        const V3 lo = v3min(v3min(wa, wb), wc), hi = v3max(v3max(wa, wb), wc);
                        const V3 qa = mk3(qf32(lo.x), qf32(lo.y), qf32(lo.z)), qb = mk3(qf32(hi.x), qf32(hi.y), qf32(hi.z));
                        V3 n;
                        float t;
                        int kind;
                        if (rq_sweep_two(tip0, base0, v, r, v3min(qa, qb), v3max(qa, qb), t, n, kind)) {
                            t = clampf_rs(t, 0.0f, 1.0f);
                            flag[2 * lane + k] = 1; tt[2 * lane + k] = t; nn[2 * lane + k] = S3(n);
                            pp[2 * lane + k] = S3(((kind == 0 ? tip0 : base0) + v * t) - n * r);
                        }
                    }
                }
                __syncthreads();
                if (lane == 0) rq_fold(best, flag, tt, pp, nn, 2 * min(RPQ_T, ncell - base), 0, true);
                __syncthreads();
            }
        }
    }
    if (lane == 0) {
        RpQueryHit h = {};
        if (best.have) {
            h.hit = 1;
            h.body = best.idx == RP_TERRAIN ? RP_TERRAIN : d.col_body[best.idx];
            h.fraction = clampf_rs(best.t, 0.0f, 1.0f);
            h.point[0] = best.p.x; h.point[1] = best.p.y; h.point[2] = best.p.z;
            h.normal[0] = best.n.x; h.normal[1] = best.n.y; h.normal[2] = best.n.z;
        }
        out[blockIdx.x] = h;
    }
}
