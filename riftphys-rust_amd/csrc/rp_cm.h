// Contact models on the device (materials.rs:75-90): compliant Hunt-Crossley rows and friction
// patches (solver.rs:219-260, 545-567, 592-619, 763-782, 848-915, 927-944), active in scenes with a
// pair whose model is compliant or patch friction (d.cm_on).
//
// Patches.  A patch is a run of two or more contacts of one patch-friction collider pair whose
// normals do not cancel; contacts are in (collider a, collider b) order, terrain last, so runs are
// contiguous and the patches come out in the reference's map order.  A patch is an island-list
// entry (RP_PBIT | patch) after its island's contacts: in each sweep it runs after every contact of
// its bodies, as the reference's patch pass follows the contact batches.  Its row is the first
// point's constraint moved to the points' centroid along their mean normal, with the mean radius
// of the patch as the torsion lever.
//
// Islands holding a compliant row or a patch are solved by tier 4 (k_sweep_cm): one workgroup per
// island, the list order as dataflow over the body chains (level order beyond the workgroup),
// velocities in HBM.  The other tiers never see these entries.  A patch whose first point writes
// no body is not in an island: k_free_patches runs it (it moves nothing; its accumulators are
// carried).  Patches carry (jt, tw) between ticks in a second cache beside the warm cache.
#pragma once

// ---- patch runs
static __device__ __forceinline__ bool cm_patch_pair(const Dev &d, uint32_t c) {
    const uint32_t ca = d.ct_a[c], cb = d.ct_b[c];
    const uint32_t ma = d.col_mat[ca], mb = cb == RP_TERRAIN ? d.hf[d.col_world[ca]].mat : d.col_mat[cb];
    return d.pair_cm[(ma & 15u) * 16u + (mb & 15u)].z != 0.0f;
}
static __device__ __forceinline__ bool same_pair(const Dev &d, uint32_t c, uint32_t e) { return d.ct_a[c] == d.ct_a[e] && d.ct_b[c] == d.ct_b[e]; }
// The run of contact c (its first contact h and length n; a pair holds at most 4 contacts), and
// whether it is a patch: two or more points whose summed normals normalise to a nonzero vector
// (solver.rs:597-601).
static __device__ __forceinline__ bool patch_run(const Dev &d, uint32_t c, uint32_t nct, uint32_t &h, uint32_t &n) {
    h = c;
    while (h > 0 && same_pair(d, h - 1, c)) h--;
    uint32_t e = c + 1;
    while (e < nct && same_pair(d, e, c)) e++;
    n = e - h;
    if (n < 2 || !cm_patch_pair(d, c)) return false;
    V3 ns = splat3(0.0f);
    for (uint32_t k = h; k < e; k++) ns = ns + L3(d.ct_n[k]);
    return !zero3(normalize_or_zero3(ns));
}
// Patch starts, then (after their exclusive scan) each contact's patch and each patch's points.
__global__ void k_patch_heads(Dev d) {
    if (aborted(d)) return;
    const uint32_t nct = min(d.cnt[CT_CONTACTS], d.cap_contacts);
    GRID_STRIDE(c, d.cap_contacts) {
        uint32_t h, n, f = 0;
        if (c < nct && patch_run(d, c, nct, h, n)) f = h == c ? 1u : 0u;
        d.pt_head[c] = f;
    }
}
__global__ void k_patch_fill(Dev d) {
    if (aborted(d)) return;
    const uint32_t nct = min(d.cnt[CT_CONTACTS], d.cap_contacts);
    GRID_STRIDE(c, nct) {
        uint32_t h, n, p = RP_NONE;
        if (patch_run(d, c, nct, h, n)) {
            p = d.pt_head[h];
            if (h == c) { d.pt_first[p] = c; d.pt_n[p] = n; }
        }
        d.ct_patch[c] = p;
    }
}
// the island lists hold the patches after the joints and contacts
__global__ void k_rows_patches(Dev d) {
    if (threadIdx.x != 0 || blockIdx.x != 0 || aborted(d)) return;
    d.cnt[CT_ROWS] += d.cnt[CT_NPATCH];
}

// solver.rs:233-260 patch_radius over the n (<= 4) plane points: sorted by sort_key6 of their
// angle (stable), each fan triangle's mean distance by the 6 x 6 centroid rule; collinear points:
// a quarter of their longest distance.
static __device__ float patch_radius(float *px, float *py, uint32_t n) {
    if (n < 2) return 0.0f;
    int64_t key[4];
    for (uint32_t i = 0; i < n; i++) key[i] = sort_key6(rp_atan2f(py[i], px[i]));
    for (uint32_t x = 1; x < n; x++) {
        const float vx = px[x], vy = py[x];
        const int64_t kv = key[x];
        uint32_t y = x;
        while (y > 0 && key[y - 1] > kv) { px[y] = px[y - 1]; py[y] = py[y - 1]; key[y] = key[y - 1]; y--; }
        px[y] = vx; py[y] = vy; key[y] = kv;
    }
    float area = 0.0f, moment = 0.0f;
    const int N = 6;
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t j = i + 1 < n ? i + 1 : 0u;
        const float ax = px[i], ay = py[i], bx = px[j], by = py[j];
        const float tri = 0.5f * fabsf(ax * by - ay * bx);
        if (tri <= 0.0f) continue;
        const float sub = tri / (float)(N * N);
        for (int r = 0; r < N; r++)
            for (int c = 0; c < N - r; c++)
                for (int up = 1; up >= 0; up--) {
                    if (!up && c + r + 1 >= N) continue;
                    float u = up ? (float)r + 1.0f / 3.0f : (float)r + 2.0f / 3.0f;
                    float v = up ? (float)c + 1.0f / 3.0f : (float)c + 2.0f / 3.0f;
                    u = u / (float)N; v = v / (float)N;
                    const float x0 = ax * u + bx * v, x1 = ay * u + by * v;
                    moment += sqrtf(x0 * x0 + x1 * x1) * sub;
                }
        area += tri;
    }
    if (area > 1.0e-12f) return moment / area;
    float len = 0.0f;
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t j = 0; j < n; j++) {
            const float dx = px[i] - px[j], dy = py[i] - py[j];
            len = fmaxf(len, sqrtf(dx * dx + dy * dy));
        }
    return 0.25f * len;
}
// solver.rs:596-617: each patch's row from its first point's constraint (as the prep left it)
__global__ void k_patch_prep(Dev d) {
    if (aborted(d)) return;
    GRID_STRIDE(p, d.cnt[CT_NPATCH]) {
        const uint32_t f = d.pt_first[p], n = min(d.pt_n[p], 4u);
        V3 sum = splat3(0.0f), ns = splat3(0.0f);
        for (uint32_t k = 0; k < n; k++) sum = sum + L3(d.ct_p[f + k]);
        const V3 cen = mk3(sum.x / (float)n, sum.y / (float)n, sum.z / (float)n);
        for (uint32_t k = 0; k < n; k++) ns = ns + L3(d.ct_n[f + k]);
        const V3 nn = normalize_or_zero3(ns);
        V3 t1, t2;
        orthonormal_basis(nn, t1, t2);
        float px[4], py[4];
        for (uint32_t k = 0; k < n; k++) {
            const V3 dd = L3(d.ct_p[f + k]) - cen;
            px[k] = dot3(dd, t1); py[k] = dot3(dd, t2);
        }
        const uint32_t ai = d.cs_ai[f], bi = d.cs_bi[f];
        const float2 inv = d.cs_inv[f];
        const M3 iia = inv.x > 0.0f ? load_m9(d.iiw + 9 * ai) : mzero();
        const M3 iib = inv.y > 0.0f ? load_m9(d.iiw + 9 * bi) : mzero();
        const V3 ra = cen - L3(d.pos[ai]), rb = bi == RP_NONE ? splat3(0.0f) : cen - L3(d.pos[bi]);
        const float kt = dot3(nn, mv(iia, nn)) + dot3(nn, mv(iib, nn));
        d.pt_ai[p] = ai; d.pt_bi[p] = bi; d.pt_inv[p] = inv;
        d.pt_ra[p] = S3(ra); d.pt_rb[p] = S3(rb); d.pt_nn[p] = S3(nn); d.pt_t1[p] = S3(t1); d.pt_t2[p] = S3(t2);
        d.pt_m[p] = make_float4(eff_mass(inv.x, inv.y, iia, iib, ra, rb, nn), eff_mass(inv.x, inv.y, iia, iib, ra, rb, t1),
                                eff_mass(inv.x, inv.y, iia, iib, ra, rb, t2), kt > 1.0e-9f ? 1.0f / kt : 0.0f);
        const float4 mu = d.cs_mu[f];
        d.pt_mu[p] = make_float4(mu.x, mu.y, mu.z, patch_radius(px, py, n));
        d.pt_acc[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        d.pt_key[p] = d.ct_key[f];
        d.pt_soft[p] = (cm_flags(d, f) & CM_SOFT) ? 1u : 0u;
    }
}
// Islands with a compliant row, a patch or a D6 joint (tier 4): flagged from their sorted entries.
__global__ void k_isl_cm(Dev d, const uint32_t *keys) {
    if (aborted(d)) return;
    const uint32_t nislc = min(d.cnt[CT_ISLC], d.cap_contacts);
    GRID_STRIDE(r, nislc) {
        const uint32_t e = d.isl_v2[r];
        const bool cm = (e & RP_PBIT) || (!(e & RP_JBIT) && (cm_flags(d, e) & CM_SOFT)) || ((e & RP_JBIT) && jx_is(d, e & ~RP_JBIT));
        if (!cm) continue;
        const bool start = r == 0 || keys[r - 1] != keys[r];
        const uint32_t isl = start ? d.isl_flag[r] : d.isl_flag[r] - 1u;  // exclusive scan of the starts
        if (!d.isl_cm[isl]) d.isl_cm[isl] = 1u;
    }
}

// ---- the rows
// torsion impulse tw about the normal: -n tw on a, +n tw on b (solver.rs:778-781, 908-911), each as
// apply_angular_impulse
static __device__ __forceinline__ void patch_twist(const Dev &d, uint32_t ai, uint32_t bi, float2 inv, V3 n, float tw) {
    if (inv.x > 0.0f) {
        const V3 tau = (-n) * tw;
        if (d.inv_mass[ai] > 0.0f && fin3(tau)) d.ang[ai] = S3(fin_v(L3(d.ang[ai]) + mv(load_m9(d.iiw + 9 * ai), tau)));
    }
    if (bi != RP_NONE && inv.y > 0.0f) {
        const V3 tau = n * tw;
        if (d.inv_mass[bi] > 0.0f && fin3(tau)) d.ang[bi] = S3(fin_v(L3(d.ang[bi]) + mv(load_m9(d.iiw + 9 * bi), tau)));
    }
}
static __device__ __forceinline__ float clamp_rs(float x, float lo, float hi) { if (x < lo) x = lo; if (x > hi) x = hi; return x; }
// the points' normal impulses summed in order (Iterator::sum, from -0)
static __device__ __forceinline__ float patch_jn(const Dev &d, uint32_t p) {
    const uint32_t f = d.pt_first[p], n = d.pt_n[p];
    float s = -0.0f;
    for (uint32_t k = 0; k < n; k++) s += d.jn[f + k];
    return s;
}
// solver.rs:763-782: the patch's carried sliding and torsion impulses, limited by the points' warm
// normal load (compliant patches start cold)
static __device__ void patch_warm(const Dev &d, const TickParams &tp, uint32_t p) {
    if (d.pt_soft[p]) return;
    const uint32_t cur = tp.cur, w = d.col_world[d.ct_a[d.pt_first[p]]];
    const uint64_t key = d.pt_key[p];
    uint32_t lo = d.wpc_off[cur][w], hi = d.wpc_off[cur][w + 1];
    while (lo < hi) { const uint32_t mid = (lo + hi) >> 1; if (d.wp_key[cur][mid] < key) lo = mid + 1; else hi = mid; }
    if (lo >= d.wpc_off[cur][w + 1] || d.wp_key[cur][lo] != key) return;
    const float4 e = d.wp_jt[cur][lo];
    const uint32_t ai = d.pt_ai[p], bi = d.pt_bi[p];
    const float2 inv = d.pt_inv[p];
    const V3 ra = L3(d.pt_ra[p]), rb = L3(d.pt_rb[p]), n = L3(d.pt_nn[p]), t1 = L3(d.pt_t1[p]), t2 = L3(d.pt_t2[p]);
    const float4 mu4 = d.pt_mu[p];
    const float js = patch_jn(d, p);
    const V3 v = rel_vel(d, ai, bi, ra, rb);
    const float vt1 = dot3(v, t1), vt2 = dot3(v, t2);
    const float mu = mu_dyn(make_float4(mu4.x, mu4.y, mu4.z, 0.0f), sqrtf(vt1 * vt1 + vt2 * vt2));
    const V3 ej = L3(e);
    const V3 t = clamp_len_max(mk3(dot3(ej, t1), dot3(ej, t2), 0.0f), mu * js);
    const float lim = (mu * js) * mu4.w;
    const float tw = clamp_rs(e.w, -lim, lim);
    d.pt_acc[p] = make_float4(t.x, t.y, tw, 0.0f);
    apply_cimp(d, ai, bi, inv.x, inv.y, ra, rb, t1 * t.x + t2 * t.y);
    if (tw != 0.0f) patch_twist(d, ai, bi, inv, n, tw);
}
// solver.rs:887-913: one pass of the patch (sliding at the centroid, then torsion about the
// normal, both limited by the points' current normal impulse).  Returns the residual-mode
// residual's bits: the largest of |d jt| / m_t and |d tw| / m_tw.
static __device__ uint32_t patch_pgs(const Dev &d, uint32_t p) {
    const uint32_t ai = d.pt_ai[p], bi = d.pt_bi[p];
    const float2 inv = d.pt_inv[p];
    const V3 ra = L3(d.pt_ra[p]), rb = L3(d.pt_rb[p]), n = L3(d.pt_nn[p]), t1 = L3(d.pt_t1[p]), t2 = L3(d.pt_t2[p]);
    const float4 m = d.pt_m[p], mu4 = d.pt_mu[p];
    const float js = patch_jn(d, p);
    const V3 v = rel_vel(d, ai, bi, ra, rb);
    const float vt1 = dot3(v, t1), vt2 = dot3(v, t2);
    const float mu = mu_dyn(make_float4(mu4.x, mu4.y, mu4.z, 0.0f), sqrtf(vt1 * vt1 + vt2 * vt2));
    float4 acc = d.pt_acc[p];
    const V3 ot = mk3(acc.x, acc.y, 0.0f);
    const V3 nt = clamp_len_max(ot - mk3(vt1 * m.y, vt2 * m.z, 0.0f), mu * js);
    acc.x = nt.x; acc.y = nt.y;
    const V3 dd = nt - ot;
    apply_cimp(d, ai, bi, inv.x, inv.y, ra, rb, t1 * dd.x + t2 * dd.y);
    const V3 wa = inv.x > 0.0f ? L3(d.ang[ai]) : splat3(0.0f);
    const V3 wb = (bi == RP_NONE || inv.y <= 0.0f) ? splat3(0.0f) : L3(d.ang[bi]);
    const float w_rel = dot3(wb - wa, n);
    const float old = acc.z, lim = (mu * js) * mu4.w;
    const float ntw = clamp_rs(old - w_rel * m.w, -lim, lim);
    acc.z = ntw;
    d.pt_acc[p] = acc;
    const float dtw = ntw - old;
    if (dtw != 0.0f) patch_twist(d, ai, bi, inv, n, dtw);
    const float r1 = m.y > 0.0f ? fabsf(dd.x) / m.y : 0.0f;
    const float r2 = m.z > 0.0f ? fabsf(dd.y) / m.z : 0.0f;
    const float r3 = m.w > 0.0f ? fabsf(dtw) / m.w : 0.0f;
    return __float_as_uint(fmaxf(r1, fmaxf(r2, r3)));
}
// A contact of tier 4 (solver.rs:839-879): the rigid normal row (entry_core's operations), the
// implicit compliant row (solver.rs:848-856, applied only when it changes) or none (a compliant
// row without one keeps its zero); a patch's point has no friction row (solver.rs:865).  Returns
// the residual's bits (the compliant row's own mass and target in place of the rigid ones).
static __device__ uint32_t cm_contact_pgs(const Dev &d, uint32_t c) {
    const Row x = load_row(d, c);
    const float4 sp = d.cs_soft[c];
    const uint32_t cmf = __float_as_uint(sp.w);
    Vel2 v = load_vel2(d, x);
    const Vel2 v0 = v;
    const float old0 = d.jn[c];
    float nw0 = old0;
    if (cmf & CM_IMPLICIT) {
        const float vn = dot3(rel_vel2(x, v), x.n);
        nw0 = fmaxf(old0 + sp.z * ((sp.y - vn) - sp.x * old0), 0.0f);
        if (nw0 != old0) apply_row2(x, v, x.n * (nw0 - old0));
    } else if (!(cmf & CM_SOFT)) {
        const float vn = dot3(rel_vel2(x, v), x.n);
        nw0 = fmaxf(old0 + (x.m4.w - vn) * x.m4.x, 0.0f);
        apply_row2(x, v, x.n * (nw0 - old0));
    }
    d.jn[c] = nw0;
    float c1 = 0.0f, c2 = 0.0f;
    if (!(cmf & CM_PATCH_PT)) {
        const V3 vf = rel_vel2(x, v);
        const float vt1 = dot3(vf, x.t1), vt2 = dot3(vf, x.t2);
        const float mu = mu_dyn(x.mu4, sqrtf(vt1 * vt1 + vt2 * vt2));
        const float2 jt = d.jt[c];
        const V3 ot = mk3(jt.x, jt.y, 0.0f);
        const V3 nt = clamp_len_max(ot - mk3(vt1 * x.m4.y, vt2 * x.m4.z, 0.0f), mu * nw0);
        d.jt[c] = make_float2(nt.x, nt.y);
        const V3 dd = nt - ot;
        c1 = dd.x; c2 = dd.y;
        apply_row2(x, v, x.t1 * dd.x + x.t2 * dd.y);
    }
    store_vel2(d, x, v, v0);
    const bool impl = (cmf & CM_IMPLICIT) != 0u;
    const float rm = impl ? sp.z : x.m4.x, rt = impl ? sp.y : x.m4.w;
    const float r0 = rm > 0.0f ? (fabsf(nw0 - old0) / rm) / (1.0f + fabsf(rt)) : 0.0f;
    const float r1 = x.m4.y > 0.0f ? fabsf(c1) / x.m4.y : 0.0f;
    const float r2 = x.m4.z > 0.0f ? fabsf(c2) / x.m4.z : 0.0f;
    return __float_as_uint(fmaxf(r0, fmaxf(r1, r2)));
}
static __device__ __forceinline__ void cm_warm(const Dev &d, const TickParams &tp, uint32_t e) {
    if (e & RP_PBIT) patch_warm(d, tp, e & ~RP_PBIT);
    else if ((e & RP_JBIT) && d6_is(d, e & ~RP_JBIT)) d6_warm(d, e & ~RP_JBIT);
    else if ((e & RP_JBIT) && dist_is(d, e & ~RP_JBIT)) return;  // a distance row starts cold
    else warm_row(d, tp, e);
}
static __device__ __forceinline__ void cm_pgs(const Dev &d, const TickParams &tp, uint32_t e, ResidSink sink) {
    uint32_t r;
    if ((e & RP_JBIT) && d6_is(d, e & ~RP_JBIT)) r = d6_pgs(d, e & ~RP_JBIT, sink.on);
    else if ((e & RP_JBIT) && dist_is(d, e & ~RP_JBIT)) r = dist_vel(d, e & ~RP_JBIT, tp.dt);
    else if (e & RP_JBIT) r = joint_pgs(d, e & ~RP_JBIT, sink.on);
    else if (e & RP_PBIT) r = patch_pgs(d, e & ~RP_PBIT);
    else r = cm_contact_pgs(d, e);
    if (sink.on) __hip_atomic_fetch_max(sink.p, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Tier 4's velocity solve, one island per workgroup: the warm start, the PGS sweeps (residual mode
// as the other tiers), the pre-shock snapshot and the shock pass, in the list order's dataflow.
__global__ void __launch_bounds__(256) k_sweep_cm(Dev d, TickParams tp) {
    if (aborted(d)) return;
    constexpr int TIER = 4;
    ISLAND_LOOP_BEGIN
        PhaseClock pc;
        const uint32_t *exb = n > blockDim.x ? level_order(I, d.ch_pa, d.ch_pb, d.ch_sa, d.ch_sb, d.ex_base, &d.ph[RP_PH_ROUNDS_BASE],
                                                           tp.wide_min, nullptr, nullptr)
                                              : nullptr;
        const Slots sb{d.bs_pa, d.bs_pb, d.bs_c};
        build_slots(d, I, d.isl_v2, d.ch_pa, d.ch_pb, exb, sb);
        if (tp.ws_mode != 1) run_sweep(I, sb, [&](uint32_t e) { cm_warm(d, tp, e); });
        __shared__ uint32_t resid_bits;
        const ResidSink rb_lds{(RP_LDS uint32_t *)&resid_bits, tp.resid_on != 0};
        uint32_t used = 0;
        bool conv = false;
        for (uint32_t it = 0; it < tp.iv; it++) {
            if (tp.resid_on) {
                if (threadIdx.x == 0) resid_bits = 0;
                __syncthreads();
            }
            run_sweep(I, sb, [&](uint32_t e) { cm_pgs(d, tp, e, rb_lds); });
            used++;
            if (tp.resid_on) {
                const float rmax = __uint_as_float(resid_bits);
                __syncthreads();
                if (rmax <= tp.resid_tol) { conv = true; break; }
            }
        }
        if (threadIdx.x == 0) {
            atomicMax(&d.cnt[CT_ITERS_MAX], used); atomicAdd(&d.cnt[CT_ITERS_SUM], used * n);
            if (conv) { atomicAdd(&d.cnt[CT_ISL_CONV], 1u); atomicAdd(&d.cnt[CT_ISLC_CONV], n); atomicAdd(&d.cnt[CT_ITERS_SUM_CONV], used * n); }
        }
        if (tp.resid_on)
            for (uint32_t r = threadIdx.x; r < n; r += blockDim.x) d.rank_conv[s + r] = conv ? 1u : 0u;
        vel_shock<TIER>(d, tp, I, lst, isl, pc, 0u);
    ISLAND_LOOP_END
}
// Patches whose first point writes no body (in no island): their warm start and the sweeps'
// passes in order; they move nothing, so each runs alone.
__global__ void k_free_patches(Dev d, TickParams tp) {
    if (aborted(d)) return;
    GRID_STRIDE(p, d.cnt[CT_NPATCH]) {
        const uint32_t f = d.pt_first[p];
        if (d.cs_wa[f] != RP_NONE || d.cs_wb[f] != RP_NONE || !d.w_solve[d.col_world[d.ct_a[f]]]) continue;
        if (tp.ws_mode != 1) patch_warm(d, tp, p);
        for (uint32_t it = 0; it < tp.iv; it++) (void)patch_pgs(d, p);
    }
}
// solver.rs:1257-1264: a patch with a nonzero impulse loads its points' bodies
__global__ void k_patch_loaded(Dev d) {
    if (aborted(d)) return;
    GRID_STRIDE(p, d.cnt[CT_NPATCH]) {
        const float4 a = d.pt_acc[p];
        if (a.x == 0.0f && a.y == 0.0f && a.z == 0.0f) continue;
        if (!d.w_solve[d.col_world[d.ct_a[d.pt_first[p]]]]) continue;
        for (uint32_t k = 0; k < d.pt_n[p]; k++) {
            const uint32_t c = d.pt_first[p] + k;
            atomicOr(&d.flags[d.cs_ai[c]], FL_LOADED);
            if (d.cs_bi[c] != RP_NONE) atomicOr(&d.flags[d.cs_bi[c]], FL_LOADED);
        }
    }
}

// ---- carried patch impulses (solver.rs:927-944): a world's new cache is its rigid patches'
// (jt, tw) by their first point's key (patch order is key order) merged with the old entries of
// idle pairs that have no new one; worlds that solved nothing keep theirs.
static __device__ __forceinline__ uint32_t patch_world(const Dev &d, uint32_t p) { return d.col_world[d.ct_a[d.pt_first[p]]]; }
// the world's patches: [pt_head[c0], pt_head[c1]) over its contacts [c0, c1)
static __device__ __forceinline__ void world_patches(const Dev &d, uint32_t w, uint32_t &p0, uint32_t &p1) {
    uint32_t c0, c1;
    world_contacts(d, w, c0, c1);
    p0 = d.pt_head[c0]; p1 = d.pt_head[c1];
}
__global__ void k_wp_newflags(Dev d) {
    if (aborted(d)) return;
    const uint32_t np = d.cnt[CT_NPATCH];
    GRID_STRIDE(p, d.cap_patch) d.wp_new[p] = (p < np && !d.pt_soft[p] && d.w_solve[patch_world(d, p)]) ? 1u : 0u;
}
__global__ void k_wp_keepflags(Dev d, TickParams tp) {
    if (aborted(d)) return;
    const uint32_t cur = tp.cur, tot = d.wpc_off[cur][d.nw];
    GRID_STRIDE(e, d.cap_warm) {
        uint32_t f = 0;
        if (e < tot) {
            const uint32_t w = upper_bound_u32(d.wpc_off[cur], d.nw, e) - 1;
            if (!d.w_solve[w]) f = 1;
            else {
                const uint64_t k = d.wp_key[cur][e];
                if (idle_body(d, w, (uint32_t)(k >> 32)) && idle_body(d, w, (uint32_t)k)) {
                    uint32_t p0, p1;
                    world_patches(d, w, p0, p1);
                    bool has_new = false;
                    for (uint32_t p = lower_bound_key(d.pt_key, p0, p1, k); p < p1 && d.pt_key[p] == k; p++) has_new |= d.wp_new[p + 1] != d.wp_new[p];
                    f = has_new ? 0u : 1u;
                }
            }
        }
        d.wp_keep[e] = f;
    }
}
__global__ void k_wp_merge(Dev d, TickParams tp) {
    if (aborted(d)) return;
    const uint32_t cur = tp.cur, nxt = cur ^ 1u, np = d.cnt[CT_NPATCH], tot = d.wpc_off[cur][d.nw];
    GRID_STRIDE(p, np) {
        if (d.wp_new[p + 1] == d.wp_new[p]) continue;
        const uint32_t w = patch_world(d, p);
        const uint64_t k = d.pt_key[p];
        const uint32_t e = lower_bound_key(d.wp_key[cur], d.wpc_off[cur][w], d.wpc_off[cur][w + 1], k);
        const uint32_t o = d.wp_new[p] + d.wp_keep[e];
        if (o >= d.cap_warm) continue;
        const float4 a = d.pt_acc[p];
        d.wp_key[nxt][o] = k;
        d.wp_jt[nxt][o] = S3(L3(d.pt_t1[p]) * a.x + L3(d.pt_t2[p]) * a.y, a.z);
    }
    GRID_STRIDE(e, tot) {
        if (d.wp_keep[e + 1] == d.wp_keep[e]) continue;
        const uint32_t w = upper_bound_u32(d.wpc_off[cur], d.nw, e) - 1;
        const uint64_t k = d.wp_key[cur][e];
        uint32_t p0, p1;
        world_patches(d, w, p0, p1);
        const uint32_t o = d.wp_keep[e] + d.wp_new[lower_bound_key(d.pt_key, p0, p1, k)];
        if (o >= d.cap_warm) continue;
        d.wp_key[nxt][o] = k;
        d.wp_jt[nxt][o] = d.wp_jt[cur][e];
    }
    GRID_STRIDE(w, d.nw + 1) {
        uint32_t p0 = np, p1;
        if (w < d.nw) world_patches(d, w, p0, p1);
        d.wpc_off[nxt][w] = d.wp_new[p0] + d.wp_keep[d.wpc_off[cur][w]];
    }
}
