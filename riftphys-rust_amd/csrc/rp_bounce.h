// W6 restitution on the device: the restitution relax pass and the time-of-impact bounce passes
// of solver.rs:995-1245 (velocity pass: bounce rows plus support rows solved as one push-only LCP
// per connected component by a primal-dual interior point on a sparse Cholesky factor in a fixed
// minimum-degree order, bounce_lcp.rs; pseudo pass: K x = u per component, summed per body in
// f64), for the worlds whose prep flagged a contact with a restitution bias.
//
// Layout: one wave per world (worlds taken in turn by a fixed set of workgroups); each workgroup
// owns one workspace slot for its component systems (up to W6_DIRECT_MAX rows: K and K + CFM diag K
// in CSR, the symbolic factor, L, the interior point's vectors).  Work that the reference sums in
// a fixed order runs on lane 0 in that order (the mean complementarity, the backward substitution,
// impulse application); gathers, row products, the factor's column updates (each position gets
// its updates in column order), the forward substitution's column updates, elementwise updates,
// and order-free reductions (max, min, closures) run across the wave.  Every f64 operation is a
// correctly rounded add, multiply, divide or square root, so the bits equal the CPU order's.
//
// Included by rp_step.hip after its contact helpers (apply_cimp, rel_vel, eff_mass).
#pragma once

#define W6_NMAX 256u
#define W6_DIRECT_MAX 250u   // solver.rs:94 DIRECT_MAX
#define W6_SWEEPS 256u       // solver.rs:97 BOUNCE_SWEEPS
#define W6_TOL 1.0e-6f       // solver.rs:98 BOUNCE_TOL
#define W6_SOLVES 64u        // solver.rs:100 BOUNCE_SOLVES
#define W6_CFM 1.0e-12       // bounce_lcp.rs:239 CFM
#define W6_KMAX (W6_NMAX * W6_NMAX)
#define W6_LMAX (W6_NMAX * (W6_NMAX - 1u) / 2u)
#define W6_NVEC 17u
// contact marks (w6_mark)
#define W6M_R 1u      // restitution bias > 0 (seen by the relax pass or the bounce)
#define W6M_T 2u      // a time-of-impact bounce row this tick
#define W6M_P 4u      // a pseudo row this tick
// body bits (w6_b[2 b + 1]); w6_b[2 b] holds a partner (rests_on) or a component slot
#define W6B_MULTI 1u
#define W6B_INV 2u
#define W6B_DIRECT 4u
#define W6_EMPTY 0xFFFFFFFEu

struct D3 { double x, y, z; };
static __device__ __forceinline__ D3 d3(double x, double y, double z) { D3 r; r.x = x; r.y = y; r.z = z; return r; }
static __device__ __forceinline__ D3 d3add(D3 a, D3 b) { return d3(a.x + b.x, a.y + b.y, a.z + b.z); }
static __device__ __forceinline__ D3 d3mul(D3 a, double s) { return d3(a.x * s, a.y * s, a.z * s); }
// glam DVec3: dot (x x' + y y') + z z', cross as Vec3
static __device__ __forceinline__ double d3dot(D3 a, D3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static __device__ __forceinline__ D3 d3cross(D3 a, D3 b) { return d3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y); }

// One workspace slot (carved from d.w6_ws; sizes in the host allocator's W6_WS_BYTES).
struct W6Ws {
    double *sd;        // [W6_NMAX][2][12] per row side: J linear, J angular, M^-1 J^T linear, angular
    uint32_t *sb;      // [W6_NMAX][2] side bodies (RP_NONE: cannot move)
    uint32_t *kptr, *kidx;
    double *k0, *k1;   // K, and K with its diagonal scaled by (1 + CFM)
    uint32_t *order, *pos, *cp, *ri, *rp, *re, *ce, *fill;
    unsigned long long *adj, *nbs;  // [W6_NMAX][4] adjacency rows of the elimination graph
    double *ldiag, *lval;
    double *v;         // W6_NVEC vectors of W6_NMAX
};
static __host__ __device__ constexpr size_t w6_ws_bytes() {
    return sizeof(double) * ((size_t)W6_NMAX * 24 + 2 * (size_t)W6_KMAX + W6_NMAX + W6_LMAX + W6_NVEC * W6_NMAX) +
           sizeof(unsigned long long) * 2 * 4 * W6_NMAX +
           sizeof(uint32_t) * (2 * W6_NMAX + (W6_NMAX + 4) + W6_KMAX + 3 * W6_NMAX + 2 * (W6_NMAX + 4) + 3 * (size_t)W6_LMAX) + 512;
}
static __device__ W6Ws w6_slot(const Dev &d, uint32_t slot) {
    char *p = (char *)d.w6_ws + (size_t)slot * w6_ws_bytes();
    W6Ws s;
    auto take = [&](size_t bytes) { char *q = p; p += (bytes + 15) & ~(size_t)15; return q; };
    s.sd = (double *)take(sizeof(double) * W6_NMAX * 24);
    s.k0 = (double *)take(sizeof(double) * W6_KMAX);
    s.k1 = (double *)take(sizeof(double) * W6_KMAX);
    s.ldiag = (double *)take(sizeof(double) * W6_NMAX);
    s.lval = (double *)take(sizeof(double) * W6_LMAX);
    s.v = (double *)take(sizeof(double) * W6_NVEC * W6_NMAX);
    s.adj = (unsigned long long *)take(sizeof(unsigned long long) * 4 * W6_NMAX);
    s.nbs = (unsigned long long *)take(sizeof(unsigned long long) * 4 * W6_NMAX);
    s.sb = (uint32_t *)take(sizeof(uint32_t) * 2 * W6_NMAX);
    s.kptr = (uint32_t *)take(sizeof(uint32_t) * (W6_NMAX + 4));
    s.kidx = (uint32_t *)take(sizeof(uint32_t) * W6_KMAX);
    s.order = (uint32_t *)take(sizeof(uint32_t) * W6_NMAX);
    s.pos = (uint32_t *)take(sizeof(uint32_t) * W6_NMAX);
    s.cp = (uint32_t *)take(sizeof(uint32_t) * (W6_NMAX + 4));
    s.rp = (uint32_t *)take(sizeof(uint32_t) * (W6_NMAX + 4));
    s.fill = (uint32_t *)take(sizeof(uint32_t) * W6_NMAX);
    s.ri = (uint32_t *)take(sizeof(uint32_t) * W6_LMAX);
    s.re = (uint32_t *)take(sizeof(uint32_t) * W6_LMAX);
    s.ce = (uint32_t *)take(sizeof(uint32_t) * W6_LMAX);
    return s;
}

// ---- wave helpers (the W6 kernel runs one 64-lane wave per workgroup)
static __device__ __forceinline__ uint32_t w6_lane() { return threadIdx.x; }
static __device__ __forceinline__ void w6_sync() { __syncthreads(); }
static __device__ __forceinline__ unsigned long long w6_ballot(bool p) { return __ballot(p); }
static __device__ __forceinline__ uint32_t w6_rank(unsigned long long m) {
    return (uint32_t)__popcll(m & ((1ull << w6_lane()) - 1ull));
}
static __device__ __forceinline__ double w6_max_all(double x) {
    for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o));
    return x;
}
static __device__ __forceinline__ double w6_min_all(double x) {
    for (int o = 32; o > 0; o >>= 1) x = fmin(x, __shfl_xor(x, o));
    return x;
}
static __device__ __forceinline__ uint32_t w6_sum_u(uint32_t x) {
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}
static __device__ __forceinline__ bool w6_all(bool p) { return w6_ballot(!p) == 0ull; }

// ---- a contact as the bounce passes see it
struct W6C {
    uint32_t ai, bi;  // bi RP_NONE for terrain
    float inv_a, inv_b, m_n, bias;
    V3 ra, rb, n;
    M3 iia, iib;
};
static __device__ __forceinline__ W6C w6_load(const Dev &d, uint32_t c) {
    W6C k;
    k.ai = d.cs_ai[c]; k.bi = d.cs_bi[c];
    const float2 inv = d.cs_inv[c];
    k.inv_a = inv.x; k.inv_b = inv.y;
    const float4 m4 = d.cs_m[c];
    k.m_n = m4.x; k.bias = m4.w;
    k.ra = L3(d.cs_ra[c]); k.rb = L3(d.cs_rb[c]); k.n = L3(d.cs_n[c]);
    k.iia = k.inv_a > 0.0f ? load_m9(d.iiw + 9 * k.ai) : mzero();
    k.iib = k.inv_b > 0.0f ? load_m9(d.iiw + 9 * k.bi) : mzero();
    return k;
}
static __device__ __forceinline__ bool w6_jointed(const Dev &d, uint32_t b) { return b != RP_NONE && (d.flags[b] & FL_JOINTED); }
// free_only (solver.rs:1034-1041): jointed sides immovable, the normal effective mass again;
// false when neither side can move
static __device__ __forceinline__ bool w6_free_only(const Dev &d, W6C &k) {
    if (w6_jointed(d, k.ai)) { k.inv_a = 0.0f; k.iia = mzero(); }
    if (k.bi == RP_NONE || w6_jointed(d, k.bi)) { k.inv_b = 0.0f; k.iib = mzero(); }
    const V3 ca = cross3(k.ra, k.n), cb = cross3(k.rb, k.n);
    const float kk = ((k.inv_a + k.inv_b) + dot3(ca, mv(k.iia, ca))) + dot3(cb, mv(k.iib, cb));
    if (!(kk > 1.0e-9f)) return false;
    k.m_n = 1.0f / kk;
    return true;
}
// a row: the contact, and whether it is taken through free_only
static __device__ __forceinline__ W6C w6_row_c(const Dev &d, uint32_t c, bool free) {
    W6C k = w6_load(d, c);
    if (free) (void)w6_free_only(d, k);
    return k;
}
static __device__ __forceinline__ float w6_vn(const Dev &d, const W6C &k) { return dot3(rel_vel(d, k.ai, k.bi, k.ra, k.rb), k.n); }
static __device__ __forceinline__ void w6_apply_n(const Dev &d, const W6C &k, float dj) {
    apply_cimp(d, k.ai, k.bi, k.inv_a, k.inv_b, k.ra, k.rb, k.n * dj);
}
// a compliant pair's contact: no support row in either pass, no growth through it (solver.rs:1020,
// 1087, 1175)
static __device__ __forceinline__ bool w6_soft(const Dev &d, uint32_t c) { return (cm_flags(d, c) & CM_SOFT) != 0u; }
static __device__ __forceinline__ bool w6_loaded(const Dev &d, uint32_t c) {
    const float2 t = d.jt[c];
    return d.jn[c] > 0.0f || t.x != 0.0f || t.y != 0.0f;
}
// row_sides (solver.rs:123-139): n and r widened to f64, the world inverse inertia symmetrised
static __device__ void w6_side(const Dev &d, uint32_t body, double sign, V3 r, float inv, M3 ii, V3 nf, double *o, uint32_t *ob) {
    if (inv <= 0.0f && zero3(ii.c0) && zero3(ii.c1) && zero3(ii.c2)) {
        *ob = RP_NONE;
        for (int k = 0; k < 12; k++) o[k] = 0.0;
        return;
    }
    const D3 n = d3(nf.x, nf.y, nf.z);
    const D3 jl = d3mul(n, sign), ja = d3mul(d3cross(d3(r.x, r.y, r.z), n), sign);
    const double f[3][3] = {{ii.c0.x, ii.c0.y, ii.c0.z}, {ii.c1.x, ii.c1.y, ii.c1.z}, {ii.c2.x, ii.c2.y, ii.c2.z}};
    double m[3][3];
    for (int c = 0; c < 3; c++)
        for (int rr = 0; rr < 3; rr++) m[c][rr] = ((double)f[c][rr] + (double)f[rr][c]) * 0.5;
    const D3 ml = d3mul(jl, (double)inv);
    const D3 ma = d3add(d3add(d3mul(d3(m[0][0], m[0][1], m[0][2]), ja.x), d3mul(d3(m[1][0], m[1][1], m[1][2]), ja.y)),
                        d3mul(d3(m[2][0], m[2][1], m[2][2]), ja.z));
    *ob = body;
    o[0] = jl.x; o[1] = jl.y; o[2] = jl.z; o[3] = ja.x; o[4] = ja.y; o[5] = ja.z;
    o[6] = ml.x; o[7] = ml.y; o[8] = ml.z; o[9] = ma.x; o[10] = ma.y; o[11] = ma.z;
}
static __device__ __forceinline__ void w6_sides(const Dev &d, const W6C &k, double *o, uint32_t *ob) {
    w6_side(d, k.ai, -1.0, k.ra, k.inv_a, k.iia, k.n, o, ob);
    w6_side(d, k.bi == RP_NONE ? RP_NONE : k.bi, 1.0, k.rb, k.inv_b, k.iib, k.n, o + 12, ob + 1);
}
// Ji . (M^-1 Jj^T), linear dot first (solver.rs:156)
static __device__ __forceinline__ double w6_kij(const double *si, const double *sj) {
    return d3dot(d3(si[0], si[1], si[2]), d3(sj[6], sj[7], sj[8])) + d3dot(d3(si[3], si[4], si[5]), d3(sj[9], sj[10], sj[11]));
}

// contact_csr (solver.rs:142-168) for the n rows whose sides are in s.sd / s.sb: lane q builds
// row q.  Row j is touched by row q's side t when they share that side's body; its value is
// (0 + the side-0 product) + the side-1 product, as the reference's accumulator receives them.
static __device__ void w6_csr(const W6Ws &s, uint32_t n) {
    for (uint32_t q = w6_lane(); q < n; q += 64) {
        uint32_t cnt = 0;
        const uint32_t b0 = s.sb[2 * q], b1 = s.sb[2 * q + 1];
        for (uint32_t j = 0; j < n; j++) {
            const uint32_t x0 = s.sb[2 * j], x1 = s.sb[2 * j + 1];
            const bool m = j == q || (b0 != RP_NONE && (x0 == b0 || x1 == b0)) || (b1 != RP_NONE && (x0 == b1 || x1 == b1));
            cnt += m;
        }
        s.kptr[q + 1] = cnt;
    }
    w6_sync();
    if (w6_lane() == 0) {
        s.kptr[0] = 0;
        for (uint32_t q = 0; q < n; q++) s.kptr[q + 1] += s.kptr[q];
    }
    w6_sync();
    for (uint32_t q = w6_lane(); q < n; q += 64) {
        uint32_t e = s.kptr[q];
        const uint32_t b0 = s.sb[2 * q], b1 = s.sb[2 * q + 1];
        const double *si0 = s.sd + 24 * q, *si1 = si0 + 12;
        for (uint32_t j = 0; j < n; j++) {
            const uint32_t x0 = s.sb[2 * j], x1 = s.sb[2 * j + 1];
            const int t0 = b0 == RP_NONE ? -1 : (x0 == b0 ? 0 : (x1 == b0 ? 1 : -1));
            const int t1 = b1 == RP_NONE ? -1 : (x0 == b1 ? 0 : (x1 == b1 ? 1 : -1));
            if (j != q && t0 < 0 && t1 < 0) continue;
            double a = 0.0;
            if (t0 >= 0) a += w6_kij(si0, s.sd + 24 * j + 12 * t0);
            if (t1 >= 0) a += w6_kij(si1, s.sd + 24 * j + 12 * t1);
            s.kidx[e] = j;
            s.k0[e] = a;
            e++;
        }
    }
    w6_sync();
}
static __device__ __forceinline__ double w6_row_dot(const W6Ws &s, const double *kv, uint32_t i, const double *x) {
    double a = 0.0;
    for (uint32_t t = s.kptr[i]; t < s.kptr[i + 1]; t++) a += kv[t] * x[s.kidx[t]];
    return a;
}
static __device__ __forceinline__ double w6_diag(const W6Ws &s, const double *kv, uint32_t i) {
    for (uint32_t t = s.kptr[i]; t < s.kptr[i + 1]; t++) if (s.kidx[t] == i) return kv[t];
    return 0.0;
}

// Symbolic::min_degree (bounce_lcp.rs:83-147): eliminate the node of least degree (lowest index
// on ties); its neighbours become a clique.  Then the factor's structure: column p holds the
// eliminated node's neighbours at their positions, ascending; re lists the entries at each
// position in column order; ce[e] ends entry e's column.
static __device__ void w6_min_degree(const W6Ws &s, uint32_t n, uint8_t *gone) {
    for (uint32_t i = w6_lane(); i < n; i += 64) {
        for (int k = 0; k < 4; k++) s.adj[4 * i + k] = 0ull;
        gone[i] = 0;
        for (uint32_t t = s.kptr[i]; t < s.kptr[i + 1]; t++) {
            const uint32_t j = s.kidx[t];
            if (j != i) s.adj[4 * i + (j >> 6)] |= 1ull << (j & 63);
        }
    }
    w6_sync();
    for (uint32_t step = 0; step < n; step++) {
        uint32_t best = 0xFFFFFFFFu;  // (degree << 9) | node
        for (uint32_t u = w6_lane(); u < n; u += 64) {
            if (gone[u]) continue;
            const uint32_t deg = (uint32_t)(__popcll(s.adj[4 * u]) + __popcll(s.adj[4 * u + 1]) + __popcll(s.adj[4 * u + 2]) + __popcll(s.adj[4 * u + 3]));
            best = min(best, (deg << 9) | u);
        }
        for (int o = 32; o > 0; o >>= 1) best = min(best, (uint32_t)__shfl_xor((int)best, o));
        const uint32_t v = best & 511u;
        if (w6_lane() < 4) { s.nbs[4 * v + w6_lane()] = s.adj[4 * v + w6_lane()]; }
        w6_sync();
        if (w6_lane() == 0) { gone[v] = 1; s.order[step] = v; }
        for (uint32_t u = w6_lane(); u < n; u += 64) {
            if (!((s.nbs[4 * v + (u >> 6)] >> (u & 63)) & 1ull)) continue;
            for (int k = 0; k < 4; k++) s.adj[4 * u + k] |= s.nbs[4 * v + k];
            s.adj[4 * u + (v >> 6)] &= ~(1ull << (v & 63));
            s.adj[4 * u + (u >> 6)] &= ~(1ull << (u & 63));
        }
        if (w6_lane() < 4) s.adj[4 * v + w6_lane()] = 0ull;
        w6_sync();
    }
    for (uint32_t p = w6_lane(); p < n; p += 64) s.pos[s.order[p]] = p;
    for (uint32_t p = w6_lane(); p < n; p += 64) {
        const uint32_t r = s.order[p];
        s.cp[p + 1] = (uint32_t)(__popcll(s.nbs[4 * r]) + __popcll(s.nbs[4 * r + 1]) + __popcll(s.nbs[4 * r + 2]) + __popcll(s.nbs[4 * r + 3]));
    }
    w6_sync();
    if (w6_lane() == 0) {
        s.cp[0] = 0;
        for (uint32_t p = 0; p < n; p++) s.cp[p + 1] += s.cp[p];
    }
    w6_sync();
    for (uint32_t p = w6_lane(); p < n; p += 64) {
        const uint32_t r = s.order[p], e0 = s.cp[p];
        uint32_t e = e0;
        for (uint32_t u = 0; u < n; u++)
            if ((s.nbs[4 * r + (u >> 6)] >> (u & 63)) & 1ull) {
                const uint32_t x = s.pos[u];
                uint32_t k = e++;
                while (k > e0 && s.ri[k - 1] > x) { s.ri[k] = s.ri[k - 1]; k--; }
                s.ri[k] = x;
            }
        for (uint32_t t = e0; t < e; t++) s.ce[t] = s.cp[p + 1];
    }
    w6_sync();
    if (w6_lane() == 0) {
        // entries at each position, in column order (columns ascending, then entries)
        const uint32_t tot = s.cp[n];
        for (uint32_t p = 0; p <= n; p++) s.rp[p] = 0;
        for (uint32_t t = 0; t < tot; t++) s.rp[s.ri[t] + 1]++;
        for (uint32_t p = 0; p < n; p++) s.rp[p + 1] += s.rp[p];
        for (uint32_t p = 0; p < n; p++) s.fill[p] = s.rp[p];
        for (uint32_t t = 0; t < tot; t++) s.re[s.fill[s.ri[t]]++] = t;
    }
    w6_sync();
}

// factor (bounce_lcp.rs:182-202), left-looking: column j gathers row order[j] of K at positions
// >= j, adds d, takes each earlier column's update in column order (the wave shares one update's
// entries), then scales.
static __device__ void w6_factor(const W6Ws &s, uint32_t n, const double *kv, const double *dd, double *work) {
    for (uint32_t i = w6_lane(); i < n; i += 64) work[i] = 0.0;
    w6_sync();
    for (uint32_t j = 0; j < n; j++) {
        const uint32_t r = s.order[j];
        for (uint32_t t = s.kptr[r] + w6_lane(); t < s.kptr[r + 1]; t += 64) {
            const uint32_t p = s.pos[s.kidx[t]];
            if (p >= j) work[p] += kv[t];
        }
        w6_sync();
        if (w6_lane() == 0) work[j] += dd[r];
        w6_sync();
        for (uint32_t q = s.rp[j]; q < s.rp[j + 1]; q++) {
            const uint32_t e = s.re[q];
            const double ljk = s.lval[e];
            const uint32_t end = s.ce[e];
            for (uint32_t t = e + w6_lane(); t < end; t += 64) work[s.ri[t]] -= s.lval[t] * ljk;
            w6_sync();
        }
        const double ljj = sqrt(fmax(work[j], 1.0e-300));
        for (uint32_t t = s.cp[j] + w6_lane(); t < s.cp[j + 1]; t += 64) {
            const uint32_t p = s.ri[t];
            s.lval[t] = work[p] / ljj;
            work[p] = 0.0;
        }
        w6_sync();
        if (w6_lane() == 0) { s.ldiag[j] = ljj; work[j] = 0.0; }
        w6_sync();
    }
}

// solve (bounce_lcp.rs:214-228): forward substitution by columns (each position takes its
// updates in column order), backward substitution by rows on lane 0 (a sum in entry order)
static __device__ void w6_solve(const W6Ws &s, uint32_t n, const double *b, double *x, double *y) {
    for (uint32_t p = w6_lane(); p < n; p += 64) y[p] = b[s.order[p]];
    w6_sync();
    for (uint32_t j = 0; j < n; j++) {
        if (w6_lane() == 0) y[j] /= s.ldiag[j];
        w6_sync();
        const double yj = y[j];
        for (uint32_t t = s.cp[j] + w6_lane(); t < s.cp[j + 1]; t += 64) y[s.ri[t]] -= s.lval[t] * yj;
        w6_sync();
    }
    if (w6_lane() == 0) {
        for (uint32_t j = n; j-- > 0;) {
            double v = y[j];
            for (uint32_t t = s.cp[j]; t < s.cp[j + 1]; t++) v -= s.lval[t] * y[s.ri[t]];
            y[j] = v / s.ldiag[j];
        }
    }
    w6_sync();
    for (uint32_t p = w6_lane(); p < n; p += 64) x[s.order[p]] = y[p];
    w6_sync();
}

// the largest a <= 1 keeping y + a dy and w + a dw nonnegative (bounce_lcp.rs:271-278)
static __device__ double w6_step(uint32_t n, const double *y, const double *dy, const double *w, const double *dw) {
    double a = 1.0;
    for (uint32_t i = w6_lane(); i < n; i += 64) {
        if (dy[i] < 0.0) a = fmin(a, -y[i] / dy[i]);
        if (dw[i] < 0.0) a = fmin(a, -w[i] / dw[i]);
    }
    return w6_min_all(a);
}

// bounce_qp (bounce_lcp.rs:256-321): returns whether the stopping test held; x = y - acc0
static __device__ bool w6_bounce_qp(const W6Ws &s, uint32_t n, const double *b, const double *acc0, double *x, uint8_t *gone,
                                    double *work, double *lds_red) {
    double *kd = s.v + 2 * W6_NMAX, *q = s.v + 3 * W6_NMAX, *y = s.v + 4 * W6_NMAX, *w = s.v + 5 * W6_NMAX,
           *r = s.v + 6 * W6_NMAX, *xt = s.v + 7 * W6_NMAX, *dd = s.v + 8 * W6_NMAX, *rhs = s.v + 9 * W6_NMAX,
           *tg = s.v + 10 * W6_NMAX, *dya = s.v + 11 * W6_NMAX, *dwa = s.v + 12 * W6_NMAX, *dy = s.v + 13 * W6_NMAX,
           *dw = s.v + 14 * W6_NMAX, *ys = s.v + 15 * W6_NMAX;
    const uint32_t nnz = s.kptr[n];
    for (uint32_t t = w6_lane(); t < nnz; t += 64) s.k1[t] = s.k0[t];
    w6_sync();
    for (uint32_t i = w6_lane(); i < n; i += 64)
        for (uint32_t t = s.kptr[i]; t < s.kptr[i + 1]; t++) if (s.kidx[t] == i) s.k1[t] *= 1.0 + W6_CFM;
    w6_sync();
    w6_min_degree(s, n, gone);
    double e0 = 0.0;
    for (uint32_t i = w6_lane(); i < n; i += 64) {
        kd[i] = w6_diag(s, s.k0, i);
        q[i] = -w6_row_dot(s, s.k1, i, acc0) - b[i];
        e0 = fmax(e0, fabs(q[i]));
    }
    const double eta = fmax(1.0, w6_max_all(e0));
    for (uint32_t i = w6_lane(); i < n; i += 64) { y[i] = kd[i] > 0.0 ? eta / kd[i] : 1.0; w[i] = eta; }
    w6_sync();
    bool conv = false;
    for (uint32_t it = 1; it <= W6_SOLVES; it++) {
        double rm = 0.0;
        for (uint32_t i = w6_lane(); i < n; i += 64) {
            r[i] = (w6_row_dot(s, s.k1, i, y) + q[i]) - w[i];
            rm = fmax(rm, fabs(r[i]));
            xt[i] = y[i] - acc0[i];
        }
        const double rmax = w6_max_all(rm);
        w6_sync();
        bool ok = true;
        for (uint32_t i = w6_lane(); i < n; i += 64) {
            const double wt = w6_row_dot(s, s.k0, i, xt) - b[i];
            const double tol = 1.0e-7 * (1.0 + fabs(b[i]));
            if (!(wt >= -tol && fmin(fabs(wt), y[i] * kd[i]) < tol)) ok = false;
        }
        if (rmax < 1.0e-7 && w6_all(ok)) { conv = true; break; }
        if (w6_lane() == 0) {
            double mu = 0.0;
            for (uint32_t i = 0; i < n; i++) mu += y[i] * w[i];
            lds_red[0] = mu / (double)n;
        }
        for (uint32_t i = w6_lane(); i < n; i += 64) dd[i] = w[i] / y[i];
        w6_sync();
        const double mu = lds_red[0];
        w6_factor(s, n, s.k1, dd, work);
        for (uint32_t i = w6_lane(); i < n; i += 64) { tg[i] = -y[i] * w[i]; rhs[i] = -r[i] + tg[i] / y[i]; }
        w6_sync();
        w6_solve(s, n, rhs, dya, ys);
        for (uint32_t i = w6_lane(); i < n; i += 64) dwa[i] = (tg[i] - w[i] * dya[i]) / y[i];
        w6_sync();
        const double aa = w6_step(n, y, dya, w, dwa);
        if (w6_lane() == 0) {
            double ma = 0.0;
            for (uint32_t i = 0; i < n; i++) ma += (y[i] + aa * dya[i]) * (w[i] + aa * dwa[i]);
            lds_red[1] = ma / (double)n;
        }
        w6_sync();
        const double ratio = lds_red[1] / fmax(mu, 1.0e-300);
        const double sigma = fmin(ratio * ratio * ratio, 1.0);
        for (uint32_t i = w6_lane(); i < n; i += 64) { tg[i] = sigma * mu - y[i] * w[i] - dya[i] * dwa[i]; rhs[i] = -r[i] + tg[i] / y[i]; }
        w6_sync();
        w6_solve(s, n, rhs, dy, ys);
        for (uint32_t i = w6_lane(); i < n; i += 64) dw[i] = (tg[i] - w[i] * dy[i]) / y[i];
        w6_sync();
        const double a = fmin(0.995 * w6_step(n, y, dy, w, dw), 1.0);
        bool fin = isfinite(a);
        for (uint32_t i = w6_lane(); i < n; i += 64) fin = fin && isfinite(dy[i]) && isfinite(dw[i]);
        if (!w6_all(fin)) break;
        for (uint32_t i = w6_lane(); i < n; i += 64) {
            y[i] = fmax(y[i] + a * dy[i], 1.0e-300);
            w[i] = fmax(w[i] + a * dw[i], 1.0e-300);
        }
        w6_sync();
    }
    for (uint32_t i = w6_lane(); i < n; i += 64) x[i] = y[i] - acc0[i];
    w6_sync();
    return conv;
}

// pseudo_solve (bounce_lcp.rs:330-337): (K + diag(CFM K_ii)) x = u
static __device__ void w6_pseudo_solve(const W6Ws &s, uint32_t n, const double *u, double *x, uint8_t *gone, double *work) {
    double *dd = s.v + 8 * W6_NMAX, *ys = s.v + 15 * W6_NMAX;
    w6_min_degree(s, n, gone);
    for (uint32_t i = w6_lane(); i < n; i += 64) dd[i] = W6_CFM * w6_diag(s, s.k0, i);
    w6_sync();
    w6_factor(s, n, s.k0, dd, work);
    w6_solve(s, n, u, x, ys);
}

// toi_bounce (solver.rs:197-212), f32
static __device__ __forceinline__ float w6_toi_bounce(float s0, float v0, float a, float e, float dt, float &pseudo, bool &has) {
    float t_c;
    if (s0 <= 0.0f) t_c = 0.0f;
    else {
        const float disc = v0 * v0 - (2.0f * a) * s0;
        const float den = disc >= 0.0f ? -v0 + sqrtf(disc) : 0.0f;
        t_c = den > 0.0f ? fminf((2.0f * s0) / den, dt) : dt;
    }
    const float v_out = fmaxf((-e) * (v0 + a * t_c), 0.0f);
    const float tau = dt - t_c;
    const float rise = v_out * tau + ((0.5f * a) * tau) * tau;
    float v_end, s_end;
    if (rise < 0.0f) { v_end = 0.0f; s_end = 0.0f; }
    else { v_end = fmaxf(v_out + a * tau, 0.0f); s_end = rise; }
    has = s0 > 0.0f;
    pseudo = (s_end - s0) / dt - v_end;
    return v_end;
}

// Ordered compaction over [lo, hi): the indices passing pred, in order, into out[base..];
// returns the count (every lane).
template <class P>
static __device__ uint32_t w6_compact(uint32_t lo, uint32_t hi, uint32_t *out, P pred) {
    uint32_t base = 0;
    for (uint32_t i0 = lo; i0 < hi; i0 += 64) {
        const uint32_t i = i0 + w6_lane();
        const bool p = i < hi && pred(i);
        const unsigned long long m = w6_ballot(p);
        if (p) out[base + w6_rank(m)] = i;
        base += (uint32_t)__popcll(m);
    }
    w6_sync();
    return base;
}

// grow (solver.rs:1018-1030): the involved bodies (W6B_INV) spread through rigid loaded contacts
// to the closure, never onto static, sleeping or jointed bodies
static __device__ void w6_grow(const Dev &d, uint32_t c0, uint32_t c1, uint32_t *flag) {
    for (;;) {
        if (w6_lane() == 0) *flag = 0;
        w6_sync();
        bool ch = false;
        for (uint32_t c = c0 + w6_lane(); c < c1; c += 64) {
            const uint32_t ai = d.cs_ai[c], bi = d.cs_bi[c];
            if (bi == RP_NONE || w6_soft(d, c) || !w6_loaded(d, c)) continue;
            const float2 inv = d.cs_inv[c];
            const bool ia = (d.w6_b[2 * ai + 1] & W6B_INV) != 0, ib = (d.w6_b[2 * bi + 1] & W6B_INV) != 0;
            if (ia && !ib && inv.y > 0.0f && !w6_jointed(d, bi)) { atomicOr(&d.w6_b[2 * bi + 1], W6B_INV); ch = true; }
            if (ib && !ia && inv.x > 0.0f && !w6_jointed(d, ai)) { atomicOr(&d.w6_b[2 * ai + 1], W6B_INV); ch = true; }
        }
        if (ch) *flag = 1;
        w6_sync();
        const bool again = *flag != 0;
        w6_sync();
        if (!again) break;
    }
}
static __device__ __forceinline__ bool w6_inv(const Dev &d, uint32_t b) { return b != RP_NONE && (d.w6_b[2 * b + 1] & W6B_INV); }

// row_components (solver.rs:174-192) for rows[0..nr) (contact indices; the sides of row i move
// when its inverse masses, after free_only for rows at or past nfree0, are > 0): comp[i], the
// components in first-row order, rows ascending; ordered[] receives the rows grouped by
// component, cstart[] (nr + 1 entries at most) each component's start.  Lane 0; returns the count.
static __device__ uint32_t w6_components(const Dev &d, const uint32_t *rows, uint32_t nr, uint32_t nfree0, bool all_free, uint32_t *comp,
                                         uint32_t *ordered, uint32_t *cstart) {
    auto find = [&](uint32_t x) {
        while (d.w6_par[x] != x) { const uint32_t g = d.w6_par[d.w6_par[x]]; d.w6_par[x] = g; x = g; }
        return x;
    };
    auto sides = [&](uint32_t i, uint32_t &a, uint32_t &b) {
        const uint32_t c = rows[i];
        W6C k = w6_load(d, c);
        if (all_free || i >= nfree0) (void)w6_free_only(d, k);
        a = k.inv_a > 0.0f ? k.ai : RP_NONE;
        b = (k.bi != RP_NONE && k.inv_b > 0.0f) ? k.bi : RP_NONE;
    };
    for (uint32_t i = 0; i < nr; i++) {
        uint32_t a, b;
        sides(i, a, b);
        if (a != RP_NONE) { d.w6_par[a] = a; d.w6_b[2 * a] = W6_EMPTY; }
        if (b != RP_NONE) { d.w6_par[b] = b; d.w6_b[2 * b] = W6_EMPTY; }
    }
    for (uint32_t i = 0; i < nr; i++) {
        uint32_t a, b;
        sides(i, a, b);
        if (a == RP_NONE || b == RP_NONE) continue;
        const uint32_t x = find(a), y = find(b);
        if (x != y) d.w6_par[x > y ? x : y] = x < y ? x : y;
    }
    uint32_t nc = 0;
    for (uint32_t i = 0; i < nr; i++) {
        uint32_t a, b;
        sides(i, a, b);
        const uint32_t body = a != RP_NONE ? a : b;
        if (body == RP_NONE) { comp[i] = nc++; continue; }
        const uint32_t r = find(body);
        if (d.w6_b[2 * r] == W6_EMPTY) d.w6_b[2 * r] = nc++;
        comp[i] = d.w6_b[2 * r];
    }
    for (uint32_t k = 0; k <= nc; k++) cstart[k] = 0;
    for (uint32_t i = 0; i < nr; i++) cstart[comp[i] + 1]++;
    for (uint32_t k = 0; k < nc; k++) cstart[k + 1] += cstart[k];
    for (uint32_t i = 0; i < nr; i++) ordered[cstart[comp[i]]++] = i;
    for (uint32_t k = nc; k > 0; k--) cstart[k] = cstart[k - 1];
    cstart[0] = 0;
    return nc;
}

// The prep's restitution bias test (solver.rs:569-570): the mid-step approach beats -1 m/s and
// -e * approach is not <= 0.  The prep stores w6_toi / w6_tv exactly for these contacts.
static __device__ __forceinline__ bool w6_has_rb(const Dev &d, uint32_t c, uint32_t w) {
    const float approach = d.cs_mu[c].w;
    if (!(approach < -1.0f)) return false;
    const uint32_t ca = d.ct_a[c], cb = d.ct_b[c];
    const uint32_t ma = d.col_mat[ca], mb = cb == RP_TERRAIN ? d.hf[w].mat : d.col_mat[cb];
    const float rest = d.pair_tab[(ma & 15u) * 16u + (mb & 15u)].z;
    return !(-rest * approach <= 0.0f);
}
// rests_on (solver.rs:1051-1057): record that body x rests on partner p; the word keeps one
// partner, the bit whether a different one exists (all the bounce asks is "a partner other than q")
static __device__ __forceinline__ void w6_rest_on(const Dev &d, uint32_t x, uint32_t p) {
    const uint32_t old = atomicCAS(&d.w6_b[2 * x], W6_EMPTY, p);
    if (old != W6_EMPTY && old != p && !(d.w6_b[2 * x + 1] & W6B_MULTI)) atomicOr(&d.w6_b[2 * x + 1], W6B_MULTI);
}
static __device__ __forceinline__ bool w6_rests_elsewhere(const Dev &d, uint32_t x, uint32_t other) {
    const uint32_t p0 = d.w6_b[2 * x];
    return p0 != W6_EMPTY && ((d.w6_b[2 * x + 1] & W6B_MULTI) || p0 != other);
}
static __device__ __forceinline__ bool w6_free_ok(const Dev &d, uint32_t c) {
    W6C k = w6_load(d, c);
    return w6_free_only(d, k);
}

// Gauss-Seidel over one component's rows from where they are, alternating direction, until no row
// changes by BOUNCE_TOL (solver.rs:1135-1154); lane 0.
static __device__ void w6_sweep_vel(const Dev &d, const uint32_t *grp, uint32_t n, const uint32_t *rc, const float *rt, float *ra,
                                    uint32_t ntoi) {
    for (uint32_t it = 0; it < W6_SWEEPS; it++) {
        float change = 0.0f;
        for (uint32_t m = 0; m < n; m++) {
            const uint32_t i = grp[it % 2 == 0 ? m : n - 1 - m];
            const W6C k = w6_row_c(d, rc[i], i >= ntoi);
            const float v = w6_vn(d, k);
            float *slot = i < ntoi ? &d.jn[rc[i]] : &ra[i];
            const float old = *slot;
            const float nv = fmaxf(old + (rt[i] - v) * k.m_n, 0.0f);
            *slot = nv;
            w6_apply_n(d, k, nv - old);
            change = fmaxf(change, fabsf(nv - old) / k.m_n / (1.0f + fabsf(rt[i])));
        }
        if (change < W6_TOL) break;
    }
}
// the pseudo pass's sweeps (solver.rs:1213-1234) on the f32 pseudo-velocities; lane 0
static __device__ void w6_sweep_pseudo(const Dev &d, const uint32_t *grp, uint32_t n, const uint32_t *rc, const float *rt) {
    for (uint32_t it = 0; it < W6_SWEEPS; it++) {
        float change = 0.0f;
        for (uint32_t m = 0; m < n; m++) {
            const uint32_t i = grp[it % 2 == 0 ? m : n - 1 - m];
            const W6C k = w6_row_c(d, rc[i], true);
            const V3 la = L3(d.w6_pl[k.ai]), wa = L3(d.w6_pa[k.ai]);
            const V3 lb = k.bi == RP_NONE ? splat3(0.0f) : L3(d.w6_pl[k.bi]), wb = k.bi == RP_NONE ? splat3(0.0f) : L3(d.w6_pa[k.bi]);
            const float rel = dot3((lb + cross3(wb, k.rb)) - (la + cross3(wa, k.ra)), k.n);
            const float dd = rt[i] - rel;
            change = fmaxf(change, fabsf(dd) / (1.0f + fabsf(rt[i])));
            const V3 j = k.n * (dd * k.m_n);
            if (k.inv_a > 0.0f) {
                d.w6_pl[k.ai] = S3(L3(d.w6_pl[k.ai]) - j * k.inv_a);
                d.w6_pa[k.ai] = S3(L3(d.w6_pa[k.ai]) - mv(k.iia, cross3(k.ra, j)));
            }
            if (k.inv_b > 0.0f) {
                d.w6_pl[k.bi] = S3(L3(d.w6_pl[k.bi]) + j * k.inv_b);
                d.w6_pa[k.bi] = S3(L3(d.w6_pa[k.bi]) + mv(k.iib, cross3(k.rb, j)));
            }
        }
        if (change < W6_TOL) break;
    }
}

// One flagged world: relax pass, velocity pass, pseudo pass.
static __device__ void w6_world(const Dev &d, const TickParams &tp, const W6Ws &s, uint32_t w, double *work, uint8_t *gone, double *red,
                                uint32_t *shu) {
    const uint32_t lane = w6_lane();
    uint32_t c0, c1;
    world_contacts(d, w, c0, c1);
    const uint32_t b0 = d.w_body_off[w], b1 = d.w_body_off[w + 1];
    const float dt = tp.dt;
    uint32_t *R = d.w6_rl + c0, *rc = d.w6_rc + c0, *pc = d.w6_pc + c0, *grp = d.w6_grp + c0, *comp = d.w6_comp + c0;
    uint32_t *cst = d.w6_cst + c0 + w;
    float *rt = d.w6_rt + c0, *ra = d.w6_racc + c0, *pu = d.w6_pu + c0;
    double *vb = s.v, *vacc0 = s.v + W6_NMAX, *vx = s.v + 16 * W6_NMAX;

    for (uint32_t c = c0 + lane; c < c1; c += 64) d.w6_mark[c] = w6_has_rb(d, c, w) ? W6M_R : 0u;
    w6_sync();
    const uint32_t nR = w6_compact(c0, c1, R, [&](uint32_t c) { return (d.w6_mark[c] & W6M_R) != 0u; });
    // restitution relax (solver.rs:1001-1008): the contacts that are not time-of-impact rows
    if (lane == 0)
        for (uint32_t k = 0; k < nR; k++) {
            const uint32_t c = R[k];
            const float4 toi = d.w6_toi[c];
            if (toi.w != 0.0f || toi.x <= 0.0f || d.jn[c] <= 0.0f) continue;
            atomicAdd(&d.cnt[CT_W6_RELAX], 1u);
            const W6C x = w6_load(d, c);
            const float v = w6_vn(d, x);
            const float old = d.jn[c];
            const float nv = fmaxf(old + (toi.x - v) * x.m_n, 0.0f);
            d.jn[c] = nv;
            w6_apply_n(d, x, nv - old);
        }
    w6_sync();
    // the bounce rows (solver.rs:1048): time-of-impact contacts with a bias that pushed
    const uint32_t ntoi = w6_compact(0, nR, pc, [&](uint32_t k) {
        const uint32_t c = R[k];
        const float4 toi = d.w6_toi[c];
        return toi.w != 0.0f && toi.x > 0.0f && d.jn[c] > 0.0f;
    });
    if (ntoi == 0) return;
    if (lane == 0) atomicAdd(&d.cnt[CT_W6_TOI], ntoi);
    for (uint32_t t = lane; t < ntoi; t += 64) {
        const uint32_t c = R[pc[t]];
        rc[t] = c;
        d.w6_mark[c] |= W6M_T;
    }
    for (uint32_t b = b0 + lane; b < b1; b += 64) { d.w6_b[2 * b] = W6_EMPTY; d.w6_b[2 * b + 1] = 0u; }
    w6_sync();
    // who the bodies rest on outside the bounce (solver.rs:1051-1057)
    for (uint32_t c = c0 + lane; c < c1; c += 64) {
        if ((d.w6_mark[c] & W6M_T) || !w6_loaded(d, c)) continue;
        const uint32_t ai = d.cs_ai[c], bi = d.cs_bi[c];
        w6_rest_on(d, ai, bi == RP_NONE ? RP_TERRAIN : bi);
        if (bi != RP_NONE) w6_rest_on(d, bi, ai);
    }
    w6_sync();
    // each bounce row's end velocity and, when it started apart, its pseudo velocity
    // (solver.rs:1058-1076); u parks in the row's accumulator slot
    for (uint32_t t = lane; t < ntoi; t += 64) {
        const uint32_t c = rc[t];
        const uint32_t ai = d.cs_ai[c], bi = d.cs_bi[c];
        const float4 toi = d.w6_toi[c], tv = d.w6_tv[c];
        const bool sa = w6_rests_elsewhere(d, ai, bi == RP_NONE ? RP_TERRAIN : bi);
        const float a0 = tv.x, a1 = sa ? tv.x : tv.y;
        float bb0 = 0.0f, bb1 = 0.0f;
        if (bi != RP_NONE) {
            const bool sb = w6_rests_elsewhere(d, bi, ai);
            bb0 = tv.z; bb1 = sb ? tv.z : tv.w;
        }
        const float v0 = bb0 - a0;
        const float acc = ((bb1 - a1) - v0) / dt;
        float u;
        bool has;
        rt[t] = w6_toi_bounce(toi.y, v0, acc, toi.z, dt, u, has);
        ra[t] = u;
    }
    w6_sync();
    // the pseudo rows, in bounce-row order (pc: contact, pu: target)
    const uint32_t nP = w6_compact(0, ntoi, grp, [&](uint32_t t) { return d.w6_toi[rc[t]].y > 0.0f; });
    for (uint32_t p = lane; p < nP; p += 64) { pc[p] = rc[grp[p]]; pu[p] = ra[grp[p]]; }
    // involved: the bounce rows' bodies, grown through the loaded contacts
    for (uint32_t t = lane; t < ntoi; t += 64) {
        const uint32_t c = rc[t];
        atomicOr(&d.w6_b[2 * d.cs_ai[c] + 1], W6B_INV);
        if (d.cs_bi[c] != RP_NONE) atomicOr(&d.w6_b[2 * d.cs_bi[c] + 1], W6B_INV);
    }
    w6_sync();
    w6_grow(d, c0, c1, shu);
    // support rows (solver.rs:1081-1092): fresh push-only rows with target min(bias, 0)
    const uint32_t nsup = w6_compact(c0, c1, rc + ntoi, [&](uint32_t c) {
        if ((d.w6_mark[c] & W6M_T) || w6_soft(d, c) || d.jn[c] <= 0.0f) return false;
        if (!w6_inv(d, d.cs_ai[c]) && !w6_inv(d, d.cs_bi[c])) return false;
        return w6_free_ok(d, c);
    });
    const uint32_t nr = ntoi + nsup;
    if (lane == 0) atomicAdd(&d.cnt[CT_W6_SUPPORT], nsup);
    for (uint32_t r = ntoi + lane; r < nr; r += 64) { rt[r] = fminf(d.cs_m[rc[r]].w, 0.0f); ra[r] = 0.0f; }
    w6_sync();
    if (lane == 0) shu[1] = w6_components(d, rc, nr, ntoi, false, comp, grp, cst);
    w6_sync();
    const uint32_t ncomp = shu[1];
    for (uint32_t cc = 0; cc < ncomp; cc++) {
        const uint32_t *g = grp + cst[cc];
        const uint32_t n = cst[cc + 1] - cst[cc];
        bool sweep = n > W6_DIRECT_MAX;
        if (!sweep) {
            for (uint32_t q = lane; q < n; q += 64) {
                const uint32_t i = g[q];
                const W6C k = w6_row_c(d, rc[i], i >= ntoi);
                w6_sides(d, k, s.sd + 24 * q, s.sb + 2 * q);
                vb[q] = (double)(rt[i] - w6_vn(d, k));
                vacc0[q] = (double)(i < ntoi ? d.jn[rc[i]] : ra[i]);
            }
            w6_sync();
            w6_csr(s, n);
            sweep = !w6_bounce_qp(s, n, vb, vacc0, vx, gone, work, red);
            if (lane == 0)
                for (uint32_t q = 0; q < n; q++) {
                    const uint32_t i = g[q];
                    const W6C k = w6_row_c(d, rc[i], i >= ntoi);
                    float *slot = i < ntoi ? &d.jn[rc[i]] : &ra[i];
                    const float old = *slot;
                    const float nv = fmaxf((float)((double)old + vx[q]), 0.0f);
                    *slot = nv;
                    w6_apply_n(d, k, nv - old);
                }
            w6_sync();
        }
        if (sweep) {
            if (lane == 0) { atomicAdd(&d.cnt[CT_W6_FALLBACK], 1u); w6_sweep_vel(d, g, n, rc, rt, ra, ntoi); }
            w6_sync();
        }
    }
    if (nP == 0) return;
    if (lane == 0) atomicAdd(&d.cnt[CT_W6_PSEUDO], nP);
    // ---- the pseudo pass (solver.rs:1163-1243)
    for (uint32_t p = lane; p < nP; p += 64) d.w6_mark[pc[p]] |= W6M_P;
    for (uint32_t b = b0 + lane; b < b1; b += 64) {
        d.w6_b[2 * b + 1] = 0u;
        d.w6_pl[b] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        d.w6_pa[b] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        for (int k = 0; k < 6; k++) d.w6_acc[6 * (size_t)b + k] = 0.0;
    }
    w6_sync();
    for (uint32_t p = lane; p < nP; p += 64) {
        const uint32_t c = pc[p];
        atomicOr(&d.w6_b[2 * d.cs_ai[c] + 1], W6B_INV);
        if (d.cs_bi[c] != RP_NONE) atomicOr(&d.w6_b[2 * d.cs_bi[c] + 1], W6B_INV);
    }
    w6_sync();
    w6_grow(d, c0, c1, shu);
    const uint32_t np2 = w6_compact(0, nP, grp, [&](uint32_t p) { return w6_free_ok(d, pc[p]); });
    for (uint32_t r = lane; r < np2; r += 64) { rc[r] = pc[grp[r]]; rt[r] = pu[grp[r]]; }
    w6_sync();
    const uint32_t nex = w6_compact(c0, c1, rc + np2, [&](uint32_t c) {
        if ((d.w6_mark[c] & W6M_P) || w6_soft(d, c) || !w6_loaded(d, c)) return false;
        if (!w6_inv(d, d.cs_ai[c]) && !w6_inv(d, d.cs_bi[c])) return false;
        return w6_free_ok(d, c);
    });
    const uint32_t npr = np2 + nex;
    for (uint32_t r = np2 + lane; r < npr; r += 64) rt[r] = 0.0f;
    w6_sync();
    if (lane == 0) shu[1] = w6_components(d, rc, npr, 0, true, comp, grp, cst);
    w6_sync();
    const uint32_t ncp = shu[1];
    for (uint32_t cc = 0; cc < ncp; cc++) {
        const uint32_t *g = grp + cst[cc];
        const uint32_t n = cst[cc + 1] - cst[cc];
        if (n <= W6_DIRECT_MAX) {
            for (uint32_t q = lane; q < n; q += 64) {
                const W6C k = w6_row_c(d, rc[g[q]], true);
                w6_sides(d, k, s.sd + 24 * q, s.sb + 2 * q);
                vb[q] = (double)rt[g[q]];
            }
            w6_sync();
            w6_csr(s, n);
            w6_pseudo_solve(s, n, vb, vx, gone, work);
            if (lane == 0)
                for (uint32_t q = 0; q < n; q++)
                    for (uint32_t t = 0; t < 2; t++) {
                        const uint32_t body = s.sb[2 * q + t];
                        if (body == RP_NONE) continue;
                        const double *sd = s.sd + 24 * q + 12 * t;
                        double *a = d.w6_acc + 6 * (size_t)body;
                        for (int k = 0; k < 3; k++) { a[k] = a[k] + sd[6 + k] * vx[q]; }
                        for (int k = 0; k < 3; k++) { a[3 + k] = a[3 + k] + sd[9 + k] * vx[q]; }
                        d.w6_b[2 * body + 1] |= W6B_DIRECT;
                    }
            w6_sync();
            continue;
        }
        if (lane == 0) { atomicAdd(&d.cnt[CT_W6_FALLBACK], 1u); w6_sweep_pseudo(d, g, n, rc, rt); }
        w6_sync();
    }
    for (uint32_t b = b0 + lane; b < b1; b += 64) {
        if (!(d.w6_b[2 * b + 1] & W6B_DIRECT)) continue;
        const double *a = d.w6_acc + 6 * (size_t)b;
        d.w6_pl[b] = make_float4((float)a[0], (float)a[1], (float)a[2], 0.0f);
        d.w6_pa[b] = make_float4((float)a[3], (float)a[4], (float)a[5], 0.0f);
    }
    if (lane == 0) d.w_w6[d.nw + w] = 1u;
    w6_sync();
}

// The bounce passes of every flagged world: workgroup k takes worlds k, k + grid, ... with
// workspace slot k.
__global__ void __launch_bounds__(64) k_w6(Dev d, TickParams tp) {
    if (aborted(d)) return;
    __shared__ double work[W6_NMAX];
    __shared__ uint8_t gone[W6_NMAX];
    __shared__ double red[4];
    __shared__ uint32_t shu[4];
    const W6Ws s = w6_slot(d, blockIdx.x);
    for (uint32_t w = blockIdx.x; w < d.nw; w += gridDim.x) {
        if (!d.w_w6[w]) continue;
        w6_world(d, tp, s, w, work, gone, red, shu);
    }
}
