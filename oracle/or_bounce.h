/* Oracle (test infrastructure only): the numerics of the W6 bounce passes, restated from
 * crates/riftphys-world/src/world/bounce_lcp.rs:1-321 and the row helpers of
 * crates/riftphys-world/src/world/solver.rs:121-192.
 *
 * f64 throughout, every sum in the reference's order:
 *   or_bside           row_sides        solver.rs:123-139  (one side: body, J lin/ang, M^-1 J^T)
 *   bn_contact_csr     contact_csr      solver.rs:142-168  (K = J M^-1 J^T, rows couple through
 *                                                           a shared moving body)
 *   bn_components      row_components   solver.rs:174-192
 *   bn_min_degree      Symbolic::min_degree  bounce_lcp.rs:83-147 (minimum degree, lowest index
 *                                                           on ties; the degree set is held as
 *                                                           256-bit rows, DIRECT_MAX = 250)
 *   bn_factor          factor (sparse, left-looking)    bounce_lcp.rs:163-203
 *   bn_solve           solve (sparse)                   bounce_lcp.rs:205-229
 *   bn_bounce_qp       bounce_qp (Mehrotra interior point on K + CFM diag(K))  :256-321
 *   bn_pseudo_solve    pseudo_solve                     bounce_lcp.rs:323-338
 * The dense factorisation and the backward-error / dense-check diagnostics of the reference feed
 * no result and are left out.
 */
#ifndef OR_BOUNCE_H
#define OR_BOUNCE_H
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define BN_DIRECT_MAX 250u  /* solver.rs:94 DIRECT_MAX */
#define BN_SWEEPS 256u      /* solver.rs:97 BOUNCE_SWEEPS */
#define BN_TOL 1.0e-6f      /* solver.rs:98 BOUNCE_TOL */
#define BN_SOLVES 64u       /* solver.rs:100 BOUNCE_SOLVES */
#define BN_CFM 1.0e-12      /* bounce_lcp.rs:239 CFM */
#define BN_WORDS 4          /* 256-bit adjacency rows */
#define BN_NONE 0xFFFFFFFFu

typedef struct { double x, y, z; } D3;
static inline D3 d3(double x, double y, double z) { D3 r = {x, y, z}; return r; }
static inline D3 d3add(D3 a, D3 b) { return d3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline D3 d3mul(D3 a, double s) { return d3(a.x * s, a.y * s, a.z * s); }
/* glam DVec3::dot: (x*x' + y*y') + z*z' */
static inline double d3dot(D3 a, D3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
/* glam DVec3::cross */
static inline D3 d3cross(D3 a, D3 b) { return d3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y); }

/* One side of a contact row (solver.rs:121 RowSide): body (BN_NONE if it cannot move), J's
 * linear and angular parts, M^-1 J^T's linear and angular parts. */
typedef struct { uint32_t body; D3 jl, ja, ml, ma; } BSide;

/* solver.rs:123-139: n and r widened to f64; the f32 world inverse inertia widened and
 * symmetrised as (M + M^T) * 0.5; M M^T's product as DMat3::mul_vec3 ((c0 x + c1 y) + c2 z). */
static BSide or_bside(uint32_t body, double sign, V3 r, float inv, M3 ii, V3 nf) {
    BSide s;
    memset(&s, 0, sizeof(s));
    if (inv <= 0.0f && m3iszero(ii)) { s.body = BN_NONE; return s; }
    D3 n = d3(nf.x, nf.y, nf.z);
    D3 jl = d3mul(n, sign);
    D3 ja = d3mul(d3cross(d3(r.x, r.y, r.z), n), sign);
    double m[3][3]; /* column-major m[col][row] */
    double f[3][3] = {{ii.c0.x, ii.c0.y, ii.c0.z}, {ii.c1.x, ii.c1.y, ii.c1.z}, {ii.c2.x, ii.c2.y, ii.c2.z}};
    for (int c = 0; c < 3; c++)
        for (int rr = 0; rr < 3; rr++) m[c][rr] = (f[c][rr] + f[rr][c]) * 0.5;
    D3 c0 = d3(m[0][0], m[0][1], m[0][2]), c1 = d3(m[1][0], m[1][1], m[1][2]), c2 = d3(m[2][0], m[2][1], m[2][2]);
    s.body = body;
    s.jl = jl;
    s.ja = ja;
    s.ml = d3mul(jl, (double)inv);
    s.ma = d3add(d3add(d3mul(c0, ja.x), d3mul(c1, ja.y)), d3mul(c2, ja.z));
    return s;
}

/* A symmetric sparse matrix, both triangles, each row's columns ascending, diagonal present. */
typedef struct { uint32_t n; uint32_t *ptr, *idx; double *val; } BCsr;
static void bn_csr_free(BCsr *k) { free(k->ptr); free(k->idx); free(k->val); memset(k, 0, sizeof(*k)); }

typedef struct { uint32_t body, row, t; } BEnt;
static int bn_ent_cmp(const void *x, const void *y) {
    const BEnt *a = (const BEnt *)x, *b = (const BEnt *)y;
    if (a->body != b->body) return a->body < b->body ? -1 : 1;
    if (a->row != b->row) return a->row < b->row ? -1 : 1;
    return a->t < b->t ? -1 : (a->t > b->t);
}
static int bn_u32_cmp(const void *x, const void *y) {
    uint32_t a = *(const uint32_t *)x, b = *(const uint32_t *)y;
    return a < b ? -1 : (a > b);
}

/* contact_csr (solver.rs:142-168): per row i, for each moving side of i (side order), for each
 * (row j, side t) on the same body in (row, side) order: acc[j] += Ji.(M^-1 Jj^T), the linear and
 * angular dot products added first.  The touched columns ascending. */
static BCsr bn_contact_csr(const BSide (*sides)[2], uint32_t n) {
    BEnt *ents = (BEnt *)malloc(sizeof(BEnt) * (2 * n + 1));
    uint32_t ne = 0;
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t t = 0; t < 2; t++)
            if (sides[i][t].body != BN_NONE) { ents[ne].body = sides[i][t].body; ents[ne].row = i; ents[ne].t = t; ne++; }
    qsort(ents, ne, sizeof(BEnt), bn_ent_cmp);
    double *acc = (double *)calloc(n + 1, sizeof(double));
    uint8_t *seen = (uint8_t *)calloc(n + 1, 1);
    uint32_t *touched = (uint32_t *)malloc(sizeof(uint32_t) * (n + 1));
    BCsr k;
    k.n = n;
    k.ptr = (uint32_t *)malloc(sizeof(uint32_t) * (n + 1));
    uint32_t cap = 4 * n + 4, nnz = 0;
    k.idx = (uint32_t *)malloc(sizeof(uint32_t) * cap);
    k.val = (double *)malloc(sizeof(double) * cap);
    k.ptr[0] = 0;
    for (uint32_t i = 0; i < n; i++) {
        uint32_t nt = 0;
        touched[nt++] = i;
        seen[i] = 1;
        for (uint32_t t = 0; t < 2; t++) {
            const BSide *si = &sides[i][t];
            if (si->body == BN_NONE) continue;
            uint32_t lo = 0, hi = ne;
            while (lo < hi) { uint32_t mid = (lo + hi) / 2; if (ents[mid].body < si->body) lo = mid + 1; else hi = mid; }
            for (uint32_t e = lo; e < ne && ents[e].body == si->body; e++) {
                uint32_t j = ents[e].row;
                const BSide *sj = &sides[j][ents[e].t];
                acc[j] += d3dot(si->jl, sj->ml) + d3dot(si->ja, sj->ma);
                if (!seen[j]) { seen[j] = 1; touched[nt++] = j; }
            }
        }
        qsort(touched, nt, sizeof(uint32_t), bn_u32_cmp);
        if (nnz + nt > cap) {
            cap = 2 * (nnz + nt);
            k.idx = (uint32_t *)realloc(k.idx, sizeof(uint32_t) * cap);
            k.val = (double *)realloc(k.val, sizeof(double) * cap);
        }
        for (uint32_t q = 0; q < nt; q++) {
            uint32_t j = touched[q];
            k.idx[nnz] = j; k.val[nnz] = acc[j]; nnz++;
            acc[j] = 0.0; seen[j] = 0;
        }
        k.ptr[i + 1] = nnz;
    }
    free(ents); free(acc); free(seen); free(touched);
    return k;
}

static uint32_t bn_find(uint32_t *p, uint32_t x) {
    while (p[x] != x) { uint32_t g = p[p[x]]; p[x] = g; x = g; }
    return x;
}

/* row_components (solver.rs:174-192): union of the two moving sides of each row (the larger root
 * points at the smaller), then rows grouped by their first moving side's root in first-row order;
 * a row with no moving side is a component alone.  comp_of[i] receives the component of row i,
 * rows ascend within each; returns the component count. */
static uint32_t bn_components(const BSide (*sides)[2], uint32_t n, uint32_t *comp_of) {
    /* body ids compacted to 0..nb-1 */
    uint32_t *ids = (uint32_t *)malloc(sizeof(uint32_t) * (2 * n + 1)), nid = 0;
    for (uint32_t i = 0; i < n; i++)
        for (int t = 0; t < 2; t++) if (sides[i][t].body != BN_NONE) ids[nid++] = sides[i][t].body;
    qsort(ids, nid, sizeof(uint32_t), bn_u32_cmp);
    uint32_t nb = 0;
    for (uint32_t i = 0; i < nid; i++) if (i == 0 || ids[i] != ids[i - 1]) ids[nb++] = ids[i];
#define BN_IX(b) ((uint32_t)((const uint32_t *)bsearch(&(b), ids, nb, sizeof(uint32_t), bn_u32_cmp) - ids))
    /* the reference's parent array is indexed by body id; the roots (smallest body id of each set)
     * map to the smallest compact index, so unions and finds agree */
    uint32_t *par = (uint32_t *)malloc(sizeof(uint32_t) * (nb + 1));
    for (uint32_t i = 0; i < nb; i++) par[i] = i;
    for (uint32_t i = 0; i < n; i++) {
        if (sides[i][0].body == BN_NONE || sides[i][1].body == BN_NONE) continue;
        uint32_t a = bn_find(par, BN_IX(sides[i][0].body)), b = bn_find(par, BN_IX(sides[i][1].body));
        if (a != b) par[a > b ? a : b] = a < b ? a : b;
    }
    uint32_t *slot = (uint32_t *)malloc(sizeof(uint32_t) * (nb + 1));
    for (uint32_t i = 0; i < nb; i++) slot[i] = BN_NONE;
    uint32_t nc = 0;
    for (uint32_t i = 0; i < n; i++) {
        uint32_t body = sides[i][0].body != BN_NONE ? sides[i][0].body : sides[i][1].body;
        if (body == BN_NONE) { comp_of[i] = nc++; continue; }
        uint32_t r = bn_find(par, BN_IX(body));
        if (slot[r] == BN_NONE) slot[r] = nc++;
        comp_of[i] = slot[r];
    }
#undef BN_IX
    free(ids); free(par); free(slot);
    return nc;
}

/* Symbolic::min_degree (bounce_lcp.rs:83-147). */
typedef struct {
    uint32_t n;
    uint32_t order[BN_DIRECT_MAX + 6], pos[BN_DIRECT_MAX + 6];
    uint32_t *cp, *ri, *rp, *re, *ce;
} BSym;
static void bn_sym_free(BSym *s) { free(s->cp); free(s->ri); free(s->rp); free(s->re); free(s->ce); }
static int bn_popc(const uint64_t *b) { int c = 0; for (int k = 0; k < BN_WORDS; k++) c += __builtin_popcountll(b[k]); return c; }

static void bn_min_degree(const BCsr *k, BSym *s) {
    uint32_t n = k->n;
    s->n = n;
    uint64_t (*adj)[BN_WORDS] = calloc(n + 1, sizeof(*adj));
    uint64_t (*nbs)[BN_WORDS] = calloc(n + 1, sizeof(*nbs));
    uint8_t *gone = (uint8_t *)calloc(n + 1, 1);
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t t = k->ptr[i]; t < k->ptr[i + 1]; t++)
            if (k->idx[t] != i) adj[i][k->idx[t] >> 6] |= 1ull << (k->idx[t] & 63);
    for (uint32_t step = 0; step < n; step++) {
        /* pop_first of the (degree, index) set */
        uint32_t v = BN_NONE; int dv = 0;
        for (uint32_t u = 0; u < n; u++) {
            if (gone[u]) continue;
            int d = bn_popc(adj[u]);
            if (v == BN_NONE || d < dv) { v = u; dv = d; }
        }
        gone[v] = 1;
        s->order[step] = v;
        memcpy(nbs[v], adj[v], sizeof(adj[v]));
        memset(adj[v], 0, sizeof(adj[v]));
        /* the neighbours become a clique; v leaves their lists */
        for (uint32_t u = 0; u < n; u++) {
            if (!((nbs[v][u >> 6] >> (u & 63)) & 1)) continue;
            for (int w = 0; w < BN_WORDS; w++) adj[u][w] |= nbs[v][w];
            adj[u][v >> 6] &= ~(1ull << (v & 63));
            adj[u][u >> 6] &= ~(1ull << (u & 63));
        }
    }
    for (uint32_t p = 0; p < n; p++) s->pos[s->order[p]] = p;
    s->cp = (uint32_t *)malloc(sizeof(uint32_t) * (n + 1));
    uint32_t tot = 0;
    for (uint32_t p = 0; p < n; p++) tot += (uint32_t)bn_popc(nbs[s->order[p]]);
    s->ri = (uint32_t *)malloc(sizeof(uint32_t) * (tot + 1));
    s->re = (uint32_t *)malloc(sizeof(uint32_t) * (tot + 1));
    s->ce = (uint32_t *)malloc(sizeof(uint32_t) * (tot + 1));
    s->rp = (uint32_t *)calloc(n + 1, sizeof(uint32_t));
    s->cp[0] = 0;
    uint32_t e = 0;
    for (uint32_t p = 0; p < n; p++) {
        uint32_t r = s->order[p], start = e;
        for (uint32_t u = 0; u < n; u++)
            if ((nbs[r][u >> 6] >> (u & 63)) & 1) s->ri[e++] = s->pos[u];
        qsort(s->ri + start, e - start, sizeof(uint32_t), bn_u32_cmp);
        s->cp[p + 1] = e;
    }
    uint32_t *count = (uint32_t *)calloc(n + 2, sizeof(uint32_t));
    for (uint32_t t = 0; t < tot; t++) count[s->ri[t] + 1]++;
    for (uint32_t p = 0; p < n; p++) count[p + 1] += count[p];
    memcpy(s->rp, count, sizeof(uint32_t) * (n + 1));
    for (uint32_t kc = 0; kc < n; kc++)
        for (uint32_t t = s->cp[kc]; t < s->cp[kc + 1]; t++) {
            s->re[count[s->ri[t]]++] = t;
            s->ce[t] = s->cp[kc + 1];
        }
    free(count); free(adj); free(nbs); free(gone);
}

/* factor, sparse (bounce_lcp.rs:182-202): L of K + diag(d) in the symbolic order. */
static void bn_factor(const BCsr *k, const BSym *s, const double *d, double *diag, double *vals) {
    uint32_t n = k->n;
    double *work = (double *)calloc(n + 1, sizeof(double));
    for (uint32_t t = 0; t < s->cp[n]; t++) vals[t] = 0.0;
    for (uint32_t j = 0; j < n; j++) {
        uint32_t r = s->order[j];
        for (uint32_t t = k->ptr[r]; t < k->ptr[r + 1]; t++) {
            uint32_t p = s->pos[k->idx[t]];
            if (p >= j) work[p] += k->val[t];
        }
        work[j] += d[r];
        for (uint32_t q = s->rp[j]; q < s->rp[j + 1]; q++) {
            uint32_t e = s->re[q];
            double ljk = vals[e];
            uint32_t end = s->ce[e];
            for (uint32_t t = e; t < end; t++) work[s->ri[t]] -= vals[t] * ljk;
        }
        double ljj = sqrt(fmax(work[j], 1.0e-300));
        diag[j] = ljj;
        work[j] = 0.0;
        for (uint32_t t = s->cp[j]; t < s->cp[j + 1]; t++) {
            uint32_t p = s->ri[t];
            vals[t] = work[p] / ljj;
            work[p] = 0.0;
        }
    }
    free(work);
}

/* solve, sparse (bounce_lcp.rs:214-228) */
static void bn_solve(const BSym *s, const double *diag, const double *vals, const double *b, double *x) {
    uint32_t n = s->n;
    double *y = (double *)malloc(sizeof(double) * (n + 1));
    for (uint32_t p = 0; p < n; p++) y[p] = b[s->order[p]];
    for (uint32_t j = 0; j < n; j++) {
        y[j] /= diag[j];
        double yj = y[j];
        for (uint32_t t = s->cp[j]; t < s->cp[j + 1]; t++) y[s->ri[t]] -= vals[t] * yj;
    }
    for (uint32_t j = n; j-- > 0;) {
        double v = y[j];
        for (uint32_t t = s->cp[j]; t < s->cp[j + 1]; t++) v -= vals[t] * y[s->ri[t]];
        y[j] = v / diag[j];
    }
    for (uint32_t p = 0; p < n; p++) x[s->order[p]] = y[p];
    free(y);
}

static double bn_diag(const BCsr *k, uint32_t i) {
    for (uint32_t t = k->ptr[i]; t < k->ptr[i + 1]; t++) if (k->idx[t] == i) return k->val[t];
    return 0.0;
}
static double bn_row_dot(const BCsr *k, uint32_t i, const double *x) {
    double s = 0.0;
    for (uint32_t t = k->ptr[i]; t < k->ptr[i + 1]; t++) s += k->val[t] * x[k->idx[t]];
    return s;
}

/* step (bounce_lcp.rs:271-278): the largest a <= 1 keeping y + a dy and w + a dw nonnegative */
static double bn_step(uint32_t n, const double *y, const double *dy, const double *w, const double *dw) {
    double a = 1.0;
    for (uint32_t i = 0; i < n; i++) {
        if (dy[i] < 0.0) a = fmin(a, -y[i] / dy[i]);
        if (dw[i] < 0.0) a = fmin(a, -w[i] / dw[i]);
    }
    return a;
}

/* bounce_qp (bounce_lcp.rs:256-321): push-only rows, y = acc0 + x >= 0, w = K y + q >= 0,
 * y w = 0, by Mehrotra predictor-corrector on K + CFM diag(K).  x receives y - acc0; returns 1 if
 * the stopping test held within `cap` solves. */
static int bn_bounce_qp(const BCsr *k0, const double *b, const double *acc0, uint32_t cap, double *x, uint32_t *solves) {
    uint32_t n = k0->n;
    BCsr k;
    k.n = n;
    k.ptr = k0->ptr; k.idx = k0->idx;
    k.val = (double *)malloc(sizeof(double) * (k0->ptr[n] + 1));
    memcpy(k.val, k0->val, sizeof(double) * k0->ptr[n]);
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t t = k.ptr[i]; t < k.ptr[i + 1]; t++) if (k.idx[t] == i) k.val[t] *= 1.0 + BN_CFM;
    BSym s;
    bn_min_degree(&k, &s);
    size_t m = (size_t)n + 1;
    double *kd = malloc(sizeof(double) * m), *q = malloc(sizeof(double) * m), *y = malloc(sizeof(double) * m),
           *w = malloc(sizeof(double) * m), *r = malloc(sizeof(double) * m), *xt = malloc(sizeof(double) * m),
           *d = malloc(sizeof(double) * m), *rhs = malloc(sizeof(double) * m), *dya = malloc(sizeof(double) * m),
           *dwa = malloc(sizeof(double) * m), *dy = malloc(sizeof(double) * m), *dw = malloc(sizeof(double) * m),
           *tg = malloc(sizeof(double) * m), *diag = malloc(sizeof(double) * m);
    double *vals = malloc(sizeof(double) * (s.cp[n] + 1));
    for (uint32_t i = 0; i < n; i++) kd[i] = bn_diag(k0, i);
    double eta0 = 0.0;
    for (uint32_t i = 0; i < n; i++) {
        q[i] = -bn_row_dot(&k, i, acc0) - b[i];
        eta0 = fmax(eta0, fabs(q[i]));
    }
    double eta = fmax(1.0, eta0);
    for (uint32_t i = 0; i < n; i++) { y[i] = kd[i] > 0.0 ? eta / kd[i] : 1.0; w[i] = eta; }
    int converged = 0;
    for (uint32_t it = 1; it <= cap; it++) {
        *solves = it;
        double rmax = 0.0, mu = 0.0;
        for (uint32_t i = 0; i < n; i++) { r[i] = (bn_row_dot(&k, i, y) + q[i]) - w[i]; rmax = fmax(rmax, fabs(r[i])); }
        for (uint32_t i = 0; i < n; i++) mu += y[i] * w[i];
        mu = mu / (double)n;
        for (uint32_t i = 0; i < n; i++) xt[i] = y[i] - acc0[i];
        int ok = rmax < 1.0e-7;
        for (uint32_t i = 0; ok && i < n; i++) {
            double wt = bn_row_dot(k0, i, xt) - b[i];
            double tol = 1.0e-7 * (1.0 + fabs(b[i]));
            if (!(wt >= -tol && fmin(fabs(wt), y[i] * kd[i]) < tol)) ok = 0;
        }
        if (ok) { converged = 1; break; }
        for (uint32_t i = 0; i < n; i++) d[i] = w[i] / y[i];
        bn_factor(&k, &s, d, diag, vals);
        /* affine direction: target -y w */
        for (uint32_t i = 0; i < n; i++) { tg[i] = -y[i] * w[i]; rhs[i] = -r[i] + tg[i] / y[i]; }
        bn_solve(&s, diag, vals, rhs, dya);
        for (uint32_t i = 0; i < n; i++) dwa[i] = (tg[i] - w[i] * dya[i]) / y[i];
        double aa = bn_step(n, y, dya, w, dwa);
        double mu_aff = 0.0;
        for (uint32_t i = 0; i < n; i++) mu_aff += (y[i] + aa * dya[i]) * (w[i] + aa * dwa[i]);
        mu_aff = mu_aff / (double)n;
        double ratio = mu_aff / fmax(mu, 1.0e-300);
        double sigma = fmin(ratio * ratio * ratio, 1.0);
        /* corrector: target sigma mu - y w - dya dwa */
        for (uint32_t i = 0; i < n; i++) { tg[i] = sigma * mu - y[i] * w[i] - dya[i] * dwa[i]; rhs[i] = -r[i] + tg[i] / y[i]; }
        bn_solve(&s, diag, vals, rhs, dy);
        for (uint32_t i = 0; i < n; i++) dw[i] = (tg[i] - w[i] * dy[i]) / y[i];
        double a = fmin(0.995 * bn_step(n, y, dy, w, dw), 1.0);
        int fin = isfinite(a);
        for (uint32_t i = 0; fin && i < n; i++) fin = isfinite(dy[i]) && isfinite(dw[i]);
        if (!fin) break;
        for (uint32_t i = 0; i < n; i++) {
            y[i] = fmax(y[i] + a * dy[i], 1.0e-300);
            w[i] = fmax(w[i] + a * dw[i], 1.0e-300);
        }
    }
    for (uint32_t i = 0; i < n; i++) x[i] = y[i] - acc0[i];
    free(kd); free(q); free(y); free(w); free(r); free(xt); free(d); free(rhs); free(dya); free(dwa); free(dy); free(dw);
    free(tg); free(diag); free(vals); free(k.val);
    bn_sym_free(&s);
    return converged;
}

/* pseudo_solve (bounce_lcp.rs:330-337): K x = u with K + diag(CFM K_ii) factored. */
static void bn_pseudo_solve(const BCsr *k, const double *u, double *x) {
    uint32_t n = k->n;
    double *d = malloc(sizeof(double) * (n + 1)), *diag = malloc(sizeof(double) * (n + 1));
    for (uint32_t i = 0; i < n; i++) d[i] = BN_CFM * bn_diag(k, i);
    BSym s;
    bn_min_degree(k, &s);
    double *vals = malloc(sizeof(double) * (s.cp[n] + 1));
    bn_factor(k, &s, d, diag, vals);
    bn_solve(&s, diag, vals, u, x);
    free(d); free(diag); free(vals);
    bn_sym_free(&s);
}

/* toi_bounce (solver.rs:197-212), f32 */
static float bn_toi_bounce(float s0, float v0, float a, float e, float dt, float *pseudo, int *has_pseudo) {
    float t_c;
    if (s0 <= 0.0f) t_c = 0.0f;
    else {
        float disc = v0 * v0 - (2.0f * a) * s0;
        float den = disc >= 0.0f ? -v0 + sqrtf(disc) : 0.0f;
        t_c = den > 0.0f ? fminf((2.0f * s0) / den, dt) : dt;
    }
    float v_out = fmaxf((-e) * (v0 + a * t_c), 0.0f);
    float tau = dt - t_c;
    float rise = v_out * tau + ((0.5f * a) * tau) * tau;
    float v_end, s_end;
    if (rise < 0.0f) { v_end = 0.0f; s_end = 0.0f; }
    else { v_end = fmaxf(v_out + a * tau, 0.0f); s_end = rise; }
    *has_pseudo = s0 > 0.0f;
    *pseudo = (s_end - s0) / dt - v_end;
    return v_end;
}
#endif
