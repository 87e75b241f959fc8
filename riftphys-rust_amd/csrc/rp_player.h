// Kinematic players on the device: World::update_players (step.rs:544-639), run first in each
// tick (step.rs:51), before the forces.
//
// One workgroup per world; the world's players in order.  For each of up to four resolution
// passes the proposed capsule (radius, height / 2) is tested against the terrain triangles under
// its box (the first within 0.05 m in emission order) and, failing that, against the colliders in
// order (the first whose stored box meets the capsule's box and whose shape lies within 0.05 m,
// other than the player's own body): the first hit is the lowest candidate index of the chunk
// (atomicMin), and lane 0 applies the push-out, the grounding and the kick.  The stored boxes
// are the ones before this tick moved any player (the reference refreshes them at the step's
// end), so the players' own colliders keep their boxes from before the pass.
//
// Included by rp_step.hip after rp_query.h.
#pragma once

#define RPP_T 256u

// The state a tick changes, saved before the players move: body state (as k_bodies_pre's backup)
// and the players' controller state.  A tick that overflows a capacity is replayed from here.
__global__ void k_player_backup(Dev d) {
    GRID_STRIDE(b, d.nb) {
        d.bk_pos[b] = d.pos[b]; d.bk_rot[b] = d.rot[b]; d.bk_lin[b] = d.lin[b]; d.bk_ang[b] = d.ang[b];
        d.bk_last[b] = d.last_lin[b]; d.bk_awake[b] = d.awake[b]; d.bk_sticks[b] = d.sticks[b];
    }
    GRID_STRIDE(p, d.np) { d.bk_pl_state[p] = d.pl_state[p]; d.bk_pl_grounded[p] = d.pl_grounded[p]; }
}

__global__ void __launch_bounds__(RPP_T) k_players(Dev d, TickParams tp) {
    if (aborted(d)) return;
    const uint32_t w = blockIdx.x, lane = threadIdx.x;
    const uint32_t p0 = d.w_player_off[w], p1 = d.w_player_off[w + 1];
    if (p0 == p1) return;
    __shared__ uint32_t first;
    __shared__ float sh_n[3], sh_pen;
    __shared__ uint32_t sh_body;
    __shared__ float sh_des[3];
    const uint32_t c0 = d.w_col_off[w], c1 = d.w_col_off[w + 1];
    const HfDesc hf = d.hf[w];
    const float dt = tp.dt;
    const Q4 qi = qid();
    // the players' colliders' boxes before any of them moves
    for (uint32_t p = p0 + lane; p < p1; p += RPP_T) {
        const uint32_t b = d.pl_body[p];
        V3 mn, mx;
        aabb_of(d.btype[b], d.bdims[b], L3(d.pos[b]), LQ(d.rot[b]), mn, mx);
        d.pl_bmin[p] = S3(mn); d.pl_bmax[p] = S3(mx);
    }
    __syncthreads();
    for (uint32_t p = p0; p < p1; p++) {
        const uint32_t pb = d.pl_body[p];
        float4 st = d.pl_state[p];  // radius, height, speed, vertical velocity
        V3 pos = L3(d.pos[pb]);
        if (lane == 0) {
            bool grounded = d.pl_grounded[p] != 0;
            float vv = st.w;
            if (!grounded) vv -= 9.81f * dt;
            else if (vv <= 0.0f) vv = -0.5f;
            const V3 in = L3(d.pl_input[p]);
            const float m2 = len2(in);
            const V3 dir = m2 > 1.0e-6f ? in / sqrtf(m2) : splat3(0.0f);
            const V3 des = (dir * st.z) * dt + mk3(0.0f, vv * dt, 0.0f);
            sh_des[0] = des.x; sh_des[1] = des.y; sh_des[2] = des.z;
            d.pl_state[p].w = vv;
            d.pl_grounded[p] = 0u;
        }
        __syncthreads();
        Shape cap;
        cap.type = 2; cap.d0 = st.x; cap.d1 = st.y * 0.5f; cap.d2 = 0.0f;
        for (int it = 0; it < 4; it++) {
            const V3 des = mk3(sh_des[0], sh_des[1], sh_des[2]);
            if (len2(des) < 1.0e-6f) break;
            const V3 xp = pos + des;
            V3 bmn, bmx;
            aabb_of(2u, make_float4(cap.d0, cap.d1, 0.0f, 0.0f), xp, qi, bmn, bmx);
            if (lane == 0) first = RP_NONE;
            __syncthreads();
            // A. the terrain, first triangle in emission order
            int gx0, gx1, gz0, gz1;
            if (hf_range(hf, bmn, bmx, gx0, gx1, gz0, gz1)) {
                const uint32_t ncx = (uint32_t)(gx1 - gx0 + 1), ncell = ncx * (uint32_t)(gz1 - gz0 + 1);
                for (uint32_t base = 0; base < ncell && first == RP_NONE; base += RPP_T) {
                    const uint32_t e = base + lane;
                    Manifold m;
                    int kh = -1;
                    if (e < ncell) {
                        const int gx = gx0 + (int)(e % ncx), gz = gz0 + (int)(e / ncx);
                        for (int k = 0; k < 2 && kh < 0; k++) {
                            V3 ta, tb, tc;
                            hf_tri(d, hf, gx, gz, k, ta, tb, tc);
                            if (collide_terrain(cap, xp, qi, ta, tb, tc, 0.05f, m)) kh = k;
                        }
                    }
                    if (kh >= 0) atomicMin(&first, 2u * e + (uint32_t)kh);
                    __syncthreads();
                    if (kh >= 0 && first == 2u * e + (uint32_t)kh) {
                        sh_n[0] = m.p[0].n.x; sh_n[1] = m.p[0].n.y; sh_n[2] = m.p[0].n.z; sh_pen = m.p[0].pen; sh_body = RP_NONE;
                    }
                    __syncthreads();
                }
            }
            // B. the colliders in order
            if (first == RP_NONE) {
                for (uint32_t base = c0; base < c1 && first == RP_NONE; base += RPP_T) {
                    const uint32_t c = base + lane;
                    Manifold m;
                    bool hit = false;
                    if (c < c1) {
                        const uint32_t b = d.col_body[c];
                        if (b != pb) {
                            V3 amn, amx;
                            const uint32_t op = d.pl_of_body[b];
                            if (op != RP_NONE) { amn = L3(d.pl_bmin[op]); amx = L3(d.pl_bmax[op]); }
                            else aabb_of(d.btype[b], d.bdims[b], L3(d.pos[b]), LQ(d.rot[b]), amn, amx);
                            // Aabb::intersects (aabb.rs:75-79)
                            const bool meet = bmn.x <= amx.x && bmx.x >= amn.x && bmn.y <= amx.y && bmx.y >= amn.y && bmn.z <= amx.z &&
                                              bmx.z >= amn.z;
                            hit = meet && collide_pair(cap, xp, qi, shape_of(d, b), L3(d.pos[b]), LQ(d.rot[b]), 0.05f, m);
                        }
                    }
                    if (hit) atomicMin(&first, c);
                    __syncthreads();
                    if (hit && first == c) {
                        sh_n[0] = m.p[0].n.x; sh_n[1] = m.p[0].n.y; sh_n[2] = m.p[0].n.z; sh_pen = m.p[0].pen; sh_body = d.col_body[c];
                    }
                    __syncthreads();
                }
            }
            if (first == RP_NONE) break;
            // C. resolution (lane 0)
            if (lane == 0) {
                const V3 n = mk3(sh_n[0], sh_n[1], sh_n[2]);
                if (n.y > 0.6f) { d.pl_grounded[p] = 1u; d.pl_state[p].w = 0.0f; }
                V3 dd = des + n * (sh_pen + 0.001f);
                const float dot = dot3(dd, n);
                if (dot < 0.0f) dd = dd - n * dot;
                sh_des[0] = dd.x; sh_des[1] = dd.y; sh_des[2] = dd.z;
                const uint32_t hb = sh_body;
                if (hb != RP_NONE && (d.bflags[hb] & BF_DYN)) {
                    d.awake[hb] = 1u; d.sticks[hb] = 0u;
                    const V3 j = (-n) * 10.0f;
                    const float im = d.inv_mass[hb];
                    if (im > 0.0f && fin3(j)) d.lin[hb] = S3(fin_v(L3(d.lin[hb]) + j * im));
                }
            }
            __syncthreads();
        }
        if (lane == 0) set_pose(d, pb, pos + mk3(sh_des[0], sh_des[1], sh_des[2]), qi);
        __syncthreads();
    }
}
