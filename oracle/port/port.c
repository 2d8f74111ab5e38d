/*
 * port.c — plain C restatement of sofren2's per-frame render path.
 *
 * TEST INFRASTRUCTURE ONLY: this is the oracle the HIP kernels are checked
 * against.  It is scalar C that replays the reference's 8-lane AVX2 code lane
 * by lane, in the reference's structure (8-meshlet cull groups, 8-triangle
 * batches, 8-pixel raster chunks, 4x2 shading blocks), single-threaded.
 * Each function cites the reference lines it follows (sofren2.c).
 *
 * Arithmetic pins: the reciprocal and reciprocal-square-root estimates are
 * the host's own VRCPPS / VRSQRTPS (the reference's _mm256_rcp_ps /
 * _mm256_rsqrt_ps), conversions are CVTTSS2SI (INT_MIN out of range, like
 * _mm256_cvttps_epi32), FMAs are explicit, min/max follow SSE operand rules,
 * and the file is compiled with -ffp-contract=off.
 */
#include <immintrin.h>
#include <stdlib.h>
#include <string.h>
#include "port.h"

#define TS SFT_TILE_SIZE
#define TP SFT_TILE_PIXELS

/* ------------------------------------------------------------ primitives
   (sofren2.c:263-385, 632-642, 1921-2030) */
static inline f32 pfma(f32 a, f32 b, f32 c) { return __builtin_fmaf(a, b, c); }
static inline f32 smin(f32 a, f32 b) { return a < b ? a : b; }     /* _mm256_min_ps(a, b) */
static inline f32 smax(f32 a, f32 b) { return a > b ? a : b; }     /* _mm256_max_ps(a, b) */
static inline i32 cvtt(f32 x) { return _mm_cvtt_ss2si(_mm_set_ss(x)); }
static inline f32 rcp_est(f32 a) { return _mm256_cvtss_f32(_mm256_rcp_ps(_mm256_set1_ps(a))); }
static inline f32 rsq_est(f32 a) { return _mm256_cvtss_f32(_mm256_rsqrt_ps(_mm256_set1_ps(a))); }
/* estimate * (2 - a * estimate), sofren2.c:633-638 */
static inline f32 rcp_refined(f32 a) { const f32 x = rcp_est(a); return x * pfma(-a, x, 2.f); }

typedef struct { f32 x, y, z; } v3;
static inline f32 vdot(v3 a, v3 b) { return a.x * b.x + (a.y * b.y + a.z * b.z); }
static inline v3 vsub(v3 a, v3 b) { v3 r = { a.x - b.x, a.y - b.y, a.z - b.z }; return r; }
static inline v3 vadd(v3 a, v3 b) { v3 r = { a.x + b.x, a.y + b.y, a.z + b.z }; return r; }
/* sft_vv_norm_fast, sofren2.c:1973-1992 */
static inline v3 norm_fast(v3 v) {
    const f32 sq = vdot(v, v);
    const f32 ap = rsq_est(sq);
    const f32 r = (0.5f * ap) * pfma(-sq, ap * ap, 3.f);
    v3 o = { v.x * r, v.y * r, v.z * r };
    return o;
}
/* sft_vv_transform, sofren2.c:1993-2012 */
static inline v3 xf(v3 v, const sftmat* m) {
    v3 r;
    r.x = pfma(v.x, m->m[0][0], pfma(v.y, m->m[1][0], pfma(v.z, m->m[2][0], m->m[3][0])));
    r.y = pfma(v.x, m->m[0][1], pfma(v.y, m->m[1][1], pfma(v.z, m->m[2][1], m->m[3][1])));
    r.z = pfma(v.x, m->m[0][2], pfma(v.y, m->m[1][2], pfma(v.z, m->m[2][2], m->m[3][2])));
    return r;
}
/* sft_vv_transform_dir, sofren2.c:2013-2030 */
static inline v3 xf_dir(v3 v, const sftmat* m) {
    v3 r;
    r.x = pfma(v.x, m->m[0][0], pfma(v.y, m->m[1][0], v.z * m->m[2][0]));
    r.y = pfma(v.x, m->m[0][1], pfma(v.y, m->m[1][1], v.z * m->m[2][1]));
    r.z = pfma(v.x, m->m[0][2], pfma(v.y, m->m[1][2], v.z * m->m[2][2]));
    return r;
}

/* ------------------------------------------------------------ frame state */
typedef struct { u64* p; u32 n, cap; } bin_t;

typedef struct {
    f32 sx[3], sy[3], inv_w[3], u[3], v[3], lu[3], lv[3], px[3], py[3], pz[3], nx[3], ny[3], nz[3];
    u32 inst_ind, m_ind;
} clip_rec;

typedef struct {
    sftvec posa, posb;
    f32 radius;
    u8 light_vis[64];
} world_proxy;

typedef struct {
    const sftx_frame_desc* f;
    sftx_port_out* out;
    f32* ss;            /* [total_ss][9][64] */
    bin_t* bins;
    u32 clip_next;
    int clip_overflow;
    world_proxy* proxies;
    struct sft_bvh_node* tlas;
    u32 tlas_count, tlas_root;
    int invalid_packets;    /* packet frusta the reference would leave uninitialised */
} port_state;

static void bin_push(bin_t* b, u64 v) {
    if (b->n == b->cap) {
        b->cap = b->cap ? b->cap * 2 : 64;
        b->p = (u64*)realloc(b->p, b->cap * sizeof(u64));
    }
    b->p[b->n++] = v;
}

/* sft__geom_bin_tiles, sofren2.c:4650-4688 */
static void bin_tiles(port_state* ps, u64 payload, i32 x0, i32 x1, i32 y0, i32 y1) {
    const i32 mx = ps->f->tile_cols - 1, my = ps->f->tile_rows - 1;
    if (x1 < 0 || x0 > mx || y1 < 0 || y0 > my) return;
    x0 = SFT_MAX(x0, 0); x1 = SFT_MIN(x1, mx);
    y0 = SFT_MAX(y0, 0); y1 = SFT_MIN(y1, my);
    for (i32 ty = y0; ty <= y1; ty += 1) {
        for (i32 tx = x0; tx <= x1; tx += 1) bin_push(&ps->bins[ty * ps->f->tile_cols + tx], payload);
    }
}

static inline f32* ss_at(port_state* ps, u32 ss_ind) { return ps->ss + (u64)ss_ind * 576u; }

/* ------------------------------------------------------------ near-plane clip
   sofren2.c:4245-4349 */
typedef struct { f32 tx, ty, tz, w, u, v, lu, lv, px, py, pz, nx, ny, nz; } cvert;

static cvert lerp_cv(cvert a, cvert b, f32 t) {
    cvert r;
    r.tx = a.tx + t * (b.tx - a.tx); r.ty = a.ty + t * (b.ty - a.ty);
    r.tz = a.tz + t * (b.tz - a.tz); r.w = a.w + t * (b.w - a.w);
    r.u = a.u + t * (b.u - a.u); r.v = a.v + t * (b.v - a.v);
    r.lu = a.lu + t * (b.lu - a.lu); r.lv = a.lv + t * (b.lv - a.lv);
    r.px = a.px + t * (b.px - a.px); r.py = a.py + t * (b.py - a.py); r.pz = a.pz + t * (b.pz - a.pz);
    r.nx = a.nx + t * (b.nx - a.nx); r.ny = a.ny + t * (b.ny - a.ny); r.nz = a.nz + t * (b.nz - a.nz);
    return r;
}

static int clip_near(cvert out[2][3], const cvert in[3]) {
    int mask = 0;
    for (int i = 0; i < 3; i += 1) mask |= (in[i].tz >= 0.f) << i;
    switch (mask) {
        case 0: return 0;
        case 7: out[0][0] = in[0]; out[0][1] = in[1]; out[0][2] = in[2]; return 1;
        case 1: case 2: case 4: {
            const int a = mask >> 1, b = (2 == a) ? 0 : a + 1, c = (0 == a) ? 2 : a - 1;
            const f32 da = in[a].tz;
            const f32 tab = da / (da - in[b].tz), tac = da / (da - in[c].tz);
            out[0][0] = in[a];
            out[0][1] = lerp_cv(in[a], in[b], tab);
            out[0][2] = lerp_cv(in[a], in[c], tac);
            return 1;
        }
        default: {
            const int o = (~mask & 7) >> 1, a = (0 == o) ? 2 : o - 1, c = (2 == o) ? 0 : o + 1;
            const f32 da = in[a].tz, db = in[o].tz, dc = in[c].tz;
            const f32 tab = da / (da - db), tcb = dc / (dc - db);
            const cvert nab = lerp_cv(in[a], in[o], tab), ncb = lerp_cv(in[c], in[o], tcb);
            out[0][0] = in[a]; out[0][1] = nab; out[0][2] = in[c];
            out[1][0] = nab; out[1][1] = ncb; out[1][2] = in[c];
            return 2;
        }
    }
}

/* ------------------------------------------------------------ geometry
   sofren2.c:4691-4884 (clip + setup + bin), 4901-5145 (cull, skin, project) */
typedef struct {
    f32 tx[64], ty[64], tz[64], w[64], sx[64], sy[64];
} clip_scratch;

static void clip_and_bin(port_state* ps, const u8 i[3][8], int lanes, const clip_scratch* cs, u32 inst_ind,
                         u32 m_ind, const struct sft_meshlet_attribs* at, const f32* ssm) {
    const sftx_frame_desc* f = ps->f;
    for (int lane = 0; lane < 8; lane += 1) {
        if (!(lanes & (1 << lane))) continue;
        cvert in[3];
        for (int k = 0; k < 3; k += 1) {
            const int v = i[k][lane];
            cvert c = { cs->tx[v], cs->ty[v], cs->tz[v], cs->w[v], at->u[v], at->v[v], at->lu[v], at->lv[v],
                        ssm[192 + v], ssm[256 + v], ssm[320 + v], ssm[384 + v], ssm[448 + v], ssm[512 + v] };
            in[k] = c;
        }
        cvert out[2][3];
        const int n = clip_near(out, in);
        for (int c = 0; c < n; c += 1) {
            const u32 slot = ps->clip_next++;
            if (slot >= ps->out->clip_cap) {
                ps->clip_overflow = 1;
                continue;
            }
            clip_rec* ct = (clip_rec*)(ps->out->clips + (u64)slot * sizeof(clip_rec));
            for (int k = 0; k < 3; k += 1) {
                const f32 iw = 1.f / out[c][k].w;
                ct->sx[k] = (out[c][k].tx * iw + 1.f) * f->half_width + f->vp_x;
                ct->sy[k] = (-out[c][k].ty * iw + 1.f) * f->half_height + f->vp_y;
                ct->inv_w[k] = iw;
            }
            ct->inst_ind = inst_ind;
            ct->m_ind = m_ind;
            for (int k = 0; k < 3; k += 1) {
                ct->u[k] = out[c][k].u; ct->v[k] = out[c][k].v;
                ct->lu[k] = out[c][k].lu; ct->lv[k] = out[c][k].lv;
                ct->px[k] = out[c][k].px; ct->py[k] = out[c][k].py; ct->pz[k] = out[c][k].pz;
                ct->nx[k] = out[c][k].nx; ct->ny[k] = out[c][k].ny; ct->nz[k] = out[c][k].nz;
            }
            const f32 area = (ct->sx[1] - ct->sx[0]) * (ct->sy[2] - ct->sy[0]) - (ct->sy[1] - ct->sy[0]) * (ct->sx[2] - ct->sx[0]);
            if (area <= 0.1f) continue;
            const i32 x0 = cvtt(floorf(SFT_MIN(ct->sx[0], SFT_MIN(ct->sx[1], ct->sx[2]))));
            const i32 x1 = cvtt(ceilf(SFT_MAX(ct->sx[0], SFT_MAX(ct->sx[1], ct->sx[2]))));
            const i32 y0 = cvtt(floorf(SFT_MIN(ct->sy[0], SFT_MIN(ct->sy[1], ct->sy[2]))));
            const i32 y1 = cvtt(ceilf(SFT_MAX(ct->sy[0], SFT_MAX(ct->sy[1], ct->sy[2]))));
            bin_tiles(ps, (u64)slot, x0 >> 6, x1 >> 6, y0 >> 6, y1 >> 6);
        }
    }
}

static void setup_and_bin_batch(port_state* ps, u32 batch, u32 inst_ind, u32 m_ind, const struct sft_meshlet_inds* mi,
                                const clip_scratch* cs, const struct sft_meshlet_attribs* at, const f32* ssm) {
    u8 i[3][8];
    int valid = 0;
    for (int l = 0; l < 8; l += 1) {
        for (int k = 0; k < 3; k += 1) i[k][l] = mi->inds[batch * 32 + k * 8 + l];
        if ((i[0][l] | i[1][l] | i[2][l]) > 0) valid |= 1 << l;
    }
    if (!valid) return;
    int inside = 0, inter = 0;
    for (int l = 0; l < 8; l += 1) {
        const f32 z0 = cs->tz[i[0][l]], z1 = cs->tz[i[1][l]], z2 = cs->tz[i[2][l]];
        const int all_in = z0 >= 0.f && z1 >= 0.f && z2 >= 0.f;
        const int all_out = z0 < 0.f && z1 < 0.f && z2 < 0.f;
        if (all_in) inside |= 1 << l;
        if ((valid & (1 << l)) && !(all_in || all_out)) inter |= 1 << l;
    }
    if (inter) clip_and_bin(ps, (const u8(*)[8])i, inter, cs, inst_ind, m_ind, at, ssm);
    if (!inside) return;
    const i32 mx = ps->f->tile_cols - 1, my = ps->f->tile_rows - 1;
    int tris = 0;
    i32 tx0[8], tx1[8], ty0[8], ty1[8];
    for (int l = 0; l < 8; l += 1) {
        const f32 sx0 = cs->sx[i[0][l]], sy0 = cs->sy[i[0][l]];
        const f32 sx1 = cs->sx[i[1][l]], sy1 = cs->sy[i[1][l]];
        const f32 sx2 = cs->sx[i[2][l]], sy2 = cs->sy[i[2][l]];
        const f32 dx1 = sx1 - sx0, dy1 = sy1 - sy0, dx2 = sx2 - sx0, dy2 = sy2 - sy0;
        const int area_ok = pfma(dx1, dy2, -(dy1 * dx2)) > 0.1f;
        tx0[l] = cvtt(smin(sx0, smin(sx1, sx2))) >> 6;
        tx1[l] = cvtt(smax(sx0, smax(sx1, sx2))) >> 6;
        ty0[l] = cvtt(smin(sy0, smin(sy1, sy2))) >> 6;
        ty1[l] = cvtt(smax(sy0, smax(sy1, sy2))) >> 6;
        const int in_grid = tx1[l] >= 0 && tx0[l] <= mx && ty1[l] >= 0 && ty0[l] <= my;
        if ((inside & (1 << l)) && area_ok && in_grid) tris |= 1 << l;
    }
    if (!tris) return;
    i32 bx0 = 9999, bx1 = -9999, by0 = 9999, by1 = -9999;
    for (int l = 0; l < 8; l += 1) {
        if (!(tris & (1 << l))) continue;
        if (tx0[l] < bx0) bx0 = tx0[l];
        if (tx1[l] > bx1) bx1 = tx1[l];
        if (ty0[l] < by0) by0 = ty0[l];
        if (ty1[l] > by1) by1 = ty1[l];
    }
    const u64 payload = (u64)inst_ind | ((u64)m_ind << 24) | ((u64)batch << 48) | ((u64)tris << 56);
    bin_tiles(ps, payload, bx0, bx1, by0, by1);
}

/* frustum + normal-cone test of one meshlet, culling volume moved by `bone`
   (the 8-meshlet group's first culling bone) then the instance (4939-4994) */
static int cull_meshlet(const sftx_frame_desc* f, const sft_mesh* mesh, u32 m, const sftmat* T, const sftmat* bone) {
    v3 c = { mesh->bounds_center_x[m], mesh->bounds_center_y[m], mesh->bounds_center_z[m] };
    v3 a = { mesh->cone_axis_x[m], mesh->cone_axis_y[m], mesh->cone_axis_z[m] };
    if (bone) {
        c = xf(c, bone);
        a = xf_dir(a, bone);
    }
    c = xf(c, T);
    a = xf_dir(a, T);
    const f32 sx = sqrtf(T->m[0][0] * T->m[0][0] + T->m[0][1] * T->m[0][1] + T->m[0][2] * T->m[0][2]);
    const f32 radius = mesh->bounds_radius[m] * sx;
    f32 len = sqrtf(vdot(a, a));
    len = smax(len, SFT_EPSILON);
    a.x = a.x / len; a.y = a.y / len; a.z = a.z / len;
    int inside = 1;
    for (int p = 0; p < 6; p += 1) {
        const v3 n = { f->planes[p][0], f->planes[p][1], f->planes[p][2] };
        const f32 dist = vdot(c, n) + f->planes[p][3];
        inside = inside && (dist + radius >= 0.f);
    }
    const v3 cam = { f->camera_pos.x, f->camera_pos.y, f->camera_pos.z };
    const v3 raw = vsub(c, cam);
    const f32 d = sqrtf(vdot(raw, raw));
    inside = inside && (vdot(raw, a) <= pfma(d, mesh->cone_cutoff[m], radius));
    return inside;
}

static void geometry_pass(port_state* ps) {
    const sftx_frame_desc* f = ps->f;
    const sftmat* vp = &f->view_proj;
    const f32 sxs = f->half_width, sxb = f->half_width + f->vp_x;
    const f32 sys = -f->half_height, syb = f->half_height + f->vp_y;
    for (i32 ii = 0; ii < f->n_instances; ii += 1) {
        const sft_instance* inst = &f->instances[ii];
        const sft_model* model = inst->model;
        for (u32 mi = 0; mi < model->mesh_count; mi += 1) {
            const sft_mesh* mesh = &model->meshes[mi];
            for (u32 start = 0; start < mesh->meshlet_count; start += SFT_MESHLETS_PER_JOB) {
                const sftmat* bone = NULL;
                if (mesh->culling_bone_inds && inst->posed_bone_matrices) {
                    bone = &inst->posed_bone_matrices[mesh->culling_bone_inds[start]];
                }
                for (u32 lane = 0; lane < 8 && start + lane < mesh->meshlet_count; lane += 1) {
                    const u32 m = start + lane;
                    const u32 gm = mesh->model_meshlet_offset + m;
                    const u32 ss_ind = inst->ss_meshlet_offset + gm;
                    const int vis = cull_meshlet(f, mesh, m, &inst->transform, bone);
                    ps->out->ss_vis[ss_ind] = (u8)vis;
                    if (!vis) continue;
                    const struct sft_meshlet* mv = &mesh->meshlets[m];
                    const struct sft_meshlet_attribs* at = &mesh->meshlet_attribs[m];
                    f32* ssm = ss_at(ps, ss_ind);
                    clip_scratch cs;
                    for (int v = 0; v < 64; v += 1) {
                        const v3 p = { mv->x[v], mv->y[v], mv->z[v] };
                        const v3 n = { at->nx[v], at->ny[v], at->nz[v] };
                        v3 pw, nw;
                        if (mesh->meshlet_skins && inst->posed_bone_matrices) {
                            /* 4-joint skinning, weights u8/255 (5026-5080) */
                            const struct sft_meshlet_skin* sk = &mesh->meshlet_skins[m];
                            const f32 inv255 = 1.f / 255.f;
                            const u16 j[4] = { sk->j0[v], sk->j1[v], sk->j2[v], sk->j3[v] };
                            const f32 w[4] = { sk->w0[v] * inv255, sk->w1[v] * inv255, sk->w2[v] * inv255, sk->w3[v] * inv255 };
                            v3 pp[4], nn[4];
                            for (int k = 0; k < 4; k += 1) {
                                pp[k] = xf(p, &inst->posed_bone_matrices[j[k]]);
                                nn[k] = xf_dir(n, &inst->posed_bone_matrices[j[k]]);
                            }
                            pw.x = pfma(pp[0].x, w[0], pfma(pp[1].x, w[1], pfma(pp[2].x, w[2], pp[3].x * w[3])));
                            pw.y = pfma(pp[0].y, w[0], pfma(pp[1].y, w[1], pfma(pp[2].y, w[2], pp[3].y * w[3])));
                            pw.z = pfma(pp[0].z, w[0], pfma(pp[1].z, w[1], pfma(pp[2].z, w[2], pp[3].z * w[3])));
                            nw.x = pfma(nn[0].x, w[0], pfma(nn[1].x, w[1], pfma(nn[2].x, w[2], nn[3].x * w[3])));
                            nw.y = pfma(nn[0].y, w[0], pfma(nn[1].y, w[1], pfma(nn[2].y, w[2], nn[3].y * w[3])));
                            nw.z = pfma(nn[0].z, w[0], pfma(nn[1].z, w[1], pfma(nn[2].z, w[2], nn[3].z * w[3])));
                            pw = xf(pw, &inst->transform);
                            nw = xf_dir(nw, &inst->transform);
                        }
                        else {
                            pw = xf(p, &inst->transform);
                            nw = xf_dir(n, &inst->transform);
                        }
                        ssm[192 + v] = pw.x; ssm[256 + v] = pw.y; ssm[320 + v] = pw.z;
                        ssm[384 + v] = nw.x; ssm[448 + v] = nw.y; ssm[512 + v] = nw.z;
                        /* projection with view_proj, FMA order z, y, x (5094-5123) */
                        const f32 w = pfma(pw.x, vp->m[0][3], pfma(pw.y, vp->m[1][3], pfma(pw.z, vp->m[2][3], vp->m[3][3])));
                        const f32 tx = pfma(pw.x, vp->m[0][0], pfma(pw.y, vp->m[1][0], pfma(pw.z, vp->m[2][0], vp->m[3][0])));
                        const f32 ty = pfma(pw.x, vp->m[0][1], pfma(pw.y, vp->m[1][1], pfma(pw.z, vp->m[2][1], vp->m[3][1])));
                        const f32 tz = pfma(pw.x, vp->m[0][2], pfma(pw.y, vp->m[1][2], pfma(pw.z, vp->m[2][2], vp->m[3][2])));
                        const f32 iw = rcp_refined(w);
                        const f32 sx = pfma(tx, iw * sxs, sxb);
                        const f32 sy = pfma(ty, iw * sys, syb);
                        cs.tx[v] = tx; cs.ty[v] = ty; cs.tz[v] = tz; cs.w[v] = w;
                        cs.sx[v] = sx; cs.sy[v] = sy;
                        ssm[v] = sx; ssm[64 + v] = sy; ssm[128 + v] = iw;
                    }
                    for (u32 b = 0; b < 16; b += 1) {
                        setup_and_bin_batch(ps, b, (u32)ii, gm, &mesh->inds[m], &cs, at, ssm);
                    }
                }
            }
        }
    }
}

/* ------------------------------------------------------------ raster
   sft__rast_simd_batch, sofren2.c:5148-5499; rotated-grid sample offsets
   from the pixel centre, sofren2.c:186-196 */
static const f32 k_msaa_ox[3][4] = { { 0.f }, { -0.25f, 0.25f }, { -0.125f, 0.375f, -0.375f, 0.125f } };
static const f32 k_msaa_oy[3][4] = { { 0.f }, { -0.25f, 0.25f }, { -0.375f, -0.125f, 0.125f, 0.375f } };
static int msaa_row(int samples) { return 4 == samples ? 2 : (2 == samples ? 1 : 0); }

static void raster_payload(port_state* ps, u64 payload, f32* depth, u32* vid, i32 tmx, i32 tmy) {
    const sftx_frame_desc* f = ps->f;
    const int S = f->samples, aa = S > 1;
    const f32* ox = k_msaa_ox[msaa_row(S)];
    const f32* oy = k_msaa_oy[msaa_row(S)];
    u32 active = (u32)((payload >> 56) & 0xFF);
    f32 sx[3][8], sy[3][8], iw[3][8];
    u32 inst_ind = 0, m_ind = 0, batch = 0, slot = 0;
    int clipped = 0;
    if (0 == active) {
        clipped = 1;
        slot = (u32)payload;
        const clip_rec* ct = (const clip_rec*)(ps->out->clips + (u64)slot * sizeof(clip_rec));
        for (int k = 0; k < 3; k += 1) {
            for (int l = 0; l < 8; l += 1) { sx[k][l] = ct->sx[k]; sy[k][l] = ct->sy[k]; iw[k][l] = ct->inv_w[k]; }
        }
        active = 1;
    }
    else {
        inst_ind = (u32)(payload & 0xFFFFFF);
        m_ind = (u32)((payload >> 24) & 0xFFFFFF);
        batch = (u32)((payload >> 48) & 0xFF);
        const sft_instance* inst = &f->instances[inst_ind];
        const sft_mesh* mesh = &inst->model->meshes[inst->model->meshlet_mesh_inds[m_ind]];
        const struct sft_meshlet_inds* mi = &mesh->inds[m_ind - mesh->model_meshlet_offset];
        const f32* ssm = ss_at(ps, inst->ss_meshlet_offset + m_ind);
        for (int k = 0; k < 3; k += 1) {
            for (int l = 0; l < 8; l += 1) {
                const int v = mi->inds[batch * 32 + k * 8 + l];
                sx[k][l] = ssm[v]; sy[k][l] = ssm[64 + v]; iw[k][l] = ssm[128 + v];
            }
        }
    }
    for (int lane = 0; lane < 8; lane += 1) {
        if (!(active & (1u << lane))) continue;
        const f32 sx0 = sx[0][lane], sx1 = sx[1][lane], sx2 = sx[2][lane];
        const f32 sy0 = sy[0][lane], sy1 = sy[1][lane], sy2 = sy[2][lane];
        const f32 w0 = iw[0][lane], w1 = iw[1][lane], w2 = iw[2][lane];
        const f32 A0 = sy0 - sy1, B0 = sx1 - sx0, A1 = sy1 - sy2, B1 = sx2 - sx1, A2 = sy2 - sy0, B2 = sx0 - sx2;
        const f32 C0 = (pfma(sx0, sy1, -(sy0 * sx1)) - pfma(sy0, sx1, -(sx0 * sy1))) * 0.5f;
        const f32 C1 = (pfma(sx1, sy2, -(sy1 * sx2)) - pfma(sy1, sx2, -(sx1 * sy2))) * 0.5f;
        const f32 C2 = (pfma(sx2, sy0, -(sy2 * sx0)) - pfma(sy2, sx0, -(sx2 * sy0))) * 0.5f;
        const f32 dx1 = sx1 - sx0, dx2 = sx2 - sx0, dy1 = sy1 - sy0, dy2 = sy2 - sy0;
        const f32 det = dx1 * dy2 - dx2 * dy1;
        if (!(det > SFT_EPSILON)) continue;
        const f32 inv_det = rcp_refined(det);
        const f32 dzdx = ((w1 - w0) * dy2 - (w2 - w0) * dy1) * inv_det;
        const f32 dzdy = ((w2 - w0) * dx1 - (w1 - w0) * dx2) * inv_det;
        const f32 c0 = pfma(B0, 0.5f, pfma(A0, 0.5f, C0));
        const f32 c1 = pfma(B1, 0.5f, pfma(A1, 0.5f, C1));
        const f32 c2 = pfma(B2, 0.5f, pfma(A2, 0.5f, C2));
        const f32 zb = pfma(0.5f - sy0, dzdy, pfma(0.5f - sx0, dzdx, w0));
        const i32 min_x = cvtt(smax(smin(sx0, smin(sx1, sx2)), (f32)tmx));
        const i32 max_x = cvtt(smin(smax(sx0, smax(sx1, sx2)), (f32)(tmx + TS - 1)));
        const i32 min_y = cvtt(smax(smin(sy0, smin(sy1, sy2)), (f32)tmy));
        const i32 max_y = cvtt(smin(smax(sy0, smax(sy1, sy2)), (f32)(tmy + TS - 1)));
        if (min_x > max_x || min_y > max_y) continue;
        const u32 id = clipped ? (slot | (1u << 31))
            : ((inst_ind << (SFT_TRIANGLE_BITS + SFT_MESHLET_BITS)) | (m_ind << SFT_TRIANGLE_BITS) | (batch * 8 + (u32)lane));
        const i32 amin = min_x & ~7, amax = (max_x & ~7) + 8;
        const f32 se0 = c0 + A0 * (f32)amin + B0 * (f32)min_y;
        const f32 se1 = c1 + A1 * (f32)amin + B1 * (f32)min_y;
        const f32 se2 = c2 + A2 * (f32)amin + B2 * (f32)min_y;
        const f32 sz = zb + dzdx * (f32)amin + dzdy * (f32)min_y;
        const f32 bw = (f32)(amax - 1 - amin), bh = (f32)(max_y - min_y);
        const f32 me0 = se0 + (A0 < 0.f ? A0 * bw : 0.f) + (B0 < 0.f ? B0 * bh : 0.f);
        const f32 me1 = se1 + (A1 < 0.f ? A1 * bw : 0.f) + (B1 < 0.f ? B1 * bh : 0.f);
        const f32 me2 = se2 + (A2 < 0.f ? A2 * bw : 0.f) + (B2 < 0.f ? B2 * bh : 0.f);
        const f32 pad0 = aa ? 0.375f * (fabsf(A0) + fabsf(B0)) : 0.f;
        const f32 pad1 = aa ? 0.375f * (fabsf(A1) + fabsf(B1)) : 0.f;
        const f32 pad2 = aa ? 0.375f * (fabsf(A2) + fabsf(B2)) : 0.f;
        const int trivial = me0 >= pad0 && me1 >= pad1 && me2 >= pad2;
        f32 de0[4], de1[4], de2[4], dz[4];
        for (int q = 0; q < S; q += 1) {
            de0[q] = A0 * ox[q] + B0 * oy[q];
            de1[q] = A1 * ox[q] + B1 * oy[q];
            de2[q] = A2 * ox[q] + B2 * oy[q];
            dz[q] = dzdx * ox[q] + dzdy * oy[q];
        }
        f32 zr[8], e0r[8], e1r[8], e2r[8];
        for (int j = 0; j < 8; j += 1) {
            zr[j] = sz + dzdx * (f32)j;
            e0r[j] = se0 + A0 * (f32)j;
            e1r[j] = se1 + A1 * (f32)j;
            e2r[j] = se2 + A2 * (f32)j;
        }
        const f32 dzc = dzdx * 8.f, a0c = A0 * 8.f, a1c = A1 * 8.f, a2c = A2 * 8.f;
        for (i32 y = min_y; y <= max_y; y += 1) {
            f32 z[8], e0[8], e1[8], e2[8];
            memcpy(z, zr, sizeof(z)); memcpy(e0, e0r, sizeof(e0)); memcpy(e1, e1r, sizeof(e1)); memcpy(e2, e2r, sizeof(e2));
            i32 p = (y - tmy) * TS + (amin - tmx);
            for (i32 x = amin; x < amax; x += 8, p += 8) {
                if (!aa) {
                    for (int j = 0; j < 8; j += 1) {
                        const int cover = trivial || (e0[j] >= 0.f && e1[j] >= 0.f && e2[j] >= 0.f);
                        if (cover && z[j] > depth[p + j]) {
                            depth[p + j] = z[j];
                            vid[p + j] = id;
                        }
                    }
                }
                else {
                    /* edge path: a chunk no pixel centre comes within the sample
                       pad of is skipped (5414-5420) */
                    int might = trivial;
                    for (int j = 0; j < 8 && !might; j += 1) might = e0[j] >= -pad0 && e1[j] >= -pad1 && e2[j] >= -pad2;
                    for (int q = 0; q < S && might; q += 1) {
                        for (int j = 0; j < 8; j += 1) {
                            const int cover = trivial || (e0[j] + de0[q] >= 0.f && e1[j] + de1[q] >= 0.f && e2[j] + de2[q] >= 0.f);
                            const f32 zs = z[j] + dz[q];
                            if (cover && zs > depth[q * TP + p + j]) {
                                depth[q * TP + p + j] = zs;
                                vid[q * TP + p + j] = id;
                            }
                        }
                    }
                }
                for (int j = 0; j < 8; j += 1) {
                    z[j] = z[j] + dzc;
                    e0[j] = e0[j] + a0c; e1[j] = e1[j] + a1c; e2[j] = e2[j] + a2c;
                }
            }
            for (int j = 0; j < 8; j += 1) {
                zr[j] = zr[j] + dzdy;
                e0r[j] = e0r[j] + B0; e1r[j] = e1r[j] + B1; e2r[j] = e2r[j] + B2;
            }
        }
    }
}

/* ------------------------------------------------------------ mip chains
   sofren2.c:9229-9289: level sizes, 4x4-block swizzle, [1 2 2 1]^2 tent with
   edge clamp and truncating division per channel */
u32 sftx_port_build_texture(const u32* level0, i32 w, i32 h, u32* out) {
    i32 lw[14], lh[14], levels = 0;
    u32 offs[15];
    u32 total = 0;
    i32 cw = w, chh = h;
    while (levels < 14) {
        lw[levels] = cw;
        lh[levels] = chh;
        offs[levels] = total;
        total += (u32)(((cw + 3) & ~3) * ((chh + 3) & ~3));
        levels += 1;
        if (1 == cw && 1 == chh) break;
        cw = cw > 1 ? cw / 2 : 1;
        chh = chh > 1 ? chh / 2 : 1;
    }
    if (!out) return total;
    memset(out, 0, (size_t)total * 4);
    u32* cur = (u32*)malloc((size_t)w * h * 4);
    memcpy(cur, level0, (size_t)w * h * 4);
    for (i32 l = 0; l < levels; l += 1) {
        const i32 bw = ((lw[l] + 3) & ~3) / 4;
        for (i32 y = 0; y < lh[l]; y += 1) {
            for (i32 x = 0; x < lw[l]; x += 1) {
                out[offs[l] + (u32)(((y >> 2) * bw + (x >> 2)) * 16 + ((y & 3) << 2) + (x & 3))] = cur[y * lw[l] + x];
            }
        }
        if (l + 1 == levels) break;
        const i32 nw = lw[l + 1], nh = lh[l + 1];
        u32* next = (u32*)malloc((size_t)nw * nh * 4);
        for (i32 y = 0; y < nh; y += 1) {
            for (i32 x = 0; x < nw; x += 1) {
                u32 r = 0, g = 0, b = 0, a = 0, ws = 0;
                for (i32 ky = -1; ky <= 2; ky += 1) {
                    const i32 sy = SFT_CLAMP(y * 2 + ky, 0, lh[l] - 1);
                    const u32 wy = (-1 == ky || 2 == ky) ? 1u : 2u;
                    for (i32 kx = -1; kx <= 2; kx += 1) {
                        const i32 sx = SFT_CLAMP(x * 2 + kx, 0, lw[l] - 1);
                        const u32 wt = ((-1 == kx || 2 == kx) ? 1u : 2u) * wy;
                        const u32 px = cur[sy * lw[l] + sx];
                        r += ((px >> 16) & 255u) * wt;
                        g += ((px >> 8) & 255u) * wt;
                        b += (px & 255u) * wt;
                        a += ((px >> 24) & 255u) * wt;
                        ws += wt;
                    }
                }
                next[y * nw + x] = ((a / ws) << 24) | ((r / ws) << 16) | ((g / ws) << 8) | (b / ws);
            }
        }
        free(cur);
        cur = next;
    }
    free(cur);
    return total;
}

/* ------------------------------------------------------------ texturing
   sofren2.c:5501-5729 */
static const sftx_texture_desc* tex_of(const sftx_frame_desc* f, u32 h) {
    if (SFT_INVALID_HANDLE == h) return NULL;
    const u32 i = h & SFT_HANDLE_INDEX_MASK;
    if ((i32)i >= f->n_textures || !f->textures[i].texels) return NULL;
    return &f->textures[i];
}

static u32 texel_addr(const sftx_texture_desc* t, int lod, i32 x, i32 y) {
    const u32 ux = (u32)x, uy = (u32)y;
    if (t->is_pot) return (((uy >> 2) << t->stride_shift[lod]) + (ux >> 2)) * 16u + ((uy & 3u) << 2) + (ux & 3u);
    return (((uy >> 2) * (u32)t->blocks_w[lod]) + (ux >> 2)) * 16u + ((uy & 3u) << 2) + (ux & 3u);
}

static inline f32 ch(u32 c, int s) { return (f32)(i32)((c >> s) & 0xFFu) * (1.f / 255.f); }

static void sample_bilinear(const sftx_frame_desc* f, u32 h, int lod, f32 cu, f32 cv, f32* r, f32* g, f32* b) {
    const sftx_texture_desc* t = tex_of(f, h);
    if (SFT_INVALID_HANDLE == h || !t) { *r = *g = *b = 1.f; return; }
    const u32* px = t->texels + t->mip_offset[lod];
    const i32 cw = SFT_MAX(t->w >> lod, 1), chh = SFT_MAX(t->h >> lod, 1);
    const f32 upx = cu * (f32)cw - 0.5f, vpx = cv * (f32)chh - 0.5f;
    const f32 uf = floorf(upx), vf = floorf(vpx);
    const f32 ufr = upx - uf, vfr = vpx - vf;
    const i32 u0 = cvtt(uf), v0 = cvtt(vf);
    i32 u0w, v0w, u1, v1;
    if (t->is_pot) {
        u1 = (u0 + 1) & (cw - 1); v1 = (v0 + 1) & (chh - 1);
        u0w = u0 & (cw - 1); v0w = v0 & (chh - 1);
    }
    else {
        u1 = SFT_MIN(u0 + 1, cw - 1); v1 = SFT_MIN(v0 + 1, chh - 1);
        u0w = SFT_MIN(u0, cw - 1); v0w = SFT_MIN(v0, chh - 1);
    }
    const u32 c00 = px[texel_addr(t, lod, u0w, v0w)], c10 = px[texel_addr(t, lod, u1, v0w)];
    const u32 c01 = px[texel_addr(t, lod, u0w, v1)], c11 = px[texel_addr(t, lod, u1, v1)];
    f32 o[3];
    const int sh[3] = { 16, 8, 0 };
    for (int k = 0; k < 3; k += 1) {
        const f32 f00 = ch(c00, sh[k]), f10 = ch(c10, sh[k]), f01 = ch(c01, sh[k]), f11 = ch(c11, sh[k]);
        const f32 top = f00 + (f10 - f00) * ufr, bot = f01 + (f11 - f01) * ufr;
        o[k] = top + (bot - top) * vfr;
    }
    *r = o[0]; *g = o[1]; *b = o[2];
}

static void sample_nearest(const sftx_frame_desc* f, u32 h, int lod, f32 cu, f32 cv, f32* r, f32* g, f32* b) {
    const sftx_texture_desc* t = tex_of(f, h);
    if (SFT_INVALID_HANDLE == h || !t) { *r = *g = *b = 1.f; return; }
    const u32* px = t->texels + t->mip_offset[lod];
    const i32 cw = SFT_MAX(t->w >> lod, 1), chh = SFT_MAX(t->h >> lod, 1);
    const i32 u0 = cvtt(floorf(cu * (f32)cw)), v0 = cvtt(floorf(cv * (f32)chh));
    const i32 uw = t->is_pot ? (u0 & (cw - 1)) : SFT_MIN(u0, cw - 1);
    const i32 vw = t->is_pot ? (v0 & (chh - 1)) : SFT_MIN(v0, chh - 1);
    const u32 c = px[texel_addr(t, lod, uw, vw)];
    *r = ch(c, 16); *g = ch(c, 8); *b = ch(c, 0);
}

/* scalar LOD from the first lane's UV gradients; the Bayer dither added before
   truncation never changes an integer log2 (5501-5541) */
static int calc_lod(const sftx_frame_desc* f, u32 h, f32 dux, f32 duy, f32 dvx, f32 dvy, i32 bx, i32 by) {
    static const f32 bayer[4] = { -0.375f, 0.125f, 0.375f, -0.125f };
    const sftx_texture_desc* t = tex_of(f, h);
    if (SFT_INVALID_HANDLE == h || !t) return 0;
    const f32 ux = dux * t->w, vx = dvx * t->h, uy = duy * t->w, vy = dvy * t->h;
    const f32 rho = SFT_MAX(SFT_MAX(fabsf(ux), fabsf(vx)), SFT_MAX(fabsf(uy), fabsf(vy)));
    i32 lod = 0;
    if (rho > 1.f) {
        const i32 ir = cvtt(rho);
        const f32 lf = (f32)(31 - __builtin_clz((u32)ir));
        lod = cvtt(lf + 0.5f + bayer[((by & 1) << 1) | (bx & 1)]);
    }
    return SFT_CLAMP(lod, 0, t->mip_levels - 1);
}

/* ------------------------------------------------------------ resolve
   sft__resolve_tile, sofren2.c:5809-6753 (non-MSAA path) */
typedef struct {
    f32 sx[3], sy[3], iw[3], u[3], v[3], lu[3], lv[3];
    v3 wp[3], wn[3];
    u32 inst_ind, m_ind;
} tri_t;

static void fetch_tri(port_state* ps, u32 vid, tri_t* t) {
    const sftx_frame_desc* f = ps->f;
    if (vid & (1u << 31)) {
        const clip_rec* ct = (const clip_rec*)(ps->out->clips + (u64)(vid & 0x7FFFFFFFu) * sizeof(clip_rec));
        t->inst_ind = ct->inst_ind;
        t->m_ind = ct->m_ind;
        for (int k = 0; k < 3; k += 1) {
            t->sx[k] = ct->sx[k]; t->sy[k] = ct->sy[k]; t->iw[k] = ct->inv_w[k];
            t->u[k] = ct->u[k]; t->v[k] = ct->v[k]; t->lu[k] = ct->lu[k]; t->lv[k] = ct->lv[k];
            t->wp[k].x = ct->px[k]; t->wp[k].y = ct->py[k]; t->wp[k].z = ct->pz[k];
            t->wn[k].x = ct->nx[k]; t->wn[k].y = ct->ny[k]; t->wn[k].z = ct->nz[k];
        }
        return;
    }
    t->inst_ind = (vid >> (SFT_TRIANGLE_BITS + SFT_MESHLET_BITS)) & ((1u << SFT_INSTANCE_BITS) - 1);
    t->m_ind = (vid >> SFT_TRIANGLE_BITS) & ((1u << SFT_MESHLET_BITS) - 1);
    const u32 tri = vid & ((1u << SFT_TRIANGLE_BITS) - 1);
    const sft_instance* inst = &f->instances[t->inst_ind];
    const sft_mesh* mesh = &inst->model->meshes[inst->model->meshlet_mesh_inds[t->m_ind]];
    const u32 lm = t->m_ind - mesh->model_meshlet_offset;
    const struct sft_meshlet_attribs* at = &mesh->meshlet_attribs[lm];
    const f32* ssm = ss_at(ps, inst->ss_meshlet_offset + t->m_ind);
    for (int k = 0; k < 3; k += 1) {
        const int v = mesh->inds[lm].inds[(tri / 8) * 32 + k * 8 + (tri % 8)];
        t->sx[k] = ssm[v]; t->sy[k] = ssm[64 + v]; t->iw[k] = ssm[128 + v];
        t->u[k] = at->u[v]; t->v[k] = at->v[v]; t->lu[k] = at->lu[v]; t->lv[k] = at->lv[v];
        t->wp[k].x = ssm[192 + v]; t->wp[k].y = ssm[256 + v]; t->wp[k].z = ssm[320 + v];
        t->wn[k].x = ssm[384 + v]; t->wn[k].y = ssm[448 + v]; t->wn[k].z = ssm[512 + v];
    }
}

/* perspective-correct weights at a pixel centre (6273-6318) */
/* with MSAA the screen barycentrics are clamped to the triangle and
   renormalised, so pixel centres outside it never extrapolate (6296-6305) */
static void bary_ms(const tri_t* t, f32 px, f32 py, int aa, f32* p0, f32* p1, f32* p2) {
    const f32 dx1 = t->sx[1] - t->sx[0], dx2 = t->sx[2] - t->sx[0];
    const f32 dy1 = t->sy[1] - t->sy[0], dy2 = t->sy[2] - t->sy[0];
    const f32 inv_det = rcp_refined(dx1 * dy2 - dx2 * dy1);
    const f32 ex = px - t->sx[0], ey = py - t->sy[0];
    f32 bu = (ex * dy2 - ey * dx2) * inv_det;
    f32 bv = (ey * dx1 - ex * dy1) * inv_det;
    f32 bw = (1.f - bu) - bv;
    if (aa) {
        bu = smax(0.f, bu);
        bv = smax(0.f, bv);
        bw = smax(0.f, bw);
        const f32 inv_sum = rcp_refined(bu + (bv + bw));
        bu = bu * inv_sum;
        bv = bv * inv_sum;
        bw = bw * inv_sum;
    }
    const f32 z = rcp_refined(pfma(bw, t->iw[0], pfma(bu, t->iw[1], bv * t->iw[2])));
    *p0 = (bw * t->iw[0]) * z;
    *p1 = (bu * t->iw[1]) * z;
    *p2 = (bv * t->iw[2]) * z;
}

static void bary(const tri_t* t, f32 px, f32 py, f32* p0, f32* p1, f32* p2) { bary_ms(t, px, py, 0, p0, p1, p2); }

static inline f32 interp3(f32 p0, f32 p1, f32 p2, const f32* a) { return pfma(p0, a[0], pfma(p1, a[1], p2 * a[2])); }

/* per-tile light list (5855-5934) */
static int tile_lights(const sftx_frame_desc* f, f32 zmin, f32 zmax, i32 tmx, i32 tmy, u16 out[16]) {
    int n = 0;
    if (!(zmin <= zmax && f->n_point > 0)) return 0;
    const sftmat* vp = &f->view_proj;
    const f32 hw = f->half_width, hh = f->half_height;
    for (i32 l = 0; l < f->n_point; l += 1) {
        const sftvec pos = f->point_lights[l].pos;
        const f32 radius = f->point_lights[l].radius;
        const f32 w = pos.x * vp->m[0][3] + pos.y * vp->m[1][3] + pos.z * vp->m[2][3] + vp->m[3][3];
        const f32 lmin = w - radius, lmax = w + radius;
        if (lmax <= SFT_EPSILON) continue;
        const f32 safe = (lmin > SFT_EPSILON) ? lmin : SFT_EPSILON;
        const f32 iwmin = 1.f / lmax, iwmax = 1.f / safe;
        if (iwmax < zmin || iwmin > zmax) continue;
        if (w > radius) {
            const f32 tx = pos.x * vp->m[0][0] + pos.y * vp->m[1][0] + pos.z * vp->m[2][0] + vp->m[3][0];
            const f32 ty = pos.x * vp->m[0][1] + pos.y * vp->m[1][1] + pos.z * vp->m[2][1] + vp->m[3][1];
            const f32 sx = (tx / w + 1.f) * hw + f->vp_x;
            const f32 sy = (-ty / w + 1.f) * hh + f->vp_y;
            const f32 psx = sqrtf(vp->m[0][0] * vp->m[0][0] + vp->m[1][0] * vp->m[1][0] + vp->m[2][0] * vp->m[2][0]);
            const f32 psy = sqrtf(vp->m[0][1] * vp->m[0][1] + vp->m[1][1] * vp->m[1][1] + vp->m[2][1] * vp->m[2][1]);
            const f32 sw = SFT_MAX(w - radius, SFT_EPSILON);
            const f32 rx = ((radius * psx) + fabsf(tx) * (radius / w)) / sw * hw;
            const f32 ry = ((radius * psy) + fabsf(ty) * (radius / w)) / sw * hh;
            if (sx + rx < (f32)tmx || sx - rx > (f32)(tmx + TS) || sy + ry < (f32)tmy || sy - ry > (f32)(tmy + TS)) continue;
        }
        out[n++] = (u16)l;
        if (16 == n) break;
    }
    return n;
}

/* ------------------------------------------------------------ TLAS
   sft__tlas_build and the 8-wide builder, sofren2.c:2885-3062, 3619-3703 */
typedef struct { u32 ind; struct { sftvec min, max; } aabb; sftvec centroid; } build_item;

static f32 axis_val(const build_item* it, int axis) {
    return 0 == axis ? it->centroid.x : (1 == axis ? it->centroid.y : it->centroid.z);
}

/* Hoare-partition quicksort of sofren2.c:2895-2924 (same pivot and swaps) */
static void sort_items(build_item* items, i32 count, int axis) {
    if (count <= 1) return;
    const f32 pivot = axis_val(&items[count / 2], axis);
    i32 i = 0, j = count - 1;
    while (i <= j) {
        while (axis_val(&items[i], axis) < pivot) i += 1;
        while (axis_val(&items[j], axis) > pivot) j -= 1;
        if (i <= j) {
            const build_item t = items[i]; items[i] = items[j]; items[j] = t;
            i += 1;
            j -= 1;
        }
    }
    if (j > 0) sort_items(items, j + 1, axis);
    if (i < count) sort_items(items + i, count - i, axis);
}

static void partition_to_8(build_item* items, u32 start, u32 count, u32 depth, u32* starts, u32* counts, u32* groups) {
    if (0 == count) return;
    if (3 == depth || 1 == count) {
        starts[*groups] = start;
        counts[*groups] = count;
        *groups += 1;
        return;
    }
    f32 mn[3] = { 1e30f, 1e30f, 1e30f }, mx[3] = { -1e30f, -1e30f, -1e30f };
    for (u32 i = 0; i < count; i += 1) {
        const f32 c[3] = { items[start + i].centroid.x, items[start + i].centroid.y, items[start + i].centroid.z };
        for (int k = 0; k < 3; k += 1) {
            if (c[k] < mn[k]) mn[k] = c[k];
        }
        for (int k = 0; k < 3; k += 1) {
            if (c[k] > mx[k]) mx[k] = c[k];
        }
    }
    const f32 dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
    const int axis = (dy > dx && dy > dz) ? 1 : ((dz > dx && dz > dy) ? 2 : 0);
    sort_items(items + start, (i32)count, axis);
    const u32 mid = count / 2;
    partition_to_8(items, start, mid, depth + 1, starts, counts, groups);
    partition_to_8(items, start + mid, count - mid, depth + 1, starts, counts, groups);
}

static u32 bvh_build(build_item* items, u32 count, struct sft_bvh_node* nodes, u32* node_count, f32 out[6]) {
    const u32 ni = (*node_count)++;
    struct sft_bvh_node* n = &nodes[ni];
    for (int i = 0; i < 8; i += 1) {
        n->min_x[i] = 1e30f; n->min_y[i] = 1e30f; n->min_z[i] = 1e30f;
        n->max_x[i] = -1e30f; n->max_y[i] = -1e30f; n->max_z[i] = -1e30f;
        n->child_indices[i] = 0xFFFFFFFFu;
    }
    f32 b[6] = { 1e30f, 1e30f, 1e30f, -1e30f, -1e30f, -1e30f };
    if (count <= 8) {
        n->is_leaf_group = 1;
        n->child_count = (u8)count;
        for (u32 i = 0; i < count; i += 1) {
            const build_item* it = &items[i];
            n->min_x[i] = it->aabb.min.x; n->min_y[i] = it->aabb.min.y; n->min_z[i] = it->aabb.min.z;
            n->max_x[i] = it->aabb.max.x; n->max_y[i] = it->aabb.max.y; n->max_z[i] = it->aabb.max.z;
            n->child_indices[i] = 0x80000000u | it->ind;
            if (it->aabb.min.x < b[0]) b[0] = it->aabb.min.x;
            if (it->aabb.min.y < b[1]) b[1] = it->aabb.min.y;
            if (it->aabb.min.z < b[2]) b[2] = it->aabb.min.z;
            if (it->aabb.max.x > b[3]) b[3] = it->aabb.max.x;
            if (it->aabb.max.y > b[4]) b[4] = it->aabb.max.y;
            if (it->aabb.max.z > b[5]) b[5] = it->aabb.max.z;
        }
    }
    else {
        n->is_leaf_group = 0;
        u32 starts[8] = {0}, counts[8] = {0}, groups = 0;
        partition_to_8(items, 0, count, 0, starts, counts, &groups);
        n->child_count = (u8)groups;
        for (u32 c = 0; c < groups; c += 1) {
            f32 cb[6];
            const u32 child = bvh_build(&items[starts[c]], counts[c], nodes, node_count, cb);
            n = &nodes[ni];
            n->child_indices[c] = child;
            n->min_x[c] = cb[0]; n->min_y[c] = cb[1]; n->min_z[c] = cb[2];
            n->max_x[c] = cb[3]; n->max_y[c] = cb[4]; n->max_z[c] = cb[5];
            if (cb[0] < b[0]) b[0] = cb[0];
            if (cb[1] < b[1]) b[1] = cb[1];
            if (cb[2] < b[2]) b[2] = cb[2];
            if (cb[3] > b[3]) b[3] = cb[3];
            if (cb[4] > b[4]) b[4] = cb[4];
            if (cb[5] > b[5]) b[5] = cb[5];
        }
    }
    if (out) memcpy(out, b, sizeof(b));
    return ni;
}

/* world AABB of each BLAS root under the instance transform; items are indexed
   by instance while only BLAS-carrying instances are counted (sofren2.c:3630-3702:
   a BLAS-less instance before a BLAS one leaves a hole the reference fills with
   scratch memory; the workloads draw BLAS instances first) */
static void tlas_build(port_state* ps) {
    const sftx_frame_desc* f = ps->f;
    const u32 n = (u32)f->n_instances;
    ps->tlas_count = 0;
    if (0 == n) return;
    build_item* items = (build_item*)calloc(n, sizeof(build_item));
    u32 valid = 0;
    for (u32 i = 0; i < n; i += 1) {
        const sft_instance* inst = &f->instances[i];
        const sft_model* m = inst->model;
        if (0 == m->blas_node_count) continue;
        const struct sft_bvh_node* r = &m->blas_nodes[m->blas_root_ind];
        f32 mnx = 1e30f, mny = 1e30f, mnz = 1e30f, mxx = -1e30f, mxy = -1e30f, mxz = -1e30f;
        for (u32 c = 0; c < r->child_count; c += 1) {
            if (r->min_x[c] < mnx) mnx = r->min_x[c];
            if (r->min_y[c] < mny) mny = r->min_y[c];
            if (r->min_z[c] < mnz) mnz = r->min_z[c];
            if (r->max_x[c] > mxx) mxx = r->max_x[c];
            if (r->max_y[c] > mxy) mxy = r->max_y[c];
            if (r->max_z[c] > mxz) mxz = r->max_z[c];
        }
        const sftvec corners[8] = { {mnx, mny, mnz}, {mxx, mny, mnz}, {mnx, mxy, mnz}, {mxx, mxy, mnz},
                                    {mnx, mny, mxz}, {mxx, mny, mxz}, {mnx, mxy, mxz}, {mxx, mxy, mxz} };
        sftvec wmn = { 1e30f, 1e30f, 1e30f }, wmx = { -1e30f, -1e30f, -1e30f };
        const sftmat* T = &inst->transform;
        for (int k = 0; k < 8; k += 1) {
            const f32 tx = corners[k].x * T->m[0][0] + corners[k].y * T->m[1][0] + corners[k].z * T->m[2][0] + T->m[3][0];
            const f32 ty = corners[k].x * T->m[0][1] + corners[k].y * T->m[1][1] + corners[k].z * T->m[2][1] + T->m[3][1];
            const f32 tz = corners[k].x * T->m[0][2] + corners[k].y * T->m[1][2] + corners[k].z * T->m[2][2] + T->m[3][2];
            if (tx < wmn.x) wmn.x = tx;
            if (ty < wmn.y) wmn.y = ty;
            if (tz < wmn.z) wmn.z = tz;
            if (tx > wmx.x) wmx.x = tx;
            if (ty > wmx.y) wmx.y = ty;
            if (tz > wmx.z) wmx.z = tz;
        }
        items[i].ind = i;
        items[i].aabb.min = wmn;
        items[i].aabb.max = wmx;
        items[i].centroid = sft_v_mul(sft_v_add(wmn, wmx), 0.5f);
        valid += 1;
    }
    ps->tlas = (struct sft_bvh_node*)calloc(valid * 2 + 1, sizeof(struct sft_bvh_node));
    ps->tlas_root = bvh_build(items, valid, ps->tlas, &ps->tlas_count, NULL);
    free(items);
}

/* ------------------------------------------------------------ occlusion rays
   sofren2.c:3840-4243 (packet any-hit, frustum vs node, packet frustum,
   TLAS/BLAS traversal) and the pass at 8163-8280 */
typedef struct { f32 n[6][3], an[6][3], d[6]; int valid[6]; } pfrustum;

/* 8 child boxes vs 6 planes: centre/extent form, FMA order of 3933-3938 */
static int frustum_vs_node(const struct sft_bvh_node* nd, const pfrustum* fr) {
    int mask = 0xFF;
    f32 cx[8], cy[8], cz[8], ex[8], ey[8], ez[8];
    for (int i = 0; i < 8; i += 1) {
        cx[i] = (nd->max_x[i] + nd->min_x[i]) * 0.5f; ex[i] = (nd->max_x[i] - nd->min_x[i]) * 0.5f;
        cy[i] = (nd->max_y[i] + nd->min_y[i]) * 0.5f; ey[i] = (nd->max_y[i] - nd->min_y[i]) * 0.5f;
        cz[i] = (nd->max_z[i] + nd->min_z[i]) * 0.5f; ez[i] = (nd->max_z[i] - nd->min_z[i]) * 0.5f;
    }
    for (int p = 0; p < 6; p += 1) {
        for (int i = 0; i < 8; i += 1) {
            const f32 r = pfma(fr->an[p][0], ex[i], pfma(fr->an[p][1], ey[i], fr->an[p][2] * ez[i]));
            const f32 dc = pfma(fr->n[p][0], cx[i], pfma(fr->n[p][1], cy[i], fr->n[p][2] * cz[i]));
            if (!((dc + fr->d[p]) + r >= 0.f)) mask &= ~(1 << i);
        }
        if (0 == mask) return 0;
    }
    return mask;
}

static void build_packet_frustum(port_state* ps, const f32 ox[8], const f32 oy[8], const f32 oz[8], int active,
                                 sftvec lp, f32 lrad, pfrustum* fr) {
    f32 mnx = 1e9f, mxx = -1e9f, mny = 1e9f, mxy = -1e9f, mnz = 1e9f, mxz = -1e9f;
    sftvec c = { 0.f, 0.f, 0.f };
    int n = 0;
    for (int i = 0; i < 8; i += 1) {
        if (!(active & (1 << i))) continue;
        if (ox[i] < mnx) mnx = ox[i];
        if (ox[i] > mxx) mxx = ox[i];
        if (oy[i] < mny) mny = oy[i];
        if (oy[i] > mxy) mxy = oy[i];
        if (oz[i] < mnz) mnz = oz[i];
        if (oz[i] > mxz) mxz = oz[i];
        c.x += ox[i]; c.y += oy[i]; c.z += oz[i];
        n += 1;
    }
    if (0 == n) return;
    c.x /= (f32)n; c.y /= (f32)n; c.z /= (f32)n;
    const sftvec corners[8] = { {mnx, mny, mnz}, {mxx, mny, mnz}, {mnx, mxy, mnz}, {mxx, mxy, mnz},
                                {mnx, mny, mxz}, {mxx, mny, mxz}, {mnx, mxy, mxz}, {mxx, mxy, mxz} };
    const sftvec fwd = sft_v_norm(sft_v_sub(c, lp));
    const sftvec up = (fabsf(fwd.y) > 0.99f) ? SFT_V(1, 0, 0) : SFT_V(0, 1, 0);
    const sftmat lv = sft_m_view_look_at(lp, c, up);
    f32 mnxz = 1e9f, mxxz = -1e9f, mnyz = 1e9f, mxyz = -1e9f;
    int invalid = 0;
    for (int i = 0; i < 8; i += 1) {
        const sftquat q = { corners[i].x, corners[i].y, corners[i].z, 1.f };
        const sftvec cl = sft_m_mul_quat(lv, q);
        if (cl.z < SFT_EPSILON) { invalid = 1; break; }
        if (cl.x / cl.z < mnxz) mnxz = cl.x / cl.z;
        if (cl.x / cl.z > mxxz) mxxz = cl.x / cl.z;
        if (cl.y / cl.z < mnyz) mnyz = cl.y / cl.z;
        if (cl.y / cl.z > mxyz) mxyz = cl.y / cl.z;
    }
    const sftvec lr = SFT_V(lv.m[0][0], lv.m[1][0], lv.m[2][0]);
    const sftvec lu = SFT_V(lv.m[0][1], lv.m[1][1], lv.m[2][1]);
    const sftvec lf = SFT_V(lv.m[0][2], lv.m[1][2], lv.m[2][2]);
    if (invalid) {
        /* the reference sets only normal.x here and leaves the rest of the side
           planes uninitialised; such packets are counted and the workload avoids them */
        ps->invalid_packets += 1;
        for (int p = 0; p < 4; p += 1) {
            fr->n[p][0] = 0.f; fr->n[p][1] = 0.f; fr->n[p][2] = 0.f;
            fr->an[p][0] = 0.f; fr->an[p][1] = 0.f; fr->an[p][2] = 0.f;
            fr->d[p] = 0.f;
        }
    }
    else {
        const sftvec nl[4] = { sft_v_norm(SFT_V(1.f, 0.f, -mnxz)), sft_v_norm(SFT_V(-1.f, 0.f, mxxz)),
                               sft_v_norm(SFT_V(0.f, 1.f, -mnyz)), sft_v_norm(SFT_V(0.f, -1.f, mxyz)) };
        for (int p = 0; p < 4; p += 1) {
            const sftvec nw = sft_v_add(sft_v_mul(lr, nl[p].x), sft_v_add(sft_v_mul(lu, nl[p].y), sft_v_mul(lf, nl[p].z)));
            fr->n[p][0] = nw.x; fr->n[p][1] = nw.y; fr->n[p][2] = nw.z;
            fr->d[p] = -sft_v_dot(nw, lp) + SFT_SHADOW_LIGHT_SIZE;
            fr->an[p][0] = fabsf(nw.x); fr->an[p][1] = fabsf(nw.y); fr->an[p][2] = fabsf(nw.z);
        }
    }
    fr->n[4][0] = lf.x; fr->n[4][1] = lf.y; fr->n[4][2] = lf.z;
    fr->d[4] = -sft_v_dot(lf, lp) + SFT_SHADOW_LIGHT_SIZE;
    fr->an[4][0] = fabsf(lf.x); fr->an[4][1] = fabsf(lf.y); fr->an[4][2] = fabsf(lf.z);
    fr->n[5][0] = -lf.x; fr->n[5][1] = -lf.y; fr->n[5][2] = -lf.z;
    fr->d[5] = -sft_v_dot(SFT_V(-lf.x, -lf.y, -lf.z), lp) + lrad;
    fr->an[5][0] = fabsf(lf.x); fr->an[5][1] = fabsf(lf.y); fr->an[5][2] = fabsf(lf.z);
}

/* two-sided Moller-Trumbore any-hit of 8 rays vs one triangle (3840-3904) */
static int packet_any_hit(sftvec v0, sftvec v1, sftvec v2, const v3* o, const v3* dir, const f32* tmax, int active) {
    const sftvec e1s = sft_v_sub(v1, v0), e2s = sft_v_sub(v2, v0);
    const v3 e1 = { e1s.x, e1s.y, e1s.z }, e2 = { e2s.x, e2s.y, e2s.z }, vv0 = { v0.x, v0.y, v0.z };
    int hit = 0;
    for (int l = 0; l < 8; l += 1) {
        if (!(active & (1 << l))) continue;
        const v3 d = dir[l];
        const v3 h = { d.y * e2.z - d.z * e2.y, d.z * e2.x - d.x * e2.z, d.x * e2.y - d.y * e2.x };
        const f32 a = vdot(e1, h);
        if (!(fabsf(a) > SFT_EPSILON)) continue;
        const f32 fi = rcp_refined(a);
        const v3 sv = vsub(o[l], vv0);
        const f32 u = fi * vdot(sv, h);
        if (!(u >= 0.f && u <= 1.f)) continue;
        const v3 q = { sv.y * e1.z - sv.z * e1.y, sv.z * e1.x - sv.x * e1.z, sv.x * e1.y - sv.y * e1.x };
        const f32 v = fi * vdot(d, q);
        if (!(v >= 0.f && u + v <= 1.f)) continue;
        const f32 t = fi * vdot(e2, q);
        if (t > SFT_EPSILON && t < tmax[l]) hit |= 1 << l;
    }
    return hit;
}

/* rays from the 8 origins to origin + L; returns the lanes that hit a static
   BLAS triangle inside the packet frustum (4086-4243) */
static int trace_packet(port_state* ps, const f32 ox[8], const f32 oy[8], const f32 oz[8],
                        const f32 lx[8], const f32 ly[8], const f32 lz[8], int active, sftvec lp, f32 lrad) {
    const sftx_frame_desc* f = ps->f;
    if (!active || 0 == ps->tlas_count) return 0;
    pfrustum fr;
    build_packet_frustum(ps, ox, oy, oz, active, lp, lrad, &fr);
    int sat = 0;
    u32 tstack[64];
    i32 ttop = 0;
    tstack[ttop++] = ps->tlas_root;
    while (ttop > 0) {
        const struct sft_bvh_node* tn = &ps->tlas[tstack[--ttop]];
        int tm = frustum_vs_node(tn, &fr);
        while (tm) {
            const int tl = __builtin_ctz(tm);
            tm &= tm - 1;
            if (0xFFFFFFFFu == tn->child_indices[tl]) continue;
            if (!tn->is_leaf_group) { tstack[ttop++] = tn->child_indices[tl]; continue; }
            const sft_instance* inst = &f->instances[tn->child_indices[tl] & ~0x80000000u];
            const sft_model* model = inst->model;
            if (model->skeleton.bone_count > 0 || 0 == model->blas_node_count) continue;
            pfrustum mf;
            const sftmat* M = &inst->transform;
            for (int p = 0; p < 6; p += 1) {
                const f32 nx = fr.n[p][0], ny = fr.n[p][1], nz = fr.n[p][2], d = fr.d[p];
                const f32 nmx = nx * M->m[0][0] + ny * M->m[0][1] + nz * M->m[0][2];
                const f32 nmy = nx * M->m[1][0] + ny * M->m[1][1] + nz * M->m[1][2];
                const f32 nmz = nx * M->m[2][0] + ny * M->m[2][1] + nz * M->m[2][2];
                const f32 dm = (nx * M->m[3][0] + ny * M->m[3][1] + nz * M->m[3][2]) + d;
                const f32 il = 1.f / sqrtf(nmx * nmx + nmy * nmy + nmz * nmz);
                mf.n[p][0] = nmx * il; mf.n[p][1] = nmy * il; mf.n[p][2] = nmz * il;
                mf.d[p] = dm * il;
                mf.an[p][0] = fabsf(nmx * il); mf.an[p][1] = fabsf(nmy * il); mf.an[p][2] = fabsf(nmz * il);
            }
            const sftmat im = sft_m_qinv(inst->transform);
            v3 lo[8], ldir[8];
            f32 ltmax[8];
            for (int l = 0; l < 8; l += 1) {
                const v3 o = { ox[l], oy[l], oz[l] };
                const v3 e = { ox[l] + lx[l], oy[l] + ly[l], oz[l] + lz[l] };
                lo[l] = xf(o, &im);
                const v3 le = xf(e, &im);
                const v3 L = vsub(le, lo[l]);
                ltmax[l] = sqrtf(vdot(L, L));
                ldir[l] = norm_fast(L);
            }
            u32 bstack[64];
            i32 btop = 0;
            bstack[btop++] = model->blas_root_ind;
            while (btop > 0) {
                const struct sft_bvh_node* bn = &model->blas_nodes[bstack[--btop]];
                int bm = frustum_vs_node(bn, &mf);
                while (bm) {
                    const int bl = __builtin_ctz(bm);
                    bm &= bm - 1;
                    if (0xFFFFFFFFu == bn->child_indices[bl]) continue;
                    if (!bn->is_leaf_group) { bstack[btop++] = bn->child_indices[bl]; continue; }
                    const u32 pl = bn->child_indices[bl] & ~0x80000000u;
                    const u32 m_ind = pl >> 7, tri = pl & 0x7F;
                    const sft_mesh* mesh = &model->meshes[model->meshlet_mesh_inds[m_ind]];
                    const u32 lm = m_ind - mesh->model_meshlet_offset;
                    const struct sft_meshlet* mv = &mesh->meshlets[lm];
                    const struct sft_meshlet_inds* mi = &mesh->inds[lm];
                    const u32 off = (tri / 8) * 32 + (tri % 8);
                    const u8 i0 = mi->inds[off], i1 = mi->inds[off + 8], i2 = mi->inds[off + 16];
                    const sftvec a0 = SFT_V(mv->x[i0], mv->y[i0], mv->z[i0]);
                    const sftvec a1 = SFT_V(mv->x[i1], mv->y[i1], mv->z[i1]);
                    const sftvec a2 = SFT_V(mv->x[i2], mv->y[i2], mv->z[i2]);
                    sat |= packet_any_hit(a0, a1, a2, lo, ldir, ltmax, active & ~sat);
                    if (sat == active) return sat;
                }
            }
        }
    }
    return sat;
}

static void occlusion_pass(port_state* ps) {
    const sftx_frame_desc* f = ps->f;
    const i32 pc = f->n_proxies;
    ps->proxies = (world_proxy*)calloc((size_t)(pc > 0 ? pc : 1), sizeof(world_proxy));
    for (i32 p = 0; p < pc; p += 1) {
        ps->proxies[p].posa = f->proxies[p].posa;
        ps->proxies[p].posb = f->proxies[p].posb;
        ps->proxies[p].radius = f->proxies[p].radius;
    }
    if (!(f->shadows_enabled && f->n_point > 0 && pc > 0)) return;
    const i32 lc = SFT_MIN(f->n_point, 64);
    for (i32 p = 0; p < pc; p += 1) {
        for (i32 l = 0; l < lc; l += 1) ps->proxies[p].light_vis[l] = 3;
    }
    for (i32 l = 0; l < lc; l += 1) {
        const sft_point_light* L = &f->point_lights[l];
        for (i32 base = 0; base < pc; base += 8) {
            sftvec bc = { 0.f, 0.f, 0.f };
            i32 act = 0;
            for (i32 k = 0; k < 8; k += 1) {
                if (base + k < pc) {
                    const world_proxy* px = &ps->proxies[base + k];
                    bc.x += (px->posa.x + px->posb.x) * 0.5f;
                    bc.y += (px->posa.y + px->posb.y) * 0.5f;
                    bc.z += (px->posa.z + px->posb.z) * 0.5f;
                    act += 1;
                }
            }
            if (act > 0) bc = sft_v_mul(bc, 1.f / (f32)act);
            const sftvec ldir = sft_v_norm(sft_v_sub(L->pos, bc));
            const sftvec lright = sft_v_norm(sft_v_perpendicular(ldir));
            const sftvec lup = sft_v_cross(lright, ldir);
            const f32 r = SFT_SHADOW_LIGHT_SIZE;
            const sftvec targets[3] = {
                sft_v_add(L->pos, sft_v_mul(lup, r)),
                sft_v_add(L->pos, sft_v_add(sft_v_mul(lright, r * 0.866025f), sft_v_mul(lup, -r * 0.5f))),
                sft_v_add(L->pos, sft_v_add(sft_v_mul(lright, -r * 0.866025f), sft_v_mul(lup, -r * 0.5f))) };
            for (int ray = 0; ray < 3; ray += 1) {
                f32 ox[8], oy[8], oz[8], lx[8], ly[8], lz[8];
                int am = 0;
                for (i32 k = 0; k < 8; k += 1) {
                    if (base + k < pc) {
                        const world_proxy* px = &ps->proxies[base + k];
                        f32 x, y, z;
                        if (0 == ray) { x = px->posa.x; y = px->posa.y; z = px->posa.z; }
                        else if (1 == ray) { x = px->posb.x; y = px->posb.y; z = px->posb.z; }
                        else { x = (px->posa.x + px->posb.x) * 0.5f; y = (px->posa.y + px->posb.y) * 0.5f; z = (px->posa.z + px->posb.z) * 0.5f; }
                        ox[k] = x; oy[k] = y; oz[k] = z;
                        lx[k] = targets[ray].x - x; ly[k] = targets[ray].y - y; lz[k] = targets[ray].z - z;
                        am |= 1 << k;
                    }
                    else {
                        ox[k] = oy[k] = oz[k] = 0.f;
                        lx[k] = ly[k] = lz[k] = 0.f;
                    }
                }
                const int occ = trace_packet(ps, ox, oy, oz, lx, ly, lz, am, L->pos, L->radius);
                for (i32 k = 0; k < 8; k += 1) {
                    if (occ & (1 << k)) ps->proxies[base + k].light_vis[l] -= 1;
                }
            }
        }
    }
}

/* ------------------------------------------------------------ shading
   sofren2.c:6190-6620; the whole 4x2 block is evaluated lane by lane for one
   triangle, as the reference's 8-wide code does, because the capsule-shadow
   blend couples the lanes of the block */
typedef struct {
    v3 pa, ab, W;
    f32 abdab, wdw, rb2, pr2, vis_mult;
} proxy_cache;

typedef struct {
    int n_lights;
    u16 lights[16];
    u8 n_proxies[16];
    proxy_cache cache[16][16];
} tile_shading;

/* first 16 proxies per tile light with visibility left, in index order, that
   reach the light's sphere (5936-5990) */
static void build_proxy_cache(port_state* ps, tile_shading* ts) {
    const sftx_frame_desc* f = ps->f;
    memset(ts->n_proxies, 0, sizeof(ts->n_proxies));
    if (!(f->shadows_enabled && f->n_proxies > 0 && ts->n_lights > 0)) return;
    for (int i = 0; i < ts->n_lights; i += 1) {
        const sft_point_light* l = &f->point_lights[ts->lights[i]];
        u8 pcnt = 0;
        for (i32 p = 0; p < f->n_proxies && pcnt < 16; p += 1) {
            const world_proxy* px = &ps->proxies[p];
            const u8 vis = px->light_vis[ts->lights[i]];
            if (0 == vis) continue;
            const sftvec cc = SFT_V((px->posa.x + px->posb.x) * 0.5f, (px->posa.y + px->posb.y) * 0.5f, (px->posa.z + px->posb.z) * 0.5f);
            const sftvec ab = sft_v_sub(px->posb, px->posa);
            const f32 len = sft_v_length(ab);
            const f32 br = px->radius + (len * 0.5f);
            const f32 cr = br + l->radius;
            const sftvec W = sft_v_sub(cc, l->pos);
            const f32 d2 = sft_v_length2(W);
            if (d2 <= cr * cr) {
                proxy_cache* c = &ts->cache[i][pcnt];
                c->pa.x = px->posa.x; c->pa.y = px->posa.y; c->pa.z = px->posa.z;
                c->ab.x = ab.x; c->ab.y = ab.y; c->ab.z = ab.z;
                c->W.x = W.x; c->W.y = W.y; c->W.z = W.z;
                c->abdab = sft_v_dot(ab, ab);
                c->wdw = d2;
                c->rb2 = br * br;
                c->pr2 = px->radius * px->radius;
                c->vis_mult = (f32)vis * 0.3333333f;
                pcnt += 1;
            }
        }
        ts->n_proxies[i] = pcnt;
    }
}

static const i32 k_ox[8] = { 0, 1, 2, 3, 0, 1, 2, 3 }, k_oy[8] = { 0, 0, 0, 0, 1, 1, 1, 1 };

/* shades one (block, triangle) group; out[c][lane] is the colour clamped to
   [0, 1] for every lane (only the lanes in `lanes` are meaningful) */
static void shade_group(port_state* ps, const tri_t* t, int lanes, i32 x, i32 y, const tile_shading* ts, f32 out[3][8]) {
    const sftx_frame_desc* f = ps->f;
    f32 cu[8], cv[8], clu[8], clv[8];
    v3 pos[8], N[8], V[8];
    const f32 ax[3] = { t->wp[0].x, t->wp[1].x, t->wp[2].x }, ay[3] = { t->wp[0].y, t->wp[1].y, t->wp[2].y }, az[3] = { t->wp[0].z, t->wp[1].z, t->wp[2].z };
    const f32 nx[3] = { t->wn[0].x, t->wn[1].x, t->wn[2].x }, ny[3] = { t->wn[0].y, t->wn[1].y, t->wn[2].y }, nz[3] = { t->wn[0].z, t->wn[1].z, t->wn[2].z };
    const v3 cam = { f->camera_pos.x, f->camera_pos.y, f->camera_pos.z };
    for (int l = 0; l < 8; l += 1) {
        f32 p0, p1, p2;
        bary_ms(t, ((f32)x + 0.5f) + (f32)k_ox[l], ((f32)y + 0.5f) + (f32)k_oy[l], f->samples > 1, &p0, &p1, &p2);
        cu[l] = interp3(p0, p1, p2, t->u); cv[l] = interp3(p0, p1, p2, t->v);
        clu[l] = interp3(p0, p1, p2, t->lu); clv[l] = interp3(p0, p1, p2, t->lv);
        pos[l].x = interp3(p0, p1, p2, ax); pos[l].y = interp3(p0, p1, p2, ay); pos[l].z = interp3(p0, p1, p2, az);
        const v3 nr = { interp3(p0, p1, p2, nx), interp3(p0, p1, p2, ny), interp3(p0, p1, p2, nz) };
        N[l] = norm_fast(nr);
        V[l] = norm_fast(vsub(cam, pos[l]));
    }
    const sft_instance* inst = &f->instances[t->inst_ind];
    const sft_mesh* mesh = &inst->model->meshes[inst->model->meshlet_mesh_inds[t->m_ind]];
    const sft_mat_handle mh = mesh->material;
    const u32 lm_page = mesh->lightmap_page;
    u32 base = 0xFFFFFFFFu, tex = SFT_INVALID_HANDLE;
    f32 shin = 4.f;
    if (SFT_INVALID_HANDLE != mh) {
        const u32 mi = SFT_MAT_INDEX(mh);
        if (SFT_MAT_IS_TRANSIENT(mh)) { base = f->tmat_color[mi]; tex = f->tmat_tex[mi]; shin = f->tmat_shininess[mi]; }
        else { base = f->mat_color[mi]; tex = f->mat_tex[mi]; shin = f->mat_shininess[mi]; }
    }
    if (SFT_DEBUG_CHECKERBOARD == f->debug_mode) { tex = f->base_tex; base = 0xFFFFFFFFu; }
    const f32 br = (f32)((base >> 16) & 0xFF) * (1.f / 255.f), bg = (f32)((base >> 8) & 0xFF) * (1.f / 255.f), bb = (f32)(base & 0xFF) * (1.f / 255.f);
    const int fl = __builtin_ctz(lanes);
    const int lod = calc_lod(f, tex, cu[fl] - cu[fl ^ 1], cu[fl] - cu[fl ^ 4], cv[fl] - cv[fl ^ 1], cv[fl] - cv[fl ^ 4], x, y);
    f32 ar[8], ag[8], ab[8], fr[8], fg[8], fb[8];
    for (int l = 0; l < 8; l += 1) {
        f32 tr, tg, tb;
        if (lod > 1) sample_nearest(f, tex, lod, cu[l], cv[l], &tr, &tg, &tb);
        else sample_bilinear(f, tex, lod, cu[l], cv[l], &tr, &tg, &tb);
        ar[l] = tr * br; ag[l] = tg * bg; ab[l] = tb * bb;
    }
    if (shin < 0.f) {
        for (int l = 0; l < 8; l += 1) { fr[l] = ar[l] * -shin; fg[l] = ag[l] * -shin; fb[l] = ab[l] * -shin; }
    }
    else {
        f32 dr[8], dg[8], db[8], sr[8] = {0}, sg[8] = {0}, sb[8] = {0}, lms[8];
        const f32 spec_scale = ((shin + 8.f) / 8.f) * 0.04f;
        const int lightmapped = 0xFFFFFFFFu != lm_page;
        for (int l = 0; l < 8; l += 1) {
            lms[l] = 1.f;
            if (lightmapped) {
                sample_bilinear(f, inst->model->lightmaps[lm_page], 0, clu[l], clv[l], &dr[l], &dg[l], &db[l]);
                dr[l] = dr[l] * f->lightmap_brightness; dg[l] = dg[l] * f->lightmap_brightness; db[l] = db[l] * f->lightmap_brightness;
            }
            else {
                dr[l] = f->ambient.x; dg[l] = f->ambient.y; db[l] = f->ambient.z;
                for (i32 d = 0; d < f->n_dir; d += 1) {
                    const v3 L = { f->dir_lights[d].dir.x, f->dir_lights[d].dir.y, f->dir_lights[d].dir.z };
                    const sftvec c = f->dir_lights[d].color;
                    const f32 ndl = smax(vdot(N[l], L), 0.f);
                    dr[l] = pfma(ndl, c.x, dr[l]); dg[l] = pfma(ndl, c.y, dg[l]); db[l] = pfma(ndl, c.z, db[l]);
                    if (ndl > 0.f) {
                        const v3 H = norm_fast(vadd(L, V[l]));
                        const f32 ndh = smax(vdot(N[l], H), 0.f);
                        const f32 spec = ndh * rcp_refined((shin - shin * ndh) + ndh);
                        const f32 term = (spec * spec_scale) * ndl;
                        sr[l] = pfma(term, c.x, sr[l]); sg[l] = pfma(term, c.y, sg[l]); sb[l] = pfma(term, c.z, sb[l]);
                    }
                }
            }
        }
        for (int i = 0; i < ts->n_lights; i += 1) {
            const sft_point_light* L = &f->point_lights[ts->lights[i]];
            const v3 lp = { L->pos.x, L->pos.y, L->pos.z };
            v3 lraw[8], Ld[8];
            f32 d2[8], att[8], ndl[8];
            int any_att = 0, lit = 0;
            for (int l = 0; l < 8; l += 1) {
                lraw[l] = vsub(lp, pos[l]);
                d2[l] = vdot(lraw[l], lraw[l]);
                f32 a = smax(0.f, pfma(-d2[l], L->inv_radius_sqr, 1.f));
                att[l] = a * a;
                if (att[l] > 0.f) any_att = 1;
            }
            if (!any_att) continue;
            for (int l = 0; l < 8; l += 1) {
                Ld[l] = norm_fast(lraw[l]);
                ndl[l] = smax(vdot(N[l], Ld[l]), 0.f);
                if (ndl[l] > 0.f && att[l] > 0.f && (lanes & (1 << l))) lit |= 1 << l;
            }
            f32 msh[8] = {0};
            for (int pi = 0; pi < ts->n_proxies[i]; pi += 1) {
                const proxy_cache* c = &ts->cache[i][pi];
                int act = 0;
                f32 wdv[8];
                for (int l = 0; l < 8; l += 1) {
                    wdv[l] = 0.f - vdot(c->W, lraw[l]);
                    const int between = wdv[l] > 0.f && wdv[l] < d2[l];
                    const f32 tf = wdv[l] * rcp_refined(d2[l]);
                    const f32 cd2 = pfma(-wdv[l], tf, c->wdw);
                    const int sel = lightmapped ? ((lanes >> l) & 1) : ((lit >> l) & 1);
                    if (between && cd2 < c->rb2 && sel) act |= 1 << l;
                }
                if (!act) continue;
                for (int l = 0; l < 8; l += 1) {
                    const f32 lrl = sqrtf(d2[l]);
                    const v3 ao = vsub(pos[l], c->pa);
                    const f32 abl = vdot(c->ab, Ld[l]), abao = vdot(c->ab, ao), lao = vdot(Ld[l], ao);
                    const f32 den = pfma(-abl, abl, c->abdab);
                    const f32 sden = smax(den, SFT_EPSILON);
                    f32 t2 = pfma(-abl, lao, abao) * rcp_refined(sden);
                    t2 = smax(0.f, smin(t2, 1.f));
                    const f32 t1 = pfma(t2, abl, -lao);
                    const int caster = t1 > SFT_EPSILON && t1 < lrl;
                    const v3 pr = { pos[l].x + Ld[l].x * t1, pos[l].y + Ld[l].y * t1, pos[l].z + Ld[l].z * t1 };
                    const v3 pcap = { c->pa.x + c->ab.x * t2, c->pa.y + c->ab.y * t2, c->pa.z + c->ab.z * t2 };
                    const v3 dd = vsub(pr, pcap);
                    const f32 dist2 = vdot(dd, dd);
                    f32 sv = smax(0.f, 1.f - dist2 / c->pr2);
                    if (!((act >> l) & 1) || !caster) sv = 0.f;
                    sv = sv * c->vis_mult;
                    const f32 diff = msh[l] - sv;
                    const f32 h = smax(0.f, 0.2f - fabsf(diff));
                    const f32 blend = (h * h) * (0.25f / 0.2f);
                    msh[l] = smax(msh[l], sv) + blend;
                }
            }
            for (int l = 0; l < 8; l += 1) {
                const f32 si = smax(0.f, 1.f - msh[l]);
                if (lightmapped) {
                    const f32 drop = (1.f - si) * att[l];
                    lms[l] = lms[l] * (1.f - drop);
                }
                else {
                    const f32 di = ndl[l] * att[l];
                    const int li = (lit >> l) & 1;
                    const f32 mdi = li ? di * si : 0.f;
                    dr[l] = pfma(mdi, L->color.x, dr[l]); dg[l] = pfma(mdi, L->color.y, dg[l]); db[l] = pfma(mdi, L->color.z, db[l]);
                    if (li) {
                        const v3 H = norm_fast(vadd(Ld[l], V[l]));
                        const f32 ndh = smax(vdot(N[l], H), 0.f);
                        const f32 spec = ndh * rcp_refined((shin - shin * ndh) + ndh);
                        const f32 term = ((spec * spec_scale) * di) * si;
                        sr[l] = pfma(term, L->color.x, sr[l]); sg[l] = pfma(term, L->color.y, sg[l]); sb[l] = pfma(term, L->color.z, sb[l]);
                    }
                }
            }
        }
        for (int l = 0; l < 8; l += 1) {
            if (lightmapped) {
                const f32 mod = pfma(lms[l], 1.f - 0.4f, 0.4f);
                fr[l] = ar[l] * (dr[l] * mod); fg[l] = ag[l] * (dg[l] * mod); fb[l] = ab[l] * (db[l] * mod);
            }
            else {
                fr[l] = pfma(ar[l], dr[l], sr[l]); fg[l] = pfma(ag[l], dg[l], sg[l]); fb[l] = pfma(ab[l], db[l], sb[l]);
            }
        }
    }
    for (int l = 0; l < 8; l += 1) {
        out[0][l] = smax(0.f, smin(fr[l], 1.f));
        out[1][l] = smax(0.f, smin(fg[l], 1.f));
        out[2][l] = smax(0.f, smin(fb[l], 1.f));
    }
}

static u32 pack_argb(f32 r, f32 g, f32 b) {
    return 0xFF000000u | ((u32)cvtt(r) << 16) | ((u32)cvtt(g) << 8) | (u32)cvtt(b);
}

/* MSAA block loop (6048-6135, 6677-6743): groups in the order of each
   triangle's first (sample, lane) occurrence; each shades once per pixel and
   adds its colour weighted by the fraction of the pixel's samples it covers,
   over a start of the background colour times the empty-sample fraction */
static void resolve_block_aa(port_state* ps, const u32* vid, i32 x, i32 y, i32 tmx, i32 tmy, const tile_shading* ts) {
    const sftx_frame_desc* f = ps->f;
    const int S = f->samples;
    const i32 sw = (i32)(f->half_width * 2.f), sh = (i32)(f->half_height * 2.f);
    const f32 inv_s = 1.f / (f32)S;
    u32 v[4][8];
    int act[4] = {0}, inb = 0, active = 0;
    i32 empty[8] = {0};
    for (int l = 0; l < 8; l += 1) {
        const i32 gx = x + k_ox[l], gy = y + k_oy[l];
        if (gx < sw && gy < sh) inb |= 1 << l;
    }
    for (int q = 0; q < S; q += 1) {
        for (int l = 0; l < 8; l += 1) {
            v[q][l] = vid[q * TP + (y + k_oy[l] - tmy) * TS + (x + k_ox[l] - tmx)];
            const int is_empty = 0xFFFFFFFFu == v[q][l];
            if (!is_empty && (inb >> l & 1)) act[q] |= 1 << l;
            if (is_empty && (inb >> l & 1)) empty[l] += 1;
        }
        active |= act[q];
    }
    if (0 == active) {
        if (f->should_clear) {
            for (int l = 0; l < 8; l += 1) {
                if (inb >> l & 1) ps->out->color[(y + k_oy[l]) * f->width + x + k_ox[l]] = f->clear_color;
            }
        }
        return;
    }
    f32 acc[3][8] = {{0}}, ew[8];
    int any_ew = 0;
    for (int l = 0; l < 8; l += 1) {
        ew[l] = (f32)empty[l] * inv_s;
        if (ew[l] > 0.f) any_ew = 1;
    }
    const u32 cc = f->clear_color;
    if (f->should_clear) {
        const f32 cr = (f32)((cc >> 16) & 0xFF) * (1.f / 255.f), cg = (f32)((cc >> 8) & 0xFF) * (1.f / 255.f), cb = (f32)(cc & 0xFF) * (1.f / 255.f);
        for (int l = 0; l < 8; l += 1) { acc[0][l] = cr * ew[l]; acc[1][l] = cg * ew[l]; acc[2][l] = cb * ew[l]; }
    }
    else if (any_ew) {
        for (int l = 0; l < 8; l += 1) {
            f32 br = 0.f, bg = 0.f, bb = 0.f;
            if (inb >> l & 1) {
                const u32 c = ps->out->color[(y + k_oy[l]) * f->width + x + k_ox[l]];
                br = (f32)((c >> 16) & 0xFF) * (1.f / 255.f); bg = (f32)((c >> 8) & 0xFF) * (1.f / 255.f); bb = (f32)(c & 0xFF) * (1.f / 255.f);
            }
            acc[0][l] = br * ew[l]; acc[1][l] = bg * ew[l]; acc[2][l] = bb * ew[l];
        }
    }
    const int write_mask = active;
    while (active) {
        int fs = 0;
        while (0 == act[fs]) fs += 1;
        const u32 target = v[fs][__builtin_ctz(act[fs])];
        int tri_mask = 0, hits[8] = {0};
        for (int q = 0; q < S; q += 1) {
            int m = 0;
            for (int l = 0; l < 8; l += 1) {
                if ((inb >> l & 1) && v[q][l] == target) { m |= 1 << l; hits[l] += 1; }
            }
            act[q] &= ~m;
            tri_mask |= m;
        }
        active = 0;
        for (int q = 0; q < S; q += 1) active |= act[q];
        tri_t t;
        fetch_tri(ps, target, &t);
        f32 col[3][8];
        shade_group(ps, &t, tri_mask, x, y, ts, col);
        for (int l = 0; l < 8; l += 1) {
            const f32 w = (tri_mask >> l & 1) ? (f32)hits[l] * inv_s : 0.f;
            for (int c = 0; c < 3; c += 1) acc[c][l] = pfma(col[c][l], w, acc[c][l]);
        }
    }
    const int store = f->should_clear ? inb : write_mask;
    for (int l = 0; l < 8; l += 1) {
        if (!(store >> l & 1)) continue;
        ps->out->color[(y + k_oy[l]) * f->width + x + k_ox[l]] =
            pack_argb(smin(acc[0][l], 1.f) * 255.f, smin(acc[1][l], 1.f) * 255.f, smin(acc[2][l], 1.f) * 255.f);
    }
}

static void resolve_tile(port_state* ps, const f32* depth, const u32* vid, i32 tmx, i32 tmy) {
    const sftx_frame_desc* f = ps->f;
    const i32 sw = (i32)(f->half_width * 2.f), sh = (i32)(f->half_height * 2.f);
    const i32 lx = SFT_MIN(tmx + TS, sw), ly = SFT_MIN(tmy + TS, sh);
    f32 zmin = 1e9f, zmax = -1e9f;
    for (i32 i = 0; i < TP * f->samples; i += 1) {
        if (depth[i] > 0.0001f) {
            if (depth[i] < zmin) zmin = depth[i];
            if (depth[i] > zmax) zmax = depth[i];
        }
    }
    static tile_shading ts;
    ts.n_lights = tile_lights(f, zmin, zmax, tmx, tmy, ts.lights);
    build_proxy_cache(ps, &ts);
    for (i32 y = tmy; y < ly; y += 2) {
        for (i32 x = tmx; x < lx; x += 4) {
            if (f->samples > 1) {
                resolve_block_aa(ps, vid, x, y, tmx, tmy, &ts);
                continue;
            }
            u32 vids[8];
            int inb = 0, active = 0;
            for (int l = 0; l < 8; l += 1) {
                const i32 gx = x + k_ox[l], gy = y + k_oy[l];
                if (gx < sw && gy < sh) inb |= 1 << l;
                vids[l] = vid[(gy - tmy) * TS + (gx - tmx)];
                if ((inb & (1 << l)) && 0xFFFFFFFFu != vids[l]) active |= 1 << l;
                if (f->should_clear && (inb & (1 << l)) && 0xFFFFFFFFu == vids[l]) ps->out->color[gy * f->width + gx] = f->clear_color;
            }
            while (active) {
                const u32 target = vids[__builtin_ctz(active)];
                int lanes = 0;
                for (int l = 0; l < 8; l += 1) {
                    if ((inb & (1 << l)) && vids[l] == target) lanes |= 1 << l;
                }
                lanes &= active;
                active &= ~lanes;
                tri_t t;
                fetch_tri(ps, target, &t);
                f32 col[3][8];
                shade_group(ps, &t, lanes, x, y, &ts, col);
                for (int l = 0; l < 8; l += 1) {
                    if (!(lanes & (1 << l))) continue;
                    ps->out->color[(y + k_oy[l]) * f->width + x + k_ox[l]] = pack_argb(col[0][l] * 255.f, col[1][l] * 255.f, col[2][l] * 255.f);
                }
            }
        }
    }
}

/* ------------------------------------------------------------ frame
   sft_present + worker phases, sofren2.c:8139-8289, 6822-6867 */
int sftx_port_render(const sftx_frame_desc* f_in, sftx_port_out* out) {
    /* textures exported as level 0 only: build their chains here */
    sftx_frame_desc fl = *f_in;
    const sftx_frame_desc* f = &fl;
    sftx_texture_desc* texs = NULL;
    if (fl.n_textures > 0) {
        texs = (sftx_texture_desc*)malloc((size_t)fl.n_textures * sizeof(sftx_texture_desc));
        memcpy(texs, f_in->textures, (size_t)fl.n_textures * sizeof(sftx_texture_desc));
        for (i32 i = 0; i < fl.n_textures; i += 1) {
            if (texs[i].texels || !texs[i].level0) continue;
            u32* chain = (u32*)malloc((size_t)texs[i].total * 4);
            sftx_port_build_texture(texs[i].level0, texs[i].w, texs[i].h, chain);
            texs[i].texels = chain;
        }
        fl.textures = texs;
    }
    port_state ps;
    memset(&ps, 0, sizeof(ps));
    ps.f = f;
    ps.out = out;
    ps.ss = (f32*)out->ss;
    const i32 nt = f->tile_cols * f->tile_rows;
    ps.bins = (bin_t*)calloc((size_t)(nt > 0 ? nt : 1), sizeof(bin_t));
    memset(out->ss_vis, 0, f->total_ss);
    tlas_build(&ps);
    occlusion_pass(&ps);
    if (out->proxy_vis) {
        for (i32 p = 0; p < f->n_proxies; p += 1) memcpy(out->proxy_vis + (u64)p * 64, ps.proxies[p].light_vis, 64);
    }
    out->invalid_packets = ps.invalid_packets;
    geometry_pass(&ps);
    out->clip_count = ps.clip_next < out->clip_cap ? ps.clip_next : out->clip_cap;

    const i32 gw = f->tile_cols * TS, gh = f->tile_rows * TS;
    const int S = f->samples;
    static f32 depth[TP * 4];
    static u32 vid[TP * 4];
    for (i32 t = 0; t < nt; t += 1) {
        const i32 tmx = (t % f->tile_cols) * TS, tmy = (t / f->tile_cols) * TS;
        memset(depth, 0, sizeof(depth));
        memset(vid, 0xFF, sizeof(vid));
        for (u32 k = 0; k < ps.bins[t].n; k += 1) raster_payload(&ps, ps.bins[t].p[k], depth, vid, tmx, tmy);
        if (out->vid) {
            for (int q = 0; q < S; q += 1) {
                for (i32 y = 0; y < TS; y += 1) {
                    for (i32 x = 0; x < TS; x += 1) {
                        const u64 g = (u64)q * gw * gh + (u64)(tmy + y) * gw + tmx + x;
                        out->vid[g] = vid[q * TP + y * TS + x];
                        out->depth[g] = depth[q * TP + y * TS + x];
                    }
                }
            }
        }
        resolve_tile(&ps, depth, vid, tmx, tmy);
        free(ps.bins[t].p);
    }
    free(ps.bins);
    free(ps.proxies);
    free(ps.tlas);
    if (texs) {
        for (i32 i = 0; i < fl.n_textures; i += 1) {
            if (!f_in->textures[i].texels && texs[i].texels) free((void*)texs[i].texels);
        }
        free(texs);
    }
    return ps.clip_overflow ? -1 : 0;
}

u32 sftx_port_ss_count(const sftx_frame_desc* f) { return f->total_ss; }

i32 sftx_port_proxy_count(const sftx_frame_desc* f) { return f->n_proxies; }

i32 sftx_port_samples(const sftx_frame_desc* f) { return f->samples; }

void sftx_port_grid(const sftx_frame_desc* f, i32* cols, i32* rows, i32* width, i32* height) {
    *cols = f->tile_cols;
    *rows = f->tile_rows;
    *width = f->width;
    *height = f->height;
}
