/*
 * scenes.c — seeded synthetic workloads for the sofren2 API (see scenes.h).
 *
 * Only public API calls and the public model structs are used, so this file
 * compiles against any implementation of the API.  It must be compiled with
 * -ffp-contract=off so the host-side float math is identical for every build.
 */
#include <unistd.h>
#include "scenes.h"

/* ---------------- PCG32 (O'Neill, pcg-random.org, XSH RR variant) -------- */
typedef struct { u64 state, inc; } sftx_rng;

static u32 sftx_rng_next(sftx_rng* r) {
    const u64 old = r->state;
    r->state = old * 6364136223846793005ULL + r->inc;
    const u32 xs = (u32)(((old >> 18u) ^ old) >> 27u);
    const u32 rot = (u32)(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}

static void sftx_rng_seed(sftx_rng* r, u64 seed, u64 stream) {
    r->state = 0u;
    r->inc = (stream << 1u) | 1u;
    sftx_rng_next(r);
    r->state += seed;
    sftx_rng_next(r);
}

/* uniform float in [0,1) with 24 random bits */
static f32 sftx_u01(sftx_rng* r) {
    return (f32)(sftx_rng_next(r) >> 8) * (1.f / 16777216.f);
}

static f32 sftx_uniform(sftx_rng* r, f32 lo, f32 hi) {
    return lo + (hi - lo) * sftx_u01(r);
}

static u32 sftx_pack_rgb(f32 r, f32 g, f32 b) {
    const u32 ri = (u32)(r * 255.f + 0.5f), gi = (u32)(g * 255.f + 0.5f), bi = (u32)(b * 255.f + 0.5f);
    return 0xFF000000u | (ri << 16) | (gi << 8) | bi;
}

/* ---------------- scene state ---------------- */
#define SFTX_MAX_SPHERES 16

static struct {
    i32 kind;
    u32 seed;
    i32 width, height;

    /* plane + spheres */
    sft_mat_handle plane_mat;
    sft_mat_handle sphere_mat[SFTX_MAX_SPHERES];
    f32 sphere_r[SFTX_MAX_SPHERES], sphere_cx[SFTX_MAX_SPHERES], sphere_cz[SFTX_MAX_SPHERES];

    u32 tri_count, meshlet_count;
} g_scene;

/* ---------------- workload 0: ground plane + 16 dmesh spheres -------------- */
static void sftx_plane_spheres_create(void) {
    sftx_rng rng;
    sftx_rng_seed(&rng, g_scene.seed, 0x5f3759dfULL);

    const f32 pr = sftx_u01(&rng), pg = sftx_u01(&rng), pb = sftx_u01(&rng);
    g_scene.plane_mat = sft_material_create(SFT_INVALID_HANDLE, sftx_pack_rgb(pr, pg, pb), 16.f);

    for (i32 i = 0; i < SFTX_MAX_SPHERES; i += 1) {
        g_scene.sphere_r[i] = sftx_uniform(&rng, 0.5f, 1.5f);
        g_scene.sphere_cx[i] = sftx_uniform(&rng, -8.f, 8.f);
        g_scene.sphere_cz[i] = sftx_uniform(&rng, -8.f, 8.f);
        const f32 r = sftx_u01(&rng), g = sftx_u01(&rng), b = sftx_u01(&rng);
        g_scene.sphere_mat[i] = sft_material_create(SFT_INVALID_HANDLE, sftx_pack_rgb(r, g, b), 16.f);
    }

    /* 400 plane quads (25 meshlets) + 16 spheres of 17*32 quads (34 meshlets each) */
    g_scene.tri_count = 400u * 2u + SFTX_MAX_SPHERES * 17u * 32u * 2u;
    g_scene.meshlet_count = 25u + SFTX_MAX_SPHERES * 34u;
}

static void sftx_plane_spheres_render(i32 frame) {
    (void)frame;
    const sftvec cam = SFT_V(0.f, 6.f, -14.f);
    sft_set_camera(cam,
        sft_m_view_look_at(cam, SFT_V(0.f, 0.f, 0.f), SFT_V(0.f, 1.f, 0.f)),
        sft_m_proj(60.f, (f32)g_scene.width / (f32)g_scene.height, 0.1f, 1000.f));

    sft_set_ambient_light(0.35f, 0.35f, 0.35f);
    sft_draw_point_light(SFT_V(0.f, 10.f, 0.f), 1.f, 1.f, 1.f, 1.f, 30.f, 1);

    /* ground plane: unit quads on [-10,10]^2 at y = 0, wound clockwise seen from above */
    sft_dmesh* plane = sft_dmesh_begin(SFT_DMESH_QUADS, g_scene.plane_mat, 400);
    for (i32 z = 0; z < 20; z += 1) {
        for (i32 x = 0; x < 20; x += 1) {
            const f32 x0 = (f32)(x - 10), x1 = (f32)(x - 9);
            const f32 z0 = (f32)(z - 10), z1 = (f32)(z - 9);
            sft_dmesh_push_quad(plane,
                SFT_V(x0, 0.f, z0), SFT_V(x0, 0.f, z1), SFT_V(x1, 0.f, z1), SFT_V(x1, 0.f, z0),
                0.f, 0.f, 1.f, 1.f, SFT_V(0.f, 1.f, 0.f));
        }
    }
    sft_dmesh_end(plane, NULL);

    for (i32 i = 0; i < SFTX_MAX_SPHERES; i += 1) {
        sft_dmesh* dm = sft_dmesh_begin(SFT_DMESH_QUADS, g_scene.sphere_mat[i], 17 * 32);
        sft_draw_sphere(dm, 16, 32);
        const f32 r = g_scene.sphere_r[i];
        const sftmat t = sft_m_mul(sft_m_scale(r, r, r), sft_m_translate(g_scene.sphere_cx[i], r, g_scene.sphere_cz[i]));
        sft_dmesh_end(dm, &t);
    }
}

/* ---------------- workload 1: textured box-room lattice -------------------
   4096 axis-aligned boxes (sizes log-uniform in [0.1,4] per axis) jittered on a
   16^3 lattice of spacing 5 inside an inward-facing room, 16 value-noise
   textures, one baked lightmap page.  Faces are (n+1)^2-vertex grids, n = 2 for
   small boxes (one meshlet per box, cone test off) and n = 3 for boxes larger
   than 2.8 (two meshlets of three faces each with a real normal cone).  The
   model is written as QSF2 and loaded with sft_model_load_qsf. */
#define SFTX_LATTICE 16
#define SFTX_BOXES (SFTX_LATTICE * SFTX_LATTICE * SFTX_LATTICE)
#define SFTX_TEXTURES 16
#define SFTX_TEX_SIZE 256
#define SFTX_LM_SIZE 1024
#define SFTX_ROOM_HALF 42.f
#define SFTX_WALL_N 7

typedef struct {
    u32 vert_count, tri_count;
    f32 x[64], y[64], z[64], nx[64], ny[64], nz[64], u[64], v[64], lu[64], lv[64];
    u8 inds[128 * 3];
    f32 cone[4];
} sftx_mbuild;

typedef struct {
    sftx_mbuild* m;
    u32 count, cap;
} sftx_mlist;

static sftx_mbuild* sftx_mlist_new(sftx_mlist* l) {
    if (l->count == l->cap) {
        l->cap = l->cap ? l->cap * 2 : 256;
        l->m = (sftx_mbuild*)realloc(l->m, l->cap * sizeof(sftx_mbuild));
    }
    sftx_mbuild* m = &l->m[l->count++];
    memset(m, 0, sizeof(*m));
    m->cone[2] = 1.f;
    m->cone[3] = 2.f;
    return m;
}

/* value noise: bilinear blend of a random lattice, three octaves */
static f32 sftx_noise_lattice(const f32* g, i32 gs, f32 x, f32 y) {
    const i32 x0 = (i32)floorf(x), y0 = (i32)floorf(y);
    const f32 fx = x - (f32)x0, fy = y - (f32)y0;
    const f32 sx = fx * fx * (3.f - 2.f * fx), sy = fy * fy * (3.f - 2.f * fy);
    const i32 xa = x0 & (gs - 1), xb = (x0 + 1) & (gs - 1), ya = y0 & (gs - 1), yb = (y0 + 1) & (gs - 1);
    const f32 a = g[ya * gs + xa], b = g[ya * gs + xb], c = g[yb * gs + xa], d = g[yb * gs + xb];
    const f32 top = a + (b - a) * sx, bot = c + (d - c) * sx;
    return top + (bot - top) * sy;
}

static void sftx_value_noise_texture(sftx_rng* rng, u32* px) {
    f32 g[3][32 * 32];
    const i32 gs[3] = { 8, 16, 32 };
    for (i32 o = 0; o < 3; o += 1) {
        for (i32 i = 0; i < gs[o] * gs[o]; i += 1) g[o][i] = sftx_u01(rng);
    }
    const f32 tint[3] = { sftx_uniform(rng, 0.3f, 1.f), sftx_uniform(rng, 0.3f, 1.f), sftx_uniform(rng, 0.3f, 1.f) };
    for (i32 y = 0; y < SFTX_TEX_SIZE; y += 1) {
        for (i32 x = 0; x < SFTX_TEX_SIZE; x += 1) {
            f32 n = 0.f, amp = 0.5f;
            for (i32 o = 0; o < 3; o += 1) {
                const f32 s = (f32)gs[o] / (f32)SFTX_TEX_SIZE;
                n += amp * sftx_noise_lattice(g[o], gs[o], (f32)x * s, (f32)y * s);
                amp *= 0.5f;
            }
            n = n / 0.875f;
            px[y * SFTX_TEX_SIZE + x] = sftx_pack_rgb(n * tint[0], n * tint[1], n * tint[2]);
        }
    }
}

/* appends an (n+1)^2 grid face: corner o, edge vectors du, dv, normal nrm,
   lightmap cell at texel (lx, ly) */
static void sftx_push_face(sftx_mbuild* m, i32 n, sftvec o, sftvec du, sftvec dv, sftvec nrm, sftvec ua, sftvec va,
                           i32 lx, i32 ly) {
    const u32 base = m->vert_count;
    for (i32 j = 0; j <= n; j += 1) {
        for (i32 i = 0; i <= n; i += 1) {
            const f32 s = (f32)i / (f32)n, t = (f32)j / (f32)n;
            const u32 k = m->vert_count++;
            const f32 px = o.x + du.x * s + dv.x * t, py = o.y + du.y * s + dv.y * t, pz = o.z + du.z * s + dv.z * t;
            m->x[k] = px; m->y[k] = py; m->z[k] = pz;
            m->nx[k] = nrm.x; m->ny[k] = nrm.y; m->nz[k] = nrm.z;
            m->u[k] = (px * ua.x + py * ua.y + pz * ua.z) * 0.5f;
            m->v[k] = (px * va.x + py * va.y + pz * va.z) * 0.5f;
            m->lu[k] = ((f32)(lx + i) + 0.5f) / (f32)SFTX_LM_SIZE;
            m->lv[k] = ((f32)(ly + j) + 0.5f) / (f32)SFTX_LM_SIZE;
        }
    }
    for (i32 j = 0; j < n; j += 1) {
        for (i32 i = 0; i < n; i += 1) {
            const u8 a = (u8)(base + j * (n + 1) + i), b = (u8)(a + 1), c = (u8)(a + n + 1), d = (u8)(c + 1);
            u8* t = &m->inds[m->tri_count * 3];
            t[0] = a; t[1] = c; t[2] = d;
            t[3] = a; t[4] = d; t[5] = b;
            m->tri_count += 2;
        }
    }
}

typedef struct { sftvec o, du, dv, n, ua, va; } sftx_face;

/* six outward faces of a box, wound clockwise as seen from outside */
static void sftx_box_faces(sftvec c, sftvec h, sftx_face f[6]) {
    const f32 x0 = c.x - h.x, x1 = c.x + h.x, y0 = c.y - h.y, y1 = c.y + h.y, z0 = c.z - h.z, z1 = c.z + h.z;
    const f32 wx = 2.f * h.x, wy = 2.f * h.y, wz = 2.f * h.z;
    /* +x, +y, +z, -x, -y, -z */
    f[0] = (sftx_face){ SFT_V(x1, y0, z0), SFT_V(0, 0, wz), SFT_V(0, wy, 0), SFT_V(1, 0, 0), SFT_V(0, 0, 1), SFT_V(0, 1, 0) };
    f[1] = (sftx_face){ SFT_V(x0, y1, z0), SFT_V(wx, 0, 0), SFT_V(0, 0, wz), SFT_V(0, 1, 0), SFT_V(1, 0, 0), SFT_V(0, 0, 1) };
    f[2] = (sftx_face){ SFT_V(x0, y0, z1), SFT_V(0, wy, 0), SFT_V(wx, 0, 0), SFT_V(0, 0, 1), SFT_V(0, 1, 0), SFT_V(1, 0, 0) };
    f[3] = (sftx_face){ SFT_V(x0, y0, z0), SFT_V(0, wy, 0), SFT_V(0, 0, wz), SFT_V(-1, 0, 0), SFT_V(0, 1, 0), SFT_V(0, 0, 1) };
    f[4] = (sftx_face){ SFT_V(x0, y0, z0), SFT_V(0, 0, wz), SFT_V(wx, 0, 0), SFT_V(0, -1, 0), SFT_V(0, 0, 1), SFT_V(1, 0, 0) };
    f[5] = (sftx_face){ SFT_V(x0, y0, z0), SFT_V(wx, 0, 0), SFT_V(0, wy, 0), SFT_V(0, 0, -1), SFT_V(1, 0, 0), SFT_V(0, 1, 0) };
}

static void sftx_meshlet_bounds(const sftx_mbuild* m, f32 out[14]) {
    f32 mn[3] = { 1e30f, 1e30f, 1e30f }, mx[3] = { -1e30f, -1e30f, -1e30f };
    for (u32 i = 0; i < m->vert_count; i += 1) {
        const f32 p[3] = { m->x[i], m->y[i], m->z[i] };
        for (int k = 0; k < 3; k += 1) {
            if (p[k] < mn[k]) mn[k] = p[k];
            if (p[k] > mx[k]) mx[k] = p[k];
        }
    }
    const f32 cx = (mn[0] + mx[0]) * 0.5f, cy = (mn[1] + mx[1]) * 0.5f, cz = (mn[2] + mx[2]) * 0.5f;
    const f32 dx = mx[0] - cx, dy = mx[1] - cy, dz = mx[2] - cz;
    out[0] = cx; out[1] = cy; out[2] = cz; out[3] = sqrtf(dx * dx + dy * dy + dz * dz);
    out[4] = mn[0]; out[5] = mn[1]; out[6] = mn[2]; out[7] = mx[0]; out[8] = mx[1]; out[9] = mx[2];
    out[10] = m->cone[0]; out[11] = m->cone[1]; out[12] = m->cone[2]; out[13] = m->cone[3];
}

static void sftx_fwrite(FILE* f, const void* p, u64 n) {
    if (n && 1 != fwrite(p, n, 1, f)) {
        fprintf(stderr, "scene file write failed\n");
        exit(2);
    }
}

/* writes meshes (one per texture) as QSF2 (layout read by sofren2.c:7816-7967) */
static void sftx_write_qsf(const char* path, sftx_mlist* meshes, i32 mesh_count, u32** tex, i32 tex_count,
                           const u32* lightmap) {
    FILE* f = fopen(path, "wb");
    if (!f) {
        fprintf(stderr, "cannot write %s\n", path);
        exit(2);
    }
    const u32 version = 2;
    const u64 game_data = 0;
    sftx_fwrite(f, "QSF2", 4);
    sftx_fwrite(f, &version, 4);
    sftx_fwrite(f, &game_data, 8);
    const u32 tc = (u32)tex_count, ts = SFTX_TEX_SIZE, ls = SFTX_LM_SIZE, one = 1;
    sftx_fwrite(f, &tc, 4);
    for (i32 t = 0; t < tex_count; t += 1) {
        sftx_fwrite(f, &ts, 4);
        sftx_fwrite(f, &ts, 4);
        sftx_fwrite(f, tex[t], (u64)ts * ts * 4);
    }
    sftx_fwrite(f, &one, 4);
    sftx_fwrite(f, &ls, 4);
    sftx_fwrite(f, &ls, 4);
    sftx_fwrite(f, lightmap, (u64)ls * ls * 4);
    u32 total = 0;
    for (i32 i = 0; i < mesh_count; i += 1) total += meshes[i].count;
    const u32 mc = (u32)mesh_count, nodes = 0, root = 0;
    sftx_fwrite(f, &mc, 4);
    sftx_fwrite(f, &total, 4);
    sftx_fwrite(f, &nodes, 4);
    sftx_fwrite(f, &root, 4);
    u32 offset = 0;
    for (i32 i = 0; i < mesh_count; i += 1) {
        const sftx_mlist* l = &meshes[i];
        const u32 tex_ind = (u32)i, page = 0;
        sftx_fwrite(f, &l->count, 4);
        sftx_fwrite(f, &offset, 4);
        sftx_fwrite(f, &tex_ind, 4);
        sftx_fwrite(f, &page, 4);
        for (u32 m = 0; m < l->count; m += 1) {
            const sftx_mbuild* b = &l->m[m];
            sftx_fwrite(f, b->x, 256); sftx_fwrite(f, b->y, 256); sftx_fwrite(f, b->z, 256);
        }
        for (u32 m = 0; m < l->count; m += 1) {
            const sftx_mbuild* b = &l->m[m];
            sftx_fwrite(f, b->nx, 256); sftx_fwrite(f, b->ny, 256); sftx_fwrite(f, b->nz, 256);
            sftx_fwrite(f, b->u, 256); sftx_fwrite(f, b->v, 256);
            sftx_fwrite(f, b->lu, 256); sftx_fwrite(f, b->lv, 256);
        }
        for (u32 m = 0; m < l->count; m += 1) sftx_fwrite(f, l->m[m].inds, 384);
        f32* bounds = (f32*)malloc((u64)l->count * 14 * sizeof(f32) + 4);
        for (u32 m = 0; m < l->count; m += 1) sftx_meshlet_bounds(&l->m[m], &bounds[m * 14]);
        for (int k = 0; k < 14; k += 1) {
            for (u32 m = 0; m < l->count; m += 1) sftx_fwrite(f, &bounds[m * 14 + k], 4);
        }
        free(bounds);
        offset += l->count;
    }
    for (i32 i = 0; i < mesh_count; i += 1) {
        const u16 mi = (u16)i;
        for (u32 m = 0; m < meshes[i].count; m += 1) sftx_fwrite(f, &mi, 2);
    }
    fclose(f);
}

static void sftx_pad_meshlet(sftx_mbuild* m) {
    for (u32 i = m->vert_count; i < 64; i += 1) {
        m->x[i] = m->x[0]; m->y[i] = m->y[0]; m->z[i] = m->z[0];
        m->nx[i] = m->nx[0]; m->ny[i] = m->ny[0]; m->nz[i] = m->nz[0];
        m->u[i] = m->u[0]; m->v[i] = m->v[0]; m->lu[i] = m->lu[0]; m->lv[i] = m->lv[0];
    }
}

static struct {
    sft_arena arena;
    sft_model* model;
    sft_instance inst;
    sftvec light_pos[8];
    sftvec light_col[8];
} g_room;

static void sftx_box_room_create(void) {
    sftx_rng rng;
    sftx_rng_seed(&rng, g_scene.seed, 0x2545F491ULL);
    u32* tex[SFTX_TEXTURES];
    for (i32 t = 0; t < SFTX_TEXTURES; t += 1) {
        tex[t] = (u32*)malloc(SFTX_TEX_SIZE * SFTX_TEX_SIZE * 4);
        sftx_value_noise_texture(&rng, tex[t]);
    }
    u32* lm = (u32*)malloc((u64)SFTX_LM_SIZE * SFTX_LM_SIZE * 4);
    for (u32 i = 0; i < (u32)SFTX_LM_SIZE * SFTX_LM_SIZE; i += 1) {
        const f32 l = sftx_uniform(&rng, 0.2f, 1.f);
        lm[i] = sftx_pack_rgb(l, l, l);
    }
    sftx_mlist meshes[SFTX_TEXTURES];
    memset(meshes, 0, sizeof(meshes));
    i32 cell = 0;   /* 4x4-texel lightmap cells, 256 per row */
    const f32 ln_lo = logf(0.1f), ln_hi = logf(4.f);
    u32 tris = 0;
    for (i32 bz = 0; bz < SFTX_LATTICE; bz += 1) {
        for (i32 by = 0; by < SFTX_LATTICE; by += 1) {
            for (i32 bx = 0; bx < SFTX_LATTICE; bx += 1) {
                const sftvec c = SFT_V((f32)bx * 5.f - 37.5f + sftx_uniform(&rng, -1.f, 1.f),
                                       (f32)by * 5.f - 37.5f + sftx_uniform(&rng, -1.f, 1.f),
                                       (f32)bz * 5.f - 37.5f + sftx_uniform(&rng, -1.f, 1.f));
                const sftvec h = SFT_V(0.5f * expf(sftx_uniform(&rng, ln_lo, ln_hi)),
                                       0.5f * expf(sftx_uniform(&rng, ln_lo, ln_hi)),
                                       0.5f * expf(sftx_uniform(&rng, ln_lo, ln_hi)));
                const i32 t = (i32)(sftx_rng_next(&rng) % SFTX_TEXTURES);
                const f32 big = SFT_MAX(h.x, SFT_MAX(h.y, h.z)) * 2.f;
                const i32 n = big > 2.8f ? 3 : 2;
                sftx_face f[6];
                sftx_box_faces(c, h, f);
                if (2 == n) {
                    sftx_mbuild* m = sftx_mlist_new(&meshes[t]);
                    for (int k = 0; k < 6; k += 1, cell += 1) {
                        sftx_push_face(m, n, f[k].o, f[k].du, f[k].dv, f[k].n, f[k].ua, f[k].va, (cell % 256) * 4, (cell / 256) * 4);
                    }
                    sftx_pad_meshlet(m);
                    tris += m->tri_count;
                }
                else {
                    for (int half = 0; half < 2; half += 1) {
                        sftx_mbuild* m = sftx_mlist_new(&meshes[t]);
                        for (int k = half * 3; k < half * 3 + 3; k += 1, cell += 1) {
                            sftx_push_face(m, n, f[k].o, f[k].du, f[k].dv, f[k].n, f[k].ua, f[k].va, (cell % 256) * 4, (cell / 256) * 4);
                        }
                        const f32 s = half ? -0.57735027f : 0.57735027f;
                        m->cone[0] = s; m->cone[1] = s; m->cone[2] = s;
                        m->cone[3] = 0.82f;   /* > sin(54.74 deg): all three face normals inside */
                        sftx_pad_meshlet(m);
                        tris += m->tri_count;
                    }
                }
            }
        }
    }
    /* the room: six inward 7x7-quad walls, lightmap cells of 8x8 texels */
    {
        const sftvec c = SFT_V(0.f, 0.f, 0.f), h = SFT_V(SFTX_ROOM_HALF, SFTX_ROOM_HALF, SFTX_ROOM_HALF);
        sftx_face f[6];
        sftx_box_faces(c, h, f);
        for (int k = 0; k < 6; k += 1) {
            /* flip winding and normal: swap the edge vectors */
            sftx_mbuild* m = sftx_mlist_new(&meshes[k % SFTX_TEXTURES]);
            const sftvec nn = SFT_V(-f[k].n.x, -f[k].n.y, -f[k].n.z);
            const i32 wc = 2048 + k;   /* wall cells after the box cells, in 8x8 units */
            sftx_push_face(m, SFTX_WALL_N, f[k].o, f[k].dv, f[k].du, nn, f[k].va, f[k].ua, (wc % 128) * 8, (wc / 128) * 8);
            m->cone[0] = nn.x; m->cone[1] = nn.y; m->cone[2] = nn.z; m->cone[3] = 0.05f;
            sftx_pad_meshlet(m);
            tris += m->tri_count;
        }
    }
    char path[256];
    const char* tmp = getenv("TMPDIR");
    snprintf(path, sizeof(path), "%s/sftx_room_%u_%d.qsf", tmp ? tmp : "/tmp", g_scene.seed, (int)getpid());
    sftx_write_qsf(path, meshes, SFTX_TEXTURES, tex, SFTX_TEXTURES, lm);
    u32 meshlets = 0;
    for (i32 i = 0; i < SFTX_TEXTURES; i += 1) {
        meshlets += meshes[i].count;
        free(meshes[i].m);
        free(tex[i]);
    }
    free(lm);
    sft_arena_init(&g_room.arena, 1ULL << 32);
    g_room.model = sft_model_load_qsf(path, &g_room.arena);
    remove(path);
    if (!g_room.model) {
        fprintf(stderr, "box room model failed to load\n");
        exit(2);
    }
    memset(&g_room.inst, 0, sizeof(g_room.inst));
    g_room.inst.model = g_room.model;
    g_room.inst.transform = sft_m_identity();
    for (int i = 0; i < 8; i += 1) {
        g_room.light_pos[i] = SFT_V(sftx_uniform(&rng, -35.f, 35.f), sftx_uniform(&rng, -35.f, 35.f), sftx_uniform(&rng, -35.f, 35.f));
        g_room.light_col[i] = SFT_V(sftx_uniform(&rng, 0.5f, 1.f), sftx_uniform(&rng, 0.5f, 1.f), sftx_uniform(&rng, 0.5f, 1.f));
    }
    g_scene.tri_count = tris;
    g_scene.meshlet_count = meshlets;
}

/* fixed 120-frame orbit through the lattice */
static void sftx_box_room_render(i32 frame) {
    const f32 th = 2.f * (f32)SFT_PI * (f32)(frame % 120) / 120.f;
    const sftvec cam = SFT_V(25.f * cosf(th), 3.f + 4.f * sinf(2.f * th), 25.f * sinf(th));
    const sftvec target = SFT_V(5.f * sinf(th), 0.f, -5.f * cosf(th));
    sft_set_camera(cam, sft_m_view_look_at(cam, target, SFT_V(0.f, 1.f, 0.f)),
                   sft_m_proj(70.f, (f32)g_scene.width / (f32)g_scene.height, 0.1f, 1000.f));
    sft_set_ambient_light(0.3f, 0.3f, 0.3f);
    for (int i = 0; i < 8; i += 1) {
        sft_draw_point_light(g_room.light_pos[i], g_room.light_col[i].x, g_room.light_col[i].y, g_room.light_col[i].z,
                             1.f, 20.f, (u32)(i + 1));
    }
    sft_draw_instance(&g_room.inst);
}

/* ---------------- dispatch ---------------- */
i32 sftx_scene_create(i32 kind, u32 seed, i32 width, i32 height) {
    memset(&g_scene, 0, sizeof(g_scene));
    g_scene.kind = kind;
    g_scene.seed = seed;
    g_scene.width = width;
    g_scene.height = height;
    sft_set_clear(1);
    switch (kind) {
        case SFTX_SCENE_PLANE_SPHERES: sftx_plane_spheres_create(); return 0;
        case SFTX_SCENE_BOX_ROOM: sftx_box_room_create(); return 0;
        default: return -1;
    }
}

void sftx_scene_record(i32 frame, u32* color, i32 pitch_bytes) {
    sft_set_color_buffer(color, pitch_bytes);
    sft_frame_begin();
    switch (g_scene.kind) {
        case SFTX_SCENE_PLANE_SPHERES: sftx_plane_spheres_render(frame); break;
        case SFTX_SCENE_BOX_ROOM: sftx_box_room_render(frame); break;
        default: break;
    }
}

void sftx_scene_render(i32 frame, u32* color, i32 pitch_bytes) {
    sftx_scene_record(frame, color, pitch_bytes);
    sft_present();
    sft_frame_end();
}

u32 sftx_scene_triangle_count(void) { return g_scene.tri_count; }
u32 sftx_scene_meshlet_count(void) { return g_scene.meshlet_count; }

void sftx_scene_destroy(void) {
    memset(&g_scene, 0, sizeof(g_scene));
}
