/*
 * scenes.c — seeded synthetic workloads for the sofren2 API (see scenes.h).
 *
 * Only public API calls and the public model structs are used, so this file
 * compiles against any implementation of the API.  It must be compiled with
 * -ffp-contract=off so the host-side float math is identical for every build.
 */
#include <stddef.h>
#include <unistd.h>
#include <sys/mman.h>
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

typedef struct {
    i32 kind;
    u32 seed;
    i32 width, height;

    /* plane + spheres */
    sft_mat_handle plane_mat;
    sft_mat_handle sphere_mat[SFTX_MAX_SPHERES];
    f32 sphere_r[SFTX_MAX_SPHERES], sphere_cx[SFTX_MAX_SPHERES], sphere_cz[SFTX_MAX_SPHERES];

    /* box lattice, lightmap page and crowd sizes (the served kind varies them) */
    i32 lattice, lm_size, n_chars, n_lights;

    u32 tri_count, meshlet_count;
} sftx_scene_common;

#define SFTX_CHARS 32
#define SFTX_MAX_CHARS 64
#define SFTX_BONES 10
#define SFTX_BANDS 96
#define SFTX_SEGS 32
#define SFTX_BONE_LEN 0.18f
#define SFTX_SHADOW_LIGHTS 16

typedef struct {
    sft_arena arena;
    sft_model* model;
    sft_instance inst;
    sftvec light_pos[8];
    sftvec light_col[8];
    /* the map workload: lights read from the file's game data */
    i32 map_lights;
    sftvec map_pos[64], map_col[64];
    f32 map_radius[64];
} sftx_room_state;

typedef struct {
    sft_model* human;
    sft_model* person[SFTX_MAX_CHARS];     /* shallow copies: own mesh record (material), shared meshlets */
    sft_mat_handle mat[SFTX_MAX_CHARS];
    sftmat xform[SFTX_MAX_CHARS];
    f32 phase[SFTX_MAX_CHARS][SFTX_BONES][2];
    sftmat posed[SFTX_MAX_CHARS][SFTX_BONES];
    sftvec light_pos[SFTX_SHADOW_LIGHTS], light_col[SFTX_SHADOW_LIGHTS];
    sftmat copy_xform[40];                 /* the instanced workload's room copies */
} sftx_crowd_state;

/* the features workload: every shading branch the other scenes miss */
#define SFTX_FEAT_OBJECTS 60
#define SFTX_FEAT_EXTRA 6
#define SFTX_FEAT_LIGHTS 24
typedef struct {
    sft_tex_handle tex[6];
    sft_mat_handle mat[8];
    i32 obj_kind[SFTX_FEAT_OBJECTS + SFTX_FEAT_EXTRA], obj_mat[SFTX_FEAT_OBJECTS + SFTX_FEAT_EXTRA];
    sftmat obj_xform[SFTX_FEAT_OBJECTS + SFTX_FEAT_EXTRA];
    f32 obj_param[SFTX_FEAT_OBJECTS + SFTX_FEAT_EXTRA][4];
    sftvec light_pos[SFTX_FEAT_LIGHTS], light_col[SFTX_FEAT_LIGHTS];
    f32 light_radius[SFTX_FEAT_LIGHTS];
} sftx_features_state;

struct sftx_scene_state {
    sftx_scene_common scene;
    sftx_room_state room;
    sftx_crowd_state crowd;
    sftx_features_state feat;
};
#define SFTX_MAX_SCENES 64
static struct sftx_scene_state* sftx_slots[SFTX_MAX_SCENES];

/* All state of one scene lives in a slot; the functions below work on the
   current slot through these names. */
struct sftx_scene_state;
static struct sftx_scene_state* sftx_cur;
#define g_scene (sftx_cur->scene)
#define g_room (sftx_cur->room)
#define g_crowd (sftx_cur->crowd)
#define g_feat (sftx_cur->feat)

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

/* camera of frame `frame` of a scene kind's path for a width x height view */
static void sftx_camera_for(i32 kind, i32 frame, i32 width, i32 height, sftvec* pos, sftmat* view, sftmat* proj) {
    if (SFTX_SCENE_PLANE_SPHERES == kind) {
        *pos = SFT_V(0.f, 6.f, -14.f);
        *view = sft_m_view_look_at(*pos, SFT_V(0.f, 0.f, 0.f), SFT_V(0.f, 1.f, 0.f));
        *proj = sft_m_proj(60.f, (f32)width / (f32)height, 0.1f, 1000.f);
        return;
    }
    const f32 th = 2.f * (f32)SFT_PI * (f32)(frame % 120) / 120.f;
    *pos = SFT_V(25.f * cosf(th), 3.f + 4.f * sinf(2.f * th), 25.f * sinf(th));
    const sftvec target = SFT_V(5.f * sinf(th), 0.f, -5.f * cosf(th));
    *view = sft_m_view_look_at(*pos, target, SFT_V(0.f, 1.f, 0.f));
    *proj = sft_m_proj(70.f, (f32)width / (f32)height, 0.1f, 1000.f);
}

/* a frame number below 1000 sets both the pose and the camera; 1000 * (p + 1)
   + c poses the scene at frame p and looks from camera frame c */
static void sftx_split_frame(i32 frame, i32* pose, i32* cam) {
    if (frame >= 1000) {
        *pose = frame / 1000 - 1;
        *cam = frame % 1000;
    }
    else {
        *pose = frame;
        *cam = frame;
    }
}

static void sftx_set_scene_camera(i32 frame) {
    i32 pose, cam;
    sftx_split_frame(frame, &pose, &cam);
    sftvec pos;
    sftmat view, proj;
    sftx_camera_for(g_scene.kind, cam, g_scene.width, g_scene.height, &pos, &view, &proj);
    sft_set_camera(pos, view, proj);
}

static void sftx_plane_spheres_render(i32 frame) {
    sftx_set_scene_camera(frame);

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
            m->lu[k] = ((f32)(lx + i) + 0.5f) / (f32)g_scene.lm_size;
            m->lv[k] = ((f32)(ly + j) + 0.5f) / (f32)g_scene.lm_size;
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
                           const u32* lightmap, const struct sft_bvh_node* nodes, u32 node_count) {
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
    const u32 tc = (u32)tex_count, ts = SFTX_TEX_SIZE, ls = (u32)g_scene.lm_size, one = 1;
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
    const u32 mc = (u32)mesh_count, root = 0;
    sftx_fwrite(f, &mc, 4);
    sftx_fwrite(f, &total, 4);
    sftx_fwrite(f, &node_count, 4);
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
    if (node_count) sftx_fwrite(f, nodes, (u64)node_count * sizeof(struct sft_bvh_node));
    fclose(f);
}

/* ---- 8-wide BVH over triangles (offline builder of the workload; leaf
   payload 0x80000000 | meshlet << 7 | triangle, the layout read at
   sofren2.c:4199-4201) */
typedef struct { f32 mn[3], mx[3], c[3]; u32 payload; } sftx_bitem;

static int sftx_bitem_axis;
static int sftx_bitem_cmp(const void* a, const void* b) {
    const f32 x = ((const sftx_bitem*)a)->c[sftx_bitem_axis], y = ((const sftx_bitem*)b)->c[sftx_bitem_axis];
    if (x < y) return -1;
    if (x > y) return 1;
    return (((const sftx_bitem*)a)->payload < ((const sftx_bitem*)b)->payload) ? -1 : 1;
}

typedef struct { struct sft_bvh_node* n; u32 count, cap; } sftx_bvh;

static u32 sftx_bvh_alloc(sftx_bvh* b) {
    if (b->count == b->cap) {
        b->cap = b->cap ? b->cap * 2 : 1024;
        b->n = (struct sft_bvh_node*)realloc(b->n, b->cap * sizeof(struct sft_bvh_node));
    }
    struct sft_bvh_node* n = &b->n[b->count];
    memset(n, 0, sizeof(*n));
    for (int i = 0; i < 8; i += 1) {
        n->min_x[i] = 1e30f; n->min_y[i] = 1e30f; n->min_z[i] = 1e30f;
        n->max_x[i] = -1e30f; n->max_y[i] = -1e30f; n->max_z[i] = -1e30f;
        n->child_indices[i] = 0xFFFFFFFFu;
    }
    return b->count++;
}

/* splits items into up to 8 groups by three levels of median cuts on the
   longest centroid axis */
static void sftx_bvh_split(sftx_bitem* it, u32 n, int depth, u32* starts, u32* counts, u32* groups, u32 base) {
    if (0 == n) return;
    if (3 == depth || 1 == n) {
        starts[*groups] = base;
        counts[*groups] = n;
        *groups += 1;
        return;
    }
    f32 mn[3] = { 1e30f, 1e30f, 1e30f }, mx[3] = { -1e30f, -1e30f, -1e30f };
    for (u32 i = 0; i < n; i += 1) {
        for (int k = 0; k < 3; k += 1) {
            if (it[i].c[k] < mn[k]) mn[k] = it[i].c[k];
            if (it[i].c[k] > mx[k]) mx[k] = it[i].c[k];
        }
    }
    int axis = 0;
    if (mx[1] - mn[1] > mx[axis] - mn[axis]) axis = 1;
    if (mx[2] - mn[2] > mx[axis] - mn[axis]) axis = 2;
    sftx_bitem_axis = axis;
    qsort(it, n, sizeof(sftx_bitem), sftx_bitem_cmp);
    const u32 h = n / 2;
    sftx_bvh_split(it, h, depth + 1, starts, counts, groups, base);
    sftx_bvh_split(it + h, n - h, depth + 1, starts, counts, groups, base + h);
}

static u32 sftx_bvh_build(sftx_bvh* b, sftx_bitem* it, u32 n, f32 box[6]) {
    const u32 ni = sftx_bvh_alloc(b);
    f32 bmn[3] = { 1e30f, 1e30f, 1e30f }, bmx[3] = { -1e30f, -1e30f, -1e30f };
    if (n <= 8) {
        b->n[ni].is_leaf_group = 1;
        b->n[ni].child_count = (u8)n;
        for (u32 i = 0; i < n; i += 1) {
            struct sft_bvh_node* nd = &b->n[ni];
            nd->min_x[i] = it[i].mn[0]; nd->min_y[i] = it[i].mn[1]; nd->min_z[i] = it[i].mn[2];
            nd->max_x[i] = it[i].mx[0]; nd->max_y[i] = it[i].mx[1]; nd->max_z[i] = it[i].mx[2];
            nd->child_indices[i] = 0x80000000u | it[i].payload;
            for (int k = 0; k < 3; k += 1) {
                if (it[i].mn[k] < bmn[k]) bmn[k] = it[i].mn[k];
                if (it[i].mx[k] > bmx[k]) bmx[k] = it[i].mx[k];
            }
        }
    }
    else {
        u32 starts[8], counts[8], groups = 0;
        sftx_bvh_split(it, n, 0, starts, counts, &groups, 0);
        b->n[ni].is_leaf_group = 0;
        b->n[ni].child_count = (u8)groups;
        for (u32 g = 0; g < groups; g += 1) {
            f32 cb[6];
            const u32 child = sftx_bvh_build(b, it + starts[g], counts[g], cb);
            struct sft_bvh_node* nd = &b->n[ni];
            nd->child_indices[g] = child;
            nd->min_x[g] = cb[0]; nd->min_y[g] = cb[1]; nd->min_z[g] = cb[2];
            nd->max_x[g] = cb[3]; nd->max_y[g] = cb[4]; nd->max_z[g] = cb[5];
            for (int k = 0; k < 3; k += 1) {
                if (cb[k] < bmn[k]) bmn[k] = cb[k];
                if (cb[3 + k] > bmx[k]) bmx[k] = cb[3 + k];
            }
        }
    }
    box[0] = bmn[0]; box[1] = bmn[1]; box[2] = bmn[2];
    box[3] = bmx[0]; box[4] = bmx[1]; box[5] = bmx[2];
    return ni;
}

/* BLAS over every triangle of the meshes (model meshlet order = mesh order) */
static sftx_bvh sftx_build_blas(sftx_mlist* meshes, i32 mesh_count) {
    u32 n = 0;
    for (i32 i = 0; i < mesh_count; i += 1) {
        for (u32 m = 0; m < meshes[i].count; m += 1) n += meshes[i].m[m].tri_count;
    }
    sftx_bitem* it = (sftx_bitem*)malloc((u64)n * sizeof(sftx_bitem) + 1);
    u32 k = 0, gm = 0;
    for (i32 i = 0; i < mesh_count; i += 1) {
        for (u32 m = 0; m < meshes[i].count; m += 1, gm += 1) {
            const sftx_mbuild* b = &meshes[i].m[m];
            for (u32 t = 0; t < b->tri_count; t += 1, k += 1) {
                sftx_bitem* e = &it[k];
                for (int a = 0; a < 3; a += 1) { e->mn[a] = 1e30f; e->mx[a] = -1e30f; }
                for (int v = 0; v < 3; v += 1) {
                    const u32 vi = b->inds[t * 3 + v];
                    const f32 p[3] = { b->x[vi], b->y[vi], b->z[vi] };
                    for (int a = 0; a < 3; a += 1) {
                        if (p[a] < e->mn[a]) e->mn[a] = p[a];
                        if (p[a] > e->mx[a]) e->mx[a] = p[a];
                    }
                }
                for (int a = 0; a < 3; a += 1) e->c[a] = (e->mn[a] + e->mx[a]) * 0.5f;
                e->payload = (gm << 7) | t;
            }
        }
    }
    sftx_bvh bvh = { NULL, 0, 0 };
    f32 box[6];
    sftx_bvh_build(&bvh, it, n, box);
    free(it);
    return bvh;
}

static void sftx_pad_meshlet(sftx_mbuild* m) {
    for (u32 i = m->vert_count; i < 64; i += 1) {
        m->x[i] = m->x[0]; m->y[i] = m->y[0]; m->z[i] = m->z[0];
        m->nx[i] = m->nx[0]; m->ny[i] = m->ny[0]; m->nz[i] = m->nz[0];
        m->u[i] = m->u[0]; m->v[i] = m->v[0]; m->lu[i] = m->lu[0]; m->lv[i] = m->lv[0];
    }
}



static void sftx_box_room_create(int with_blas) {
    sftx_rng rng;
    sftx_rng_seed(&rng, g_scene.seed, 0x2545F491ULL);
    u32* tex[SFTX_TEXTURES];
    for (i32 t = 0; t < SFTX_TEXTURES; t += 1) {
        tex[t] = (u32*)malloc(SFTX_TEX_SIZE * SFTX_TEX_SIZE * 4);
        sftx_value_noise_texture(&rng, tex[t]);
    }
    const i32 L = g_scene.lattice;
    const f32 spacing = 80.f / (f32)L, origin = spacing * (f32)(L - 1) * 0.5f;
    u32* lm = (u32*)malloc((u64)g_scene.lm_size * g_scene.lm_size * 4);
    for (u32 i = 0; i < (u32)g_scene.lm_size * g_scene.lm_size; i += 1) {
        const f32 l = sftx_uniform(&rng, 0.2f, 1.f);
        lm[i] = sftx_pack_rgb(l, l, l);
    }
    sftx_mlist meshes[SFTX_TEXTURES];
    memset(meshes, 0, sizeof(meshes));
    i32 cell = 0;   /* 4x4-texel lightmap cells, 256 per row */
    const f32 ln_lo = logf(0.1f), ln_hi = logf(4.f);
    u32 tris = 0;
    for (i32 bz = 0; bz < L; bz += 1) {
        for (i32 by = 0; by < L; by += 1) {
            for (i32 bx = 0; bx < L; bx += 1) {
                const sftvec c = SFT_V((f32)bx * spacing - origin + sftx_uniform(&rng, -1.f, 1.f),
                                       (f32)by * spacing - origin + sftx_uniform(&rng, -1.f, 1.f),
                                       (f32)bz * spacing - origin + sftx_uniform(&rng, -1.f, 1.f));
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
    sftx_bvh blas = { NULL, 0, 0 };
    if (with_blas) blas = sftx_build_blas(meshes, SFTX_TEXTURES);
    sftx_write_qsf(path, meshes, SFTX_TEXTURES, tex, SFTX_TEXTURES, lm, blas.n, blas.count);
    free(blas.n);
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
    sftx_set_scene_camera(frame);
    sft_set_ambient_light(0.3f, 0.3f, 0.3f);
    for (int i = 0; i < 8; i += 1) {
        sft_draw_point_light(g_room.light_pos[i], g_room.light_col[i].x, g_room.light_col[i].y, g_room.light_col[i].z,
                             1.f, 20.f, (u32)(i + 1));
    }
    sft_draw_instance(&g_room.inst);
}

/* ---------------- workload 2: box room + 32 skinned humanoids + soft shadows
   Each humanoid is a 96-band, 32-segment tube (6144 triangles, 96 meshlets of
   64 vertices) around a 10-joint vertical chain.  Per vertex, Dirichlet(0.5)
   weights over the six joints nearest its height keep their top four,
   renormalised and quantised as w*255+0.5 (the reference's import rule,
   sofren2.c:7407-7410).  Joint rotations are sinusoids of the frame with
   random phase.  Ten capsule proxies per character follow the bones.  The
   room's QSF2 carries a triangle BLAS for the proxy-occlusion rays; 16
   shadow-casting point lights hang 15-25 units above the crowd, outside every
   proxy cluster. */


/* Gamma(a) for a < 1 via Gamma(a+1) * U^(1/a); Marsaglia-Tsang for a+1 */
static f32 sftx_gauss(sftx_rng* r) {
    const f32 u1 = (sftx_u01(r) + 1e-7f), u2 = sftx_u01(r);
    return sqrtf(-2.f * logf(u1)) * cosf(2.f * (f32)SFT_PI * u2);
}

static f32 sftx_gamma(sftx_rng* r, f32 a) {
    const f32 d = (a + 1.f) - 1.f / 3.f, c = 1.f / sqrtf(9.f * d);
    f32 g;
    for (;;) {
        f32 x, v;
        do {
            x = sftx_gauss(r);
            v = 1.f + c * x;
        } while (v <= 0.f);
        v = v * v * v;
        const f32 u = sftx_u01(r) + 1e-7f;
        if (logf(u) < 0.5f * x * x + d - d * v + d * logf(v)) {
            g = d * v;
            break;
        }
    }
    return g * powf(sftx_u01(r) + 1e-7f, 1.f / a);
}



static sft_model* sftx_build_humanoid(sft_arena* ar, sftx_rng* rng) {
    sft_model* m = (sft_model*)sft_arena_alloc(ar, sizeof(sft_model));
    memset(m, 0, sizeof(*m));
    sft_mesh* mesh = (sft_mesh*)sft_arena_alloc(ar, sizeof(sft_mesh));
    memset(mesh, 0, sizeof(*mesh));
    const u32 mc = SFTX_BANDS;
    mesh->meshlets = (struct sft_meshlet*)sft_arena_alloc(ar, mc * sizeof(struct sft_meshlet));
    mesh->meshlet_attribs = (struct sft_meshlet_attribs*)sft_arena_alloc(ar, mc * sizeof(struct sft_meshlet_attribs));
    mesh->meshlet_skins = (struct sft_meshlet_skin*)sft_arena_alloc(ar, mc * sizeof(struct sft_meshlet_skin));
    mesh->inds = (struct sft_meshlet_inds*)sft_arena_alloc(ar, mc * sizeof(struct sft_meshlet_inds));
    f32** arr[] = { &mesh->bounds_center_x, &mesh->bounds_center_y, &mesh->bounds_center_z, &mesh->bounds_radius,
                    &mesh->bounds_min_x, &mesh->bounds_max_x, &mesh->bounds_min_y, &mesh->bounds_max_y,
                    &mesh->bounds_min_z, &mesh->bounds_max_z,
                    &mesh->cone_axis_x, &mesh->cone_axis_y, &mesh->cone_axis_z, &mesh->cone_cutoff };
    for (u32 i = 0; i < sizeof(arr) / sizeof(arr[0]); i += 1) *arr[i] = (f32*)sft_arena_alloc(ar, ((mc + 7) & ~7u) * sizeof(f32));
    mesh->culling_bone_inds = (i32*)sft_arena_alloc(ar, ((mc + 7) & ~7u) * sizeof(i32));
    mesh->meshlet_count = mc;
    mesh->model_meshlet_offset = 0;
    mesh->lightmap_page = 0xFFFFFFFFu;
    mesh->material = SFT_INVALID_HANDLE;
    const f32 height = SFTX_BONE_LEN * SFTX_BONES;
    for (u32 k = 0; k < mc; k += 1) {
        struct sft_meshlet* mv = &mesh->meshlets[k];
        struct sft_meshlet_attribs* at = &mesh->meshlet_attribs[k];
        struct sft_meshlet_skin* sk = &mesh->meshlet_skins[k];
        memset(at, 0, sizeof(*at));
        memset(sk, 0, sizeof(*sk));
        for (u32 ring = 0; ring < 2; ring += 1) {
            const f32 y = height * (f32)(k + ring) / (f32)SFTX_BANDS;
            const f32 rad = 0.14f + 0.04f * sinf(y * 5.f);
            i32 near_j = (i32)(y / SFTX_BONE_LEN + 0.5f);
            if (near_j > SFTX_BONES - 1) near_j = SFTX_BONES - 1;
            for (u32 sgi = 0; sgi < SFTX_SEGS; sgi += 1) {
                const u32 v = ring * SFTX_SEGS + sgi;
                const f32 th = 2.f * (f32)SFT_PI * (f32)sgi / (f32)SFTX_SEGS;
                mv->x[v] = rad * cosf(th); mv->y[v] = y; mv->z[v] = rad * sinf(th);
                at->nx[v] = cosf(th); at->ny[v] = 0.f; at->nz[v] = sinf(th);
                at->u[v] = (f32)sgi / (f32)SFTX_SEGS; at->v[v] = y / height;
                /* Dirichlet(0.5) over the six nearest joints, top four kept */
                i32 cand[6];
                f32 w[6], sum = 0.f;
                i32 lo = near_j - 2;
                if (lo < 0) lo = 0;
                if (lo > SFTX_BONES - 6) lo = SFTX_BONES - 6;
                for (int c = 0; c < 6; c += 1) {
                    cand[c] = lo + c;
                    w[c] = sftx_gamma(rng, 0.5f);
                    sum += w[c];
                }
                for (int c = 0; c < 6; c += 1) w[c] = w[c] / sum;
                for (int a = 0; a < 4; a += 1) {
                    for (int b2 = a + 1; b2 < 6; b2 += 1) {
                        if (w[b2] > w[a]) {
                            const f32 tw = w[a]; w[a] = w[b2]; w[b2] = tw;
                            const i32 tc = cand[a]; cand[a] = cand[b2]; cand[b2] = tc;
                        }
                    }
                }
                const f32 top = w[0] + w[1] + w[2] + w[3];
                sk->j0[v] = (u16)cand[0]; sk->w0[v] = (u8)(w[0] / top * 255.f + 0.5f);
                sk->j1[v] = (u16)cand[1]; sk->w1[v] = (u8)(w[1] / top * 255.f + 0.5f);
                sk->j2[v] = (u16)cand[2]; sk->w2[v] = (u8)(w[2] / top * 255.f + 0.5f);
                sk->j3[v] = (u16)cand[3]; sk->w3[v] = (u8)(w[3] / top * 255.f + 0.5f);
            }
        }
        /* band quads, front faces clockwise from outside */
        struct sft_meshlet_inds* mi = &mesh->inds[k];
        memset(mi, 0, sizeof(*mi));
        for (u32 sgi = 0; sgi < SFTX_SEGS; sgi += 1) {
            const u8 a = (u8)sgi, b2 = (u8)((sgi + 1) % SFTX_SEGS), c = (u8)(SFTX_SEGS + sgi), d = (u8)(SFTX_SEGS + (sgi + 1) % SFTX_SEGS);
            const u8 tri[2][3] = { { a, c, d }, { a, d, b2 } };
            for (int h = 0; h < 2; h += 1) {
                const u32 t = sgi * 2 + h;
                const u32 off = (t / 8) * 32 + (t % 8);
                mi->inds[off] = tri[h][0];
                mi->inds[off + 8] = tri[h][1];
                mi->inds[off + 16] = tri[h][2];
            }
        }
        const f32 y0 = height * (f32)k / (f32)SFTX_BANDS, y1 = height * (f32)(k + 1) / (f32)SFTX_BANDS;
        mesh->bounds_min_x[k] = -0.2f; mesh->bounds_max_x[k] = 0.2f;
        mesh->bounds_min_y[k] = y0; mesh->bounds_max_y[k] = y1;
        mesh->bounds_min_z[k] = -0.2f; mesh->bounds_max_z[k] = 0.2f;
        mesh->bounds_center_x[k] = 0.f;
        mesh->bounds_center_y[k] = (y0 + y1) * 0.5f;
        mesh->bounds_center_z[k] = 0.f;
        mesh->bounds_radius[k] = 0.6f;   /* generous: the pose can bend the band away from its cull bone */
        mesh->cone_axis_x[k] = 0.f; mesh->cone_axis_y[k] = 0.f; mesh->cone_axis_z[k] = 1.f;
        mesh->cone_cutoff[k] = 2.f;
        i32 cb = (i32)(((y0 + y1) * 0.5f) / SFTX_BONE_LEN);
        mesh->culling_bone_inds[k] = cb > SFTX_BONES - 1 ? SFTX_BONES - 1 : cb;
    }
    m->meshes = mesh;
    m->mesh_count = 1;
    m->total_meshlets = mc;
    m->meshlet_mesh_inds = (u16*)sft_arena_alloc(ar, mc * sizeof(u16));
    memset(m->meshlet_mesh_inds, 0, mc * sizeof(u16));
    m->skeleton.bone_count = SFTX_BONES;
    m->skeleton.bones = (struct sft_bone*)sft_arena_alloc(ar, SFTX_BONES * sizeof(struct sft_bone));
    memset(m->skeleton.bones, 0, SFTX_BONES * sizeof(struct sft_bone));
    m->shadow_proxy_count = SFTX_BONES;
    m->shadow_proxies = (struct sft_shadow_proxy*)sft_arena_alloc(ar, SFTX_BONES * sizeof(struct sft_shadow_proxy));
    for (i32 j = 0; j < SFTX_BONES; j += 1) {
        struct sft_bone* bn = &m->skeleton.bones[j];
        bn->inverse_bind_matrix = sft_m_translate(0.f, -SFTX_BONE_LEN * (f32)j, 0.f);
        bn->local_bind_matrix = sft_m_translate(0.f, j ? SFTX_BONE_LEN : 0.f, 0.f);
        bn->bind_pos = SFT_V(0.f, SFTX_BONE_LEN * (f32)j, 0.f);
        bn->parent_ind = j - 1;
        struct sft_shadow_proxy* p = &m->shadow_proxies[j];
        p->posa = SFT_V(0.f, SFTX_BONE_LEN * (f32)j, 0.f);
        p->posb = SFT_V(0.f, SFTX_BONE_LEN * (f32)(j + 1), 0.f);
        p->radius = 0.17f;
        p->bone_ind = (u32)j;
    }
    return m;
}

static void sftx_crowd_create(void) {
    sftx_box_room_create(1);
    sftx_rng rng;
    sftx_rng_seed(&rng, g_scene.seed, 0x9E3779B9ULL);
    if (g_scene.n_chars > 0) g_crowd.human = sftx_build_humanoid(&g_room.arena, &rng);
    for (i32 c = 0; c < g_scene.n_chars; c += 1) {
        /* 8 x 4 grid around the origin, jittered, rotated about y */
        const f32 x = -6.f + 12.f * (f32)(c % 8) / 7.f + sftx_uniform(&rng, -0.3f, 0.3f);
        const f32 z = -3.f + 6.f * (f32)(c / 8) / 3.f + sftx_uniform(&rng, -0.3f, 0.3f);
        g_crowd.xform[c] = sft_m_mul(sft_m_rot_y(sftx_uniform(&rng, 0.f, 6.2831853f)), sft_m_translate(x, -1.f, z));
        g_crowd.mat[c] = sft_material_create(SFT_INVALID_HANDLE,
            sftx_pack_rgb(sftx_uniform(&rng, 0.3f, 1.f), sftx_uniform(&rng, 0.3f, 1.f), sftx_uniform(&rng, 0.3f, 1.f)), 24.f);
        sft_model* pm = (sft_model*)sft_arena_alloc(&g_room.arena, sizeof(sft_model));
        sft_mesh* pmesh = (sft_mesh*)sft_arena_alloc(&g_room.arena, sizeof(sft_mesh));
        *pm = *g_crowd.human;
        *pmesh = g_crowd.human->meshes[0];
        pmesh->material = g_crowd.mat[c];
        pm->meshes = pmesh;
        g_crowd.person[c] = pm;
        for (i32 j = 0; j < SFTX_BONES; j += 1) {
            g_crowd.phase[c][j][0] = sftx_uniform(&rng, 0.f, 6.2831853f);
            g_crowd.phase[c][j][1] = sftx_uniform(&rng, 0.f, 6.2831853f);
        }
    }
    for (i32 l = 0; l < g_scene.n_lights; l += 1) {
        g_crowd.light_pos[l] = SFT_V(sftx_uniform(&rng, -15.f, 15.f), sftx_uniform(&rng, 15.f, 25.f), sftx_uniform(&rng, -15.f, 15.f));
        g_crowd.light_col[l] = SFT_V(sftx_uniform(&rng, 0.4f, 1.f), sftx_uniform(&rng, 0.4f, 1.f), sftx_uniform(&rng, 0.4f, 1.f));
    }
    sft_set_shadows_enabled(1);
    g_scene.tri_count += (u32)g_scene.n_chars * SFTX_BANDS * SFTX_SEGS * 2;
    g_scene.meshlet_count += (u32)g_scene.n_chars * SFTX_BANDS;
}

/* posed = inverse_bind * global, global_j = local_j * global_(j-1)
   (row vectors, sofren2.c:9102-9104) */
static void sftx_crowd_pose(i32 frame) {
    const f32 t = 2.f * (f32)SFT_PI * (f32)(frame % 120) / 120.f;
    for (i32 c = 0; c < g_scene.n_chars; c += 1) {
        sftmat global = sft_m_identity();
        for (i32 j = 0; j < SFTX_BONES; j += 1) {
            const f32 a = 0.35f * sinf(t * 2.f + g_crowd.phase[c][j][0]);
            const f32 b = 0.25f * sinf(t * 3.f + g_crowd.phase[c][j][1]);
            const sftmat local = sft_m_mul(sft_m_mul(sft_m_rot_z(a), sft_m_rot_x(b)), sft_m_translate(0.f, j ? SFTX_BONE_LEN : 0.f, 0.f));
            global = sft_m_mul(local, global);
            g_crowd.posed[c][j] = sft_m_mul(g_crowd.human->skeleton.bones[j].inverse_bind_matrix, global);
        }
    }
}

static void sftx_crowd_render(i32 frame) {
    i32 pose, cam;
    sftx_split_frame(frame, &pose, &cam);
    sftx_set_scene_camera(frame);
    sft_set_ambient_light(0.3f, 0.3f, 0.3f);
    for (i32 l = 0; l < g_scene.n_lights; l += 1) {
        sft_draw_point_light(g_crowd.light_pos[l], g_crowd.light_col[l].x, g_crowd.light_col[l].y, g_crowd.light_col[l].z,
                             1.f, 40.f, (u32)(l + 1));
    }
    sft_draw_instance(&g_room.inst);
    sftx_crowd_pose(pose);
    for (i32 c = 0; c < g_scene.n_chars; c += 1) {
        sft_instance inst;
        memset(&inst, 0, sizeof(inst));
        inst.model = g_crowd.person[c];
        inst.transform = g_crowd.xform[c];
        inst.posed_bone_matrices = g_crowd.posed[c];
        sft_draw_instance(&inst);
    }
}

/* ---------------- workload 6: instanced BLAS models -----------------------
   The crowd scene on a 6^3 lattice with 8 humanoids and 12 shadow-casting
   lights, its room model (with its BLAS) drawn 40 more times under rigid
   transforms (a yaw and an offset of up to 30 units), so that the frame's
   TLAS groups many instances with a BLAS.  4 humanoids (no BLAS) are drawn
   before the room: they leave zeroed items in the TLAS's item range, as
   the reference numbers items by draw slot (sofren2.c:3630-3660). */
#define SFTX_INST_COPIES 40
static void sftx_instanced_create(void) {
    g_scene.lattice = 6;
    g_scene.lm_size = 1024;
    g_scene.n_chars = 8;
    g_scene.n_lights = 12;
    sftx_crowd_create();
    sftx_rng rng;
    sftx_rng_seed(&rng, g_scene.seed, 0x27D4EB2FULL);
    for (i32 k = 0; k < SFTX_INST_COPIES; k += 1) {
        const f32 yaw = sftx_uniform(&rng, 0.f, 6.2831853f);
        const f32 x = sftx_uniform(&rng, -30.f, 30.f), y = sftx_uniform(&rng, -10.f, 10.f);
        const f32 z = sftx_uniform(&rng, -30.f, 30.f);
        g_crowd.copy_xform[k] = sft_m_mul(sft_m_rot_y(yaw), sft_m_translate(x, y, z));
    }
    const u32 room_tris = g_scene.tri_count - (u32)g_scene.n_chars * SFTX_BANDS * SFTX_SEGS * 2;
    const u32 room_meshlets = g_scene.meshlet_count - (u32)g_scene.n_chars * SFTX_BANDS;
    g_scene.tri_count += room_tris * SFTX_INST_COPIES;
    g_scene.meshlet_count += room_meshlets * SFTX_INST_COPIES;
}

static void sftx_instanced_render(i32 frame) {
    i32 pose, cam;
    sftx_split_frame(frame, &pose, &cam);
    sftx_set_scene_camera(frame);
    sft_set_ambient_light(0.3f, 0.3f, 0.3f);
    for (i32 l = 0; l < g_scene.n_lights; l += 1) {
        sft_draw_point_light(g_crowd.light_pos[l], g_crowd.light_col[l].x, g_crowd.light_col[l].y, g_crowd.light_col[l].z,
                             1.f, 40.f, (u32)(l + 1));
    }
    sftx_crowd_pose(pose);
    for (i32 c = 0; c < g_scene.n_chars; c += 1) {
        if (c == g_scene.n_chars / 2) {
            sft_draw_instance(&g_room.inst);
            for (i32 k = 0; k < SFTX_INST_COPIES; k += 1) {
                sft_instance copy = g_room.inst;
                copy.transform = g_crowd.copy_xform[k];
                sft_draw_instance(&copy);
            }
        }
        sft_instance inst;
        memset(&inst, 0, sizeof(inst));
        inst.model = g_crowd.person[c];
        inst.transform = g_crowd.xform[c];
        inst.posed_bone_matrices = g_crowd.posed[c];
        sft_draw_instance(&inst);
    }
}

/* ---------------- workload 3: served scene variants -----------------------
   The crowd scene with sizes drawn from the seed: a box lattice of L^3 boxes
   with L = round(cbrt(T / 63)) for T log-uniform in [50k, 1M] triangles
   (L in [9, 25]; a 2048^2 lightmap page above L = 16), 0-64 skinned humanoids
   (64 of them and the room are 65 instances: such a scene needs the 7/17
   visibility-ID split, sftx_scene_characters tells the caller) and 0-16
   shadow-casting point lights.  Without humanoids or lights the view is
   lit by the box room's 8 point lights without shadows. */
static void sftx_served_create(void) {
    sftx_rng rng;
    sftx_rng_seed(&rng, g_scene.seed, 0xC2B2AE35ULL);
    const f32 tris = expf(sftx_uniform(&rng, logf(50000.f), logf(1000000.f)));
    i32 L = (i32)(cbrtf(tris / 63.f) + 0.5f);
    L = L < 9 ? 9 : (L > 25 ? 25 : L);
    g_scene.lattice = L;
    g_scene.lm_size = L > 16 ? 2048 : 1024;
    g_scene.n_chars = (i32)(sftx_rng_next(&rng) % (SFTX_MAX_CHARS + 1));
    g_scene.n_lights = (i32)(sftx_rng_next(&rng) % (SFTX_SHADOW_LIGHTS + 1));
    sftx_crowd_create();
}

static void sftx_served_render(i32 frame) {
    if (g_scene.n_chars > 0 && g_scene.n_lights > 0) {
        sft_set_shadows_enabled(1);
        sftx_crowd_render(frame);
        return;
    }
    sft_set_shadows_enabled(0);
    i32 pose, cam;
    sftx_split_frame(frame, &pose, &cam);
    sftx_box_room_render(frame);
    sftx_crowd_pose(pose);
    for (i32 c = 0; c < g_scene.n_chars; c += 1) {
        sft_instance inst;
        memset(&inst, 0, sizeof(inst));
        inst.model = g_crowd.person[c];
        inst.transform = g_crowd.xform[c];
        inst.posed_bone_matrices = g_crowd.posed[c];
        sft_draw_instance(&inst);
    }
}

/* ---------------- workload 4: a QSF2 map file ------------------------------
   The model of the QSF2 file named by the SFTX_QSF environment variable, lit
   as the reference's demo lights it (main.c:213-236, 396-406): ambient 0.35,
   the file's game-data point lights (type, position, colour, intensity,
   attenuation; radius = attenuation), or one directional light when it has
   none.  The camera stands at (0, 2, 0) and turns once around the vertical
   axis over 120 frames. */
static void sftx_map_create(void) {
    const char* path = getenv("SFTX_QSF");
    if (!path) {
        fprintf(stderr, "SFTX_QSF names no map file\n");
        exit(2);
    }
    sft_arena_init(&g_room.arena, 1ULL << 32);
    g_room.model = sft_model_load_qsf(path, &g_room.arena);
    if (!g_room.model) {
        fprintf(stderr, "map %s failed to load\n", path);
        exit(2);
    }
    memset(&g_room.inst, 0, sizeof(g_room.inst));
    g_room.inst.model = g_room.model;
    g_room.inst.transform = sft_m_identity();
    g_room.map_lights = 0;
    FILE* f = fopen(path, "rb");
    if (f) {
        u64 off = 0;
        i32 n = 0;
        if (0 == fseek(f, 8, SEEK_SET) && 1 == fread(&off, 8, 1, f) && 0 == fseek(f, (long)off, SEEK_SET) &&
            1 == fread(&n, 4, 1, f)) {
            for (i32 i = 0; i < n && i < 64; i += 1) {
                struct { u32 type; f32 pos[3], color[3], intensity, attenuation; } l;
                if (1 != fread(&l, sizeof(l), 1, f)) break;
                g_room.map_pos[i] = SFT_V(l.pos[0], l.pos[1], l.pos[2]);
                g_room.map_col[i] = SFT_V(l.color[0], l.color[1], l.color[2]);
                g_room.map_radius[i] = (l.attenuation * 2.f) / 2.f;
                g_room.map_lights = i + 1;
            }
        }
        fclose(f);
    }
    g_scene.tri_count = 0;
    g_scene.meshlet_count = g_room.model->total_meshlets;
}

static void sftx_map_render(i32 frame) {
    i32 pose, cam;
    sftx_split_frame(frame, &pose, &cam);
    const f32 yaw = 2.f * (f32)SFT_PI * (f32)(cam % 120) / 120.f, pitch = -0.15f;
    const sftvec pos = SFT_V(0.f, 2.f, 0.f);
    const sftvec fwd = SFT_V(cosf(pitch) * sinf(yaw), sinf(pitch), cosf(pitch) * cosf(yaw));
    sft_set_camera(pos, sft_m_view_look_at(pos, sft_v_add(pos, fwd), SFT_V(0.f, 1.f, 0.f)),
                   sft_m_proj(70.f, (f32)g_scene.width / (f32)g_scene.height, 0.1f, 1000.f));
    sft_set_ambient_light(0.35f, 0.35f, 0.35f);
    if (g_room.map_lights > 0) {
        for (i32 i = 0; i < g_room.map_lights; i += 1) {
            sft_draw_point_light(g_room.map_pos[i], g_room.map_col[i].x, g_room.map_col[i].y, g_room.map_col[i].z, 1.f,
                                 g_room.map_radius[i], (u32)(i + 1));
        }
    }
    else {
        sft_draw_directional_light(sft_v_norm(SFT_V(1.f, 0.5f, 0.f)), 1.f, 1.f, 1.f, 0.8f);
    }
    sft_draw_instance(&g_room.inst);
}

/* ---------------- workload 5: shading features --------------------------
   A 640x360 scene built to reach the branches of the shading path that the
   other workloads leave out: textures of non-power-of-two sizes (edge clamp
   instead of wrap, sofren2.c:5594-5599; bilinear near the camera, nearest on
   the far floor), materials with negative shininess (unlit, 6416-6423), two
   directional lights (6440-6470), 24 point lights of which 20 overlap the
   middle tiles (the 16-light cut of a tile's list, 5929-5931), and 64 draw
   calls, the most the default ID split holds.  Frames 500-999 (camera frame
   f - 500) add six draws that stand between the camera and the scene and
   must be dropped (8412-8417); the reference itself cannot render those
   frames: its TLAS build walks every counted instance, dropped ones
   included, and reads their never-written slots (3630-3634).  The view does not clear
   (sft_set_clear(0)): uncovered pixels keep the caller's buffer and, with
   MSAA, partly covered ones blend with it (6067-6124, 6732-6743).  Every
   object is dmesh geometry: ground plane, three poster quads, and spheres,
   cubes, inverted cubes, capsules, cylinders and triangle fans.  UVs on
   the non-power-of-two textures stay well above 0, also where the
   reference extrapolates them to block lanes off the triangle, whose
   samples it computes and discards: it indexes before the level for
   negative texel coordinates (a 3-texel-wide texture is 1 texel wide at
   level 1, so u stays above 0.5 there). */
static void sftx_noise_image(sftx_rng* rng, i32 w, i32 h, u32* px) {
    f32 g[3][32 * 32];
    const i32 gs[3] = { 4, 8, 16 };
    for (i32 o = 0; o < 3; o += 1) {
        for (i32 i = 0; i < gs[o] * gs[o]; i += 1) g[o][i] = sftx_u01(rng);
    }
    const f32 tint[3] = { sftx_uniform(rng, 0.3f, 1.f), sftx_uniform(rng, 0.3f, 1.f), sftx_uniform(rng, 0.3f, 1.f) };
    for (i32 y = 0; y < h; y += 1) {
        for (i32 x = 0; x < w; x += 1) {
            f32 n = 0.f, amp = 0.5f;
            for (i32 o = 0; o < 3; o += 1) {
                n += amp * sftx_noise_lattice(g[o], gs[o], (f32)x * (f32)gs[o] / (f32)w, (f32)y * (f32)gs[o] / (f32)h);
                amp *= 0.5f;
            }
            n = n / 0.875f;
            px[y * w + x] = sftx_pack_rgb(n * tint[0], n * tint[1], n * tint[2]);
        }
    }
}

static void sftx_features_create(void) {
    sftx_rng rng;
    sftx_rng_seed(&rng, g_scene.seed, 0x85EBCA6BULL);
    sft_set_clear(0);
    /* 0-3 non-power-of-two, 4-5 power-of-two */
    static const i32 tw[6] = { 300, 3, 37, 129, 64, 32 }, th[6] = { 111, 200, 100, 65, 64, 32 };
    for (i32 t = 0; t < 6; t += 1) {
        u32* px = (u32*)malloc((size_t)tw[t] * th[t] * 4);
        sftx_noise_image(&rng, tw[t], th[t], px);
        g_feat.tex[t] = sft_texture_create(tw[t], th[t], px);
        free(px);
    }
    /* floor (NPOT, lit), unlit NPOT poster, shiny NPOT poster, NPOT lit,
       POT lit, POT unlit, untextured lit, untextured unlit */
    g_feat.mat[0] = sft_material_create(g_feat.tex[0], 0xFFFFFFFFu, 16.f);
    g_feat.mat[1] = sft_material_create(g_feat.tex[1], sftx_pack_rgb(0.9f, 0.7f, 0.5f), -1.25f);
    g_feat.mat[2] = sft_material_create(g_feat.tex[2], sftx_pack_rgb(sftx_u01(&rng), sftx_u01(&rng), sftx_u01(&rng)), 48.f);
    g_feat.mat[3] = sft_material_create(g_feat.tex[3], 0xFFE0E0E0u, 4.f);
    g_feat.mat[4] = sft_material_create(g_feat.tex[4], sftx_pack_rgb(sftx_u01(&rng), sftx_u01(&rng), sftx_u01(&rng)), 24.f);
    g_feat.mat[5] = sft_material_create(g_feat.tex[5], 0xFFFFFFFFu, -0.8f);
    g_feat.mat[6] = sft_material_create(SFT_INVALID_HANDLE, sftx_pack_rgb(sftx_u01(&rng), sftx_u01(&rng), sftx_u01(&rng)), 8.f);
    g_feat.mat[7] = sft_material_create(SFT_INVALID_HANDLE, sftx_pack_rgb(0.2f, 0.9f, 0.6f), -1.f);
    for (i32 i = 0; i < SFTX_FEAT_OBJECTS; i += 1) {
        g_feat.obj_kind[i] = i % 6;
        /* primitives have UVs from 0: power-of-two textures or none */
        static const i32 prim_mats[5] = { 4, 5, 6, 7, 4 };
        g_feat.obj_mat[i] = (5 == i % 6) ? ((i / 6) % 2 ? 4 : 6) : prim_mats[(i / 6) % 5];
        const f32 sc = sftx_uniform(&rng, 0.4f, 1.2f);
        const f32 x = sftx_uniform(&rng, -9.f, 9.f), z = sftx_uniform(&rng, -6.f, 10.f);
        g_feat.obj_xform[i] = sft_m_mul(sft_m_mul(sft_m_scale(sc, sc, sc), sft_m_rot_y(sftx_uniform(&rng, 0.f, 6.2831853f))),
                                        sft_m_translate(x, sc * 0.6f, z));
        for (i32 k = 0; k < 4; k += 1) g_feat.obj_param[i][k] = sftx_uniform(&rng, 0.2f, 1.f);
    }
    /* the extra draws stand between the camera and the scene: were they
       not dropped, they would cover much of the frame */
    for (i32 i = SFTX_FEAT_OBJECTS; i < SFTX_FEAT_OBJECTS + SFTX_FEAT_EXTRA; i += 1) {
        g_feat.obj_kind[i] = 0;
        g_feat.obj_mat[i] = 6;
        const f32 x = -5.f + 2.f * (f32)(i - SFTX_FEAT_OBJECTS);
        g_feat.obj_xform[i] = sft_m_mul(sft_m_scale(1.5f, 1.5f, 1.5f), sft_m_translate(x, 2.5f, -6.f));
    }
    for (i32 l = 0; l < SFTX_FEAT_LIGHTS; l += 1) {
        const int clustered = l < 20;
        g_feat.light_pos[l] = clustered
            ? SFT_V(sftx_uniform(&rng, -1.5f, 1.5f), sftx_uniform(&rng, 1.f, 3.f), sftx_uniform(&rng, -1.5f, 1.5f))
            : SFT_V(sftx_uniform(&rng, -9.f, 9.f), sftx_uniform(&rng, 1.f, 4.f), sftx_uniform(&rng, -6.f, 10.f));
        g_feat.light_col[l] = SFT_V(sftx_uniform(&rng, 0.05f, 0.35f), sftx_uniform(&rng, 0.05f, 0.35f), sftx_uniform(&rng, 0.05f, 0.35f));
        g_feat.light_radius[l] = clustered ? sftx_uniform(&rng, 6.f, 12.f) : sftx_uniform(&rng, 3.f, 6.f);
    }
    g_scene.tri_count = 0;
    g_scene.meshlet_count = 0;
}

static void sftx_features_render(i32 frame) {
    const int extra = frame >= 500 && frame < 1000;
    if (extra) frame -= 500;
    i32 pose, cam;
    sftx_split_frame(frame, &pose, &cam);
    /* a front arc of +-0.6 rad: the posters never turn edge-on */
    const f32 th = 0.6f * sinf(2.f * (f32)SFT_PI * (f32)(cam % 120) / 120.f);
    const sftvec pos = SFT_V(13.f * sinf(th), 5.f, -13.f * cosf(th));
    sft_set_camera(pos, sft_m_view_look_at(pos, SFT_V(0.f, 1.f, 0.f), SFT_V(0.f, 1.f, 0.f)),
                   sft_m_proj(60.f, (f32)g_scene.width / (f32)g_scene.height, 0.1f, 1000.f));
    sft_set_ambient_light(0.15f, 0.15f, 0.18f);
    sft_draw_directional_light(SFT_V(1.f, -1.f, 0.5f), 0.6f, 0.5f, 0.4f, 1.f);
    sft_draw_directional_light(SFT_V(-0.3f, -1.f, -0.8f), 0.2f, 0.3f, 0.5f, 0.9f);
    for (i32 l = 0; l < SFTX_FEAT_LIGHTS; l += 1) {
        sft_draw_point_light(g_feat.light_pos[l], g_feat.light_col[l].x, g_feat.light_col[l].y, g_feat.light_col[l].z, 1.f,
                             g_feat.light_radius[l], (u32)(l + 1));
    }
    /* floor: 20x16 unit quads, UV continuous over the plane in [0.4, 1.6] x
       [0.35, 1.55] (clamped past 1 on the 300x111 texture) */
    sft_dmesh* fl = sft_dmesh_begin(SFT_DMESH_QUADS, g_feat.mat[0], 320);
    for (i32 z = 0; z < 16; z += 1) {
        for (i32 x = 0; x < 20; x += 1) {
            const f32 x0 = (f32)(x - 10), x1 = (f32)(x - 9), z0 = (f32)(z - 6), z1 = (f32)(z - 5);
            sft_dmesh_push_quad(fl, SFT_V(x0, 0.f, z0), SFT_V(x0, 0.f, z1), SFT_V(x1, 0.f, z1), SFT_V(x1, 0.f, z0),
                                0.4f + 0.06f * (f32)x, 0.35f + 0.075f * (f32)z, 0.4f + 0.06f * (f32)(x + 1),
                                0.35f + 0.075f * (f32)(z + 1), SFT_V(0.f, 1.f, 0.f));
        }
    }
    sft_dmesh_end(fl, NULL);
    /* posters: unlit NPOT, shiny NPOT through a transient clone of its
       material, and a 4-normal quad */
    sft_dmesh* p0 = sft_dmesh_begin(SFT_DMESH_QUADS, g_feat.mat[1], 1);
    sft_dmesh_push_quad(p0, SFT_V(-7.f, 0.5f, 6.f), SFT_V(-7.f, 4.5f, 6.f), SFT_V(-3.f, 4.5f, 7.f), SFT_V(-3.f, 0.5f, 7.f),
                        0.6f, 0.3f, 1.4f, 1.2f, SFT_V(0.24f, 0.f, -0.97f));
    sft_dmesh_end(p0, NULL);
    sft_dmesh* p1 = sft_dmesh_begin(SFT_DMESH_QUADS, sft_material_clone(g_feat.mat[2]), 1);
    sft_dmesh_push_quad(p1, SFT_V(2.f, 0.2f, 7.f), SFT_V(2.f, 3.8f, 7.f), SFT_V(6.5f, 3.8f, 5.5f), SFT_V(6.5f, 0.2f, 5.5f),
                        0.35f, 0.3f, 2.2f, 1.3f, SFT_V(-0.32f, 0.f, -0.95f));
    sft_dmesh_end(p1, NULL);
    sft_dmesh* p2 = sft_dmesh_begin(SFT_DMESH_QUADS, g_feat.mat[3], 1);
    sft_dmesh_push_quad_4n(p2, SFT_V(-2.f, 0.05f, 2.f), SFT_V(-2.f, 2.5f, 3.f), SFT_V(1.5f, 2.5f, 3.f), SFT_V(1.5f, 0.05f, 2.f),
                           0.45f, 0.5f, 1.1f, 1.3f, sft_v_norm(SFT_V(-0.5f, 0.2f, -1.f)), sft_v_norm(SFT_V(-0.4f, 0.6f, -1.f)),
                           sft_v_norm(SFT_V(0.5f, 0.6f, -1.f)), sft_v_norm(SFT_V(0.4f, 0.2f, -1.f)));
    sft_dmesh_end(p2, NULL);
    for (i32 i = 0; i < SFTX_FEAT_OBJECTS + (extra ? SFTX_FEAT_EXTRA : 0); i += 1) {
        const f32* q = g_feat.obj_param[i];
        const sft_mat_handle mat = g_feat.mat[g_feat.obj_mat[i]];
        sft_dmesh* dm;
        switch (g_feat.obj_kind[i]) {
            case 0:
                dm = sft_dmesh_begin(SFT_DMESH_QUADS, mat, 7 * 10);
                sft_draw_sphere(dm, 6, 10);
                break;
            case 1:
                dm = sft_dmesh_begin(SFT_DMESH_QUADS, mat, 6);
                sft_draw_cube(dm);
                break;
            case 2:
                dm = sft_dmesh_begin(SFT_DMESH_QUADS, mat, 6);
                sft_draw_cube_inv(dm);
                break;
            case 3:
                dm = sft_dmesh_begin(SFT_DMESH_QUADS, mat, 256);
                sft_draw_capsule(dm, SFT_V(0.f, -0.3f * q[0], 0.f), SFT_V(0.2f * q[1], 0.5f * q[2], 0.f), 0.25f + 0.2f * q[3], 8, 4);
                break;
            case 4:
                dm = sft_dmesh_begin(SFT_DMESH_QUADS, mat, 64);
                sft_draw_cylinder(dm, SFT_V(0.f, -0.5f, 0.f), SFT_V(0.f, 0.4f + 0.4f * q[0], 0.f), 0.3f + 0.3f * q[1],
                                  0.1f + 0.3f * q[2], 10);
                break;
            default: {
                /* a fan of 12 triangles (a power-of-two texture or none: far,
                   small fans extrapolate UVs below 0 in lanes the reference
                   computes but does not store) */
                dm = sft_dmesh_begin(SFT_DMESH_TRIANGLES, mat, 12);
                for (i32 k = 0; k < 12; k += 1) {
                    const f32 a0 = 6.2831853f * (f32)k / 12.f, a1 = 6.2831853f * (f32)(k + 1) / 12.f;
                    const sftvec c = SFT_V(0.f, 0.2f + 0.3f * q[0], 0.f);
                    const sftvec e0 = SFT_V(0.8f * cosf(a0), 0.f, 0.8f * sinf(a0));
                    const sftvec e1 = SFT_V(0.8f * cosf(a1), 0.f, 0.8f * sinf(a1));
                    sft_dmesh_push_tri(dm, c, e1, e0, 0.9f, 0.9f, 0.9f + 0.6f * cosf(a1), 0.9f + 0.6f * sinf(a1),
                                       0.9f + 0.6f * cosf(a0), 0.9f + 0.6f * sinf(a0), sft_v_face_normal(c, e1, e0));
                }
                break;
            }
        }
        sft_dmesh_end(dm, &g_feat.obj_xform[i]);
    }
}

/* ---------------- shared checks of host-side queries ---------------------
   Compiled into every build of the API, so the results of two builds can
   be compared word for word. */

/* A seeded skeleton of 12 bones (parents anywhere before the child, bone 0
   and one more a root) with two animations: channels of every path with
   linear and step samplers of 1-6 keys.  For each of the n times, samples
   animation (k % 2) with sft_pose_update and writes the bones' global
   matrices then their posed matrices (2 * 12 * 16 floats per time).
   Returns the floats written. */
i32 sftx_pose_check(u32 seed, i32 n_times, const f32* times, f32* out) {
    enum { NB = 12, NCH = 14 };
    sftx_rng rng;
    sftx_rng_seed(&rng, seed, 0x27D4EB2FULL);
    sft_arena ar;
    sft_arena_init(&ar, 1ULL << 24);
    sft_model* m = (sft_model*)sft_arena_alloc(&ar, sizeof(sft_model));
    memset(m, 0, sizeof(*m));
    m->skeleton.bone_count = NB;
    m->skeleton.bones = (struct sft_bone*)sft_arena_alloc(&ar, NB * sizeof(struct sft_bone));
    memset(m->skeleton.bones, 0, NB * sizeof(struct sft_bone));
    for (i32 i = 0; i < NB; i += 1) {
        struct sft_bone* b = &m->skeleton.bones[i];
        b->parent_ind = (0 == i || 5 == i) ? -1 : (i32)(sftx_rng_next(&rng) % (u32)i);
        b->base_t = SFT_V(sftx_uniform(&rng, -1.f, 1.f), sftx_uniform(&rng, 0.f, 1.f), sftx_uniform(&rng, -1.f, 1.f));
        const sftvec ax = sft_v_norm(SFT_V(sftx_uniform(&rng, -1.f, 1.f), sftx_uniform(&rng, -1.f, 1.f), 0.5f));
        const f32 ang = sftx_uniform(&rng, -1.5f, 1.5f);
        b->base_r.x = ax.x * sinf(ang * 0.5f);
        b->base_r.y = ax.y * sinf(ang * 0.5f);
        b->base_r.z = ax.z * sinf(ang * 0.5f);
        b->base_r.w = cosf(ang * 0.5f);
        b->base_s = SFT_V(sftx_uniform(&rng, 0.8f, 1.2f), sftx_uniform(&rng, 0.8f, 1.2f), sftx_uniform(&rng, 0.8f, 1.2f));
        b->local_bind_matrix = sft_m_mul(sft_m_scale(b->base_s.x, b->base_s.y, b->base_s.z),
                                         sft_m_mul(sft_m_from_quat(b->base_r), sft_m_translate(b->base_t.x, b->base_t.y, b->base_t.z)));
        b->inverse_bind_matrix = sft_m_qinv(sft_m_mul(sft_m_rot_y(sftx_uniform(&rng, 0.f, 6.f)),
                                                      sft_m_translate(sftx_uniform(&rng, -2.f, 2.f), sftx_uniform(&rng, 0.f, 2.f), 0.f)));
    }
    m->anim_count = 2;
    m->animations = (struct sft_animation*)sft_arena_alloc(&ar, 2 * sizeof(struct sft_animation));
    for (i32 a = 0; a < 2; a += 1) {
        struct sft_animation* an = &m->animations[a];
        an->channel_count = NCH;
        an->sampler_count = NCH;
        an->channels = (struct sft_anim_channel*)sft_arena_alloc(&ar, NCH * sizeof(struct sft_anim_channel));
        an->samplers = (struct sft_anim_sampler*)sft_arena_alloc(&ar, NCH * sizeof(struct sft_anim_sampler));
        an->duration = 0.f;
        for (i32 c = 0; c < NCH; c += 1) {
            struct sft_anim_channel* ch = &an->channels[c];
            struct sft_anim_sampler* sm = &an->samplers[c];
            ch->sampler_ind = c;
            /* one channel names a bone the skeleton lacks */
            ch->node_ind = (NCH - 1 == c) ? NB + 3 : (i32)(sftx_rng_next(&rng) % NB);
            ch->path = (enum sft_anim_path)(sftx_rng_next(&rng) % 3);
            sm->count = 1 + (i32)(sftx_rng_next(&rng) % 6);
            sm->is_step = (u8)(0 == sftx_rng_next(&rng) % 3);
            sm->is_cubic = 0;
            sm->inputs = (f32*)sft_arena_alloc(&ar, (u64)sm->count * sizeof(f32));
            sm->outputs = (f32*)sft_arena_alloc(&ar, (u64)sm->count * 4 * sizeof(f32));
            f32 t = sftx_uniform(&rng, 0.f, 0.3f);
            for (i32 k = 0; k < sm->count; k += 1) {
                sm->inputs[k] = t;
                t += sftx_uniform(&rng, 0.05f, 0.6f);
                f32* o = &sm->outputs[k * 4];
                if (SFT_ANIM_PATH_ROTATION == ch->path) {
                    const sftvec ax = sft_v_norm(SFT_V(sftx_uniform(&rng, -1.f, 1.f), sftx_uniform(&rng, -1.f, 1.f),
                                                       sftx_uniform(&rng, -1.f, 1.f)));
                    const f32 ang = sftx_uniform(&rng, -3.f, 3.f);
                    o[0] = ax.x * sinf(ang * 0.5f); o[1] = ax.y * sinf(ang * 0.5f); o[2] = ax.z * sinf(ang * 0.5f);
                    o[3] = cosf(ang * 0.5f);
                }
                else if (SFT_ANIM_PATH_SCALE == ch->path) {
                    o[0] = sftx_uniform(&rng, 0.5f, 1.5f); o[1] = sftx_uniform(&rng, 0.5f, 1.5f);
                    o[2] = sftx_uniform(&rng, 0.5f, 1.5f); o[3] = 0.f;
                }
                else {
                    o[0] = sftx_uniform(&rng, -1.f, 1.f); o[1] = sftx_uniform(&rng, -1.f, 1.f);
                    o[2] = sftx_uniform(&rng, -1.f, 1.f); o[3] = 0.f;
                }
            }
            if (sm->inputs[sm->count - 1] > an->duration) an->duration = sm->inputs[sm->count - 1];
        }
    }
    sft_pose* pose = sft_pose_create(m, &ar);
    i32 n = 0;
    for (i32 k = 0; k < n_times && pose; k += 1) {
        sft_pose_update(pose, k % 2, times[k]);
        memcpy(out + n, pose->global_matrices, NB * sizeof(sftmat));
        n += NB * 16;
        memcpy(out + n, pose->posed_bone_matrices, NB * sizeof(sftmat));
        n += NB * 16;
    }
    munmap(ar.base, ar.max_size);
    return n;
}

/* n seeded rays against the current scene's room model (it needs a BLAS:
   the crowd and served workloads and map files have one) placed by a
   rotated, translated instance transform; per ray 8 floats: hit, distance,
   position, normal.  Returns the rays cast (0 without a room model). */
i32 sftx_raycast_check(u32 seed, i32 n_rays, f32* out) {
    if (!sftx_cur || !g_room.model) return 0;
    sftx_rng rng;
    sftx_rng_seed(&rng, seed, 0x165667B1ULL);
    sft_instance inst = g_room.inst;
    inst.transform = sft_m_mul(sft_m_rot_y(0.7f), sft_m_translate(3.f, -1.f, 2.f));
    for (i32 i = 0; i < n_rays; i += 1) {
        const sftvec o = SFT_V(sftx_uniform(&rng, -30.f, 30.f), sftx_uniform(&rng, -2.f, 10.f), sftx_uniform(&rng, -30.f, 30.f));
        sftvec d = SFT_V(sftx_uniform(&rng, -1.f, 1.f), sftx_uniform(&rng, -1.f, 1.f), sftx_uniform(&rng, -1.f, 1.f));
        if (0 == i % 7) d = SFT_V(0.f, -1.f, 0.f);  /* axis-aligned: the epsilon guard of 1/dir */
        const f32 maxd = (0 == i % 5) ? sftx_uniform(&rng, 1.f, 20.f) : 1000.f;
        const sft_ray_hit h = sft_raycast_instance(&inst, o, d, maxd);
        f32* r = out + (size_t)i * 8;
        r[0] = (f32)h.hit; r[1] = h.distance;
        r[2] = h.pos.x; r[3] = h.pos.y; r[4] = h.pos.z;
        r[5] = h.normal.x; r[6] = h.normal.y; r[7] = h.normal.z;
    }
    return n_rays;
}

/* ---------------- shared check of the offline asset builders -------------
   The builders are additive entry points of the GPU library; the reference
   build provides the same names as thin calls of its own static builders
   (sofren2.c:2526, 3526, 3511), so these prototypes serve both. */
void sftx_mesh_build(sft_mesh* out_mesh, sft_arena* arena, const f32* vx, const f32* vy, const f32* vz,
                     const i32* inds, i32 tri_count);
void sftx_blas_build(sft_model* model, sft_arena* arena);
void sftx_shadow_proxies_build(sft_model* model, sft_arena* arena, i32 budget);

typedef struct {
    f32* x; f32* y; f32* z;
    i32* inds;
    i32 nv, nt, cap_v, cap_t;
} sftx_soup;

static i32 sftx_soup_vert(sftx_soup* s, f32 x, f32 y, f32 z) {
    if (s->nv == s->cap_v) return 0;
    s->x[s->nv] = x; s->y[s->nv] = y; s->z[s->nv] = z;
    return s->nv++;
}

static void sftx_soup_tri(sftx_soup* s, i32 a, i32 b, i32 c) {
    if (s->nt == s->cap_t) return;
    s->inds[s->nt * 3] = a; s->inds[s->nt * 3 + 1] = b; s->inds[s->nt * 3 + 2] = c;
    s->nt += 1;
}

/* Triangle soups that reach the builders' branches:
   0: a jittered closed sphere (32 x 48), shared vertices, curved normals;
   1: 300 boxes of log-uniform size under random rotations about y, eight
      shared corners each (disconnected parts: the candidate list runs dry);
   2: a 60 x 60 height field with spikes (the 100-degree cone limit rejects);
   3: random triangles over a pool of 900 vertices, with repeated corners and
      collinear points (degenerate normals and sphere fall-backs). */
static void sftx_make_soup(sftx_soup* s, i32 kind, sftx_rng* r) {
    if (0 == kind) {
        const i32 rings = 32, segs = 48;
        for (i32 i = 0; i <= rings; i += 1) {
            for (i32 j = 0; j < segs; j += 1) {
                const f32 th = (f32)SFT_PI * (f32)i / (f32)rings, ph = 2.f * (f32)SFT_PI * (f32)j / (f32)segs;
                const f32 rr = 2.f + sftx_uniform(r, -0.05f, 0.05f);
                sftx_soup_vert(s, rr * sinf(th) * cosf(ph), rr * cosf(th), rr * sinf(th) * sinf(ph));
            }
        }
        for (i32 i = 0; i < rings; i += 1) {
            for (i32 j = 0; j < segs; j += 1) {
                const i32 a = i * segs + j, b = i * segs + (j + 1) % segs, c = a + segs, d = b + segs;
                sftx_soup_tri(s, a, c, d);
                sftx_soup_tri(s, a, d, b);
            }
        }
    }
    else if (1 == kind) {
        static const u8 faces[12][3] = { {0,2,3},{0,3,1},{4,5,7},{4,7,6},{0,1,5},{0,5,4},
                                         {2,6,7},{2,7,3},{0,4,6},{0,6,2},{1,3,7},{1,7,5} };
        for (i32 b = 0; b < 300; b += 1) {
            const f32 cx = sftx_uniform(r, -20.f, 20.f), cy = sftx_uniform(r, 0.f, 8.f), cz = sftx_uniform(r, -20.f, 20.f);
            const f32 hx = 0.05f * expf(sftx_uniform(r, 0.f, 3.7f)), hy = 0.05f * expf(sftx_uniform(r, 0.f, 3.7f));
            const f32 hz = 0.05f * expf(sftx_uniform(r, 0.f, 3.7f)), ang = sftx_uniform(r, 0.f, 6.28f);
            const f32 ca = cosf(ang), sa = sinf(ang);
            i32 base = 0;
            for (i32 k = 0; k < 8; k += 1) {
                const f32 lx = (k & 1) ? hx : -hx, ly = (k & 2) ? hy : -hy, lz = (k & 4) ? hz : -hz;
                const i32 v = sftx_soup_vert(s, cx + lx * ca - lz * sa, cy + ly, cz + lx * sa + lz * ca);
                if (0 == k) base = v;
            }
            for (i32 f = 0; f < 12; f += 1) sftx_soup_tri(s, base + faces[f][0], base + faces[f][1], base + faces[f][2]);
        }
    }
    else if (2 == kind) {
        const i32 n = 60;
        for (i32 i = 0; i <= n; i += 1) {
            for (i32 j = 0; j <= n; j += 1) {
                f32 h = 0.3f * sinf(0.4f * (f32)i) * cosf(0.3f * (f32)j) + sftx_uniform(r, -0.02f, 0.02f);
                if (0 == sftx_rng_next(r) % 23) h += sftx_uniform(r, -3.f, 3.f);
                sftx_soup_vert(s, (f32)j * 0.5f, h, (f32)i * 0.5f);
            }
        }
        for (i32 i = 0; i < n; i += 1) {
            for (i32 j = 0; j < n; j += 1) {
                const i32 a = i * (n + 1) + j, b = a + 1, c = a + n + 1, d = c + 1;
                sftx_soup_tri(s, a, c, d);
                sftx_soup_tri(s, a, d, b);
            }
        }
    }
    else {
        for (i32 v = 0; v < 900; v += 1) {
            if (v >= 3 && 0 == v % 17) {
                /* on the line through the previous two */
                const f32 t = sftx_uniform(r, -1.f, 2.f);
                sftx_soup_vert(s, s->x[v - 2] + t * (s->x[v - 1] - s->x[v - 2]), s->y[v - 2] + t * (s->y[v - 1] - s->y[v - 2]),
                               s->z[v - 2] + t * (s->z[v - 1] - s->z[v - 2]));
            }
            else {
                sftx_soup_vert(s, sftx_uniform(r, -5.f, 5.f), sftx_uniform(r, -5.f, 5.f), sftx_uniform(r, -5.f, 5.f));
            }
        }
        for (i32 t = 0; t < 2000; t += 1) {
            const i32 a = (i32)(sftx_rng_next(r) % 900u);
            i32 b = (i32)(sftx_rng_next(r) % 900u), c = (i32)(sftx_rng_next(r) % 900u);
            if (0 == t % 29) b = a;                       /* repeated corner */
            if (0 == t % 31) { b = a; c = a; }            /* a point */
            if (0 == t % 13 && a >= 2) { b = a - 1; c = a - 2; }
            sftx_soup_tri(s, a, b, c);
        }
    }
}

static u64 sftx_put(u8* out, u64 at, u64 cap, const void* p, u64 n) {
    if (at + n <= cap) memcpy(out + at, p, n);
    return at + n;
}

/* every BLAS node (fields only: the bytes after is_leaf_group are zero) and
   every proxy of the model */
static u64 sftx_put_blas_proxies(const sft_model* m, u8* out, u64 at, u64 cap) {
    const u32 head[3] = { m->blas_node_count, m->blas_root_ind, m->shadow_proxy_count };
    at = sftx_put(out, at, cap, head, sizeof(head));
    for (u32 i = 0; i < m->blas_node_count; i += 1) {
        u8 node[sizeof(struct sft_bvh_node)];
        memset(node, 0, sizeof(node));
        memcpy(node, &m->blas_nodes[i], offsetof(struct sft_bvh_node, is_leaf_group) + 1);
        at = sftx_put(out, at, cap, node, sizeof(node));
    }
    return sftx_put(out, at, cap, m->shadow_proxies, (u64)m->shadow_proxy_count * sizeof(struct sft_shadow_proxy));
}

/* Runs the builders on a seeded input and writes what they produced.
   Kinds 0-3: the soups above through sftx_mesh_build (rand() seeded with
   `seed` first), the one-mesh model's BLAS and BVH-subtree proxies with
   budgets 1, 8 and 40.  Kind 4: the crowd humanoid's BLAS and bone-cluster
   proxies with budgets 10, 4 and 1.  Kind 5: the same mesh under a seeded
   skeleton with two roots and scattered bind positions (merges into a
   parent's cluster and into the largest one).  Returns the bytes the output
   needs (nothing past `cap` is written). */
i64 sftx_builder_check(u32 seed, i32 kind, u8* out, u64 cap) {
    sftx_rng rng;
    sftx_rng_seed(&rng, seed, 0x61C88647ULL + (u64)kind);
    sft_arena ar;
    sft_arena_init(&ar, 1ULL << 30);
    sft_model* m = NULL;
    u64 at = 0;
    if (kind >= 0 && kind <= 3) {
        sftx_soup s;
        s.cap_v = 20000; s.cap_t = 20000; s.nv = s.nt = 0;
        s.x = (f32*)malloc((u64)s.cap_v * 3 * sizeof(f32));
        s.y = s.x + s.cap_v; s.z = s.y + s.cap_v;
        s.inds = (i32*)malloc((u64)s.cap_t * 3 * sizeof(i32));
        sftx_make_soup(&s, kind, &rng);
        m = (sft_model*)sft_arena_alloc(&ar, sizeof(sft_model));
        memset(m, 0, sizeof(*m));
        sft_mesh* mesh = (sft_mesh*)sft_arena_alloc(&ar, sizeof(sft_mesh));
        memset(mesh, 0, sizeof(*mesh));
        srand(seed);
        sftx_mesh_build(mesh, &ar, s.x, s.y, s.z, s.inds, s.nt);
        free(s.x);
        free(s.inds);
        const u32 mc = mesh->meshlet_count;
        const u32 head[2] = { (u32)s.nt, mc };
        at = sftx_put(out, at, cap, head, sizeof(head));
        for (u32 i = 0; i < mc; i += 1) {
            at = sftx_put(out, at, cap, &mesh->meshlets[i], sizeof(struct sft_meshlet));
            at = sftx_put(out, at, cap, &mesh->inds[i], sizeof(struct sft_meshlet_inds));
            const f32 b[14] = { mesh->bounds_center_x[i], mesh->bounds_center_y[i], mesh->bounds_center_z[i],
                                mesh->bounds_radius[i], mesh->bounds_min_x[i], mesh->bounds_min_y[i],
                                mesh->bounds_min_z[i], mesh->bounds_max_x[i], mesh->bounds_max_y[i],
                                mesh->bounds_max_z[i], mesh->cone_axis_x[i], mesh->cone_axis_y[i],
                                mesh->cone_axis_z[i], mesh->cone_cutoff[i] };
            at = sftx_put(out, at, cap, b, sizeof(b));
        }
        m->meshes = mesh;
        m->mesh_count = 1;
        m->total_meshlets = mc;
        mesh->model_meshlet_offset = 0;
        m->meshlet_mesh_inds = (u16*)sft_arena_alloc(&ar, ((u64)mc + 1) * sizeof(u16));
        memset(m->meshlet_mesh_inds, 0, ((u64)mc + 1) * sizeof(u16));
        sftx_blas_build(m, &ar);
        static const i32 budgets[3] = { 1, 8, 40 };
        for (i32 b = 0; b < 3; b += 1) {
            sftx_shadow_proxies_build(m, &ar, budgets[b]);
            at = sftx_put_blas_proxies(m, out, at, cap);
        }
    }
    else if (4 == kind || 5 == kind) {
        m = sftx_build_humanoid(&ar, &rng);
        if (5 == kind) {
            for (i32 j = 0; j < m->skeleton.bone_count; j += 1) {
                struct sft_bone* bn = &m->skeleton.bones[j];
                bn->parent_ind = (0 == j || 4 == j) ? -1 : (i32)(sftx_rng_next(&rng) % (u32)j);
                bn->bind_pos = SFT_V(sftx_uniform(&rng, -0.3f, 0.3f), SFTX_BONE_LEN * (f32)j + sftx_uniform(&rng, -0.1f, 0.1f),
                                     sftx_uniform(&rng, -0.3f, 0.3f));
            }
        }
        sftx_blas_build(m, &ar);
        static const i32 budgets[3] = { 10, 4, 1 };
        for (i32 b = 0; b < 3; b += 1) {
            sftx_shadow_proxies_build(m, &ar, budgets[b]);
            at = sftx_put_blas_proxies(m, out, at, cap);
        }
    }
    else {
        at = 0;
    }
    munmap(ar.base, ar.max_size);
    return (i64)at;
}

/* ---------------- dispatch ---------------- */
static i32 sftx_build(i32 kind) {
    sft_set_clear(1);
    g_scene.lattice = SFTX_LATTICE;
    g_scene.lm_size = SFTX_LM_SIZE;
    g_scene.n_chars = SFTX_CHARS;
    g_scene.n_lights = SFTX_SHADOW_LIGHTS;
    switch (kind) {
        case SFTX_SCENE_PLANE_SPHERES: sftx_plane_spheres_create(); return 0;
        case SFTX_SCENE_BOX_ROOM: sftx_box_room_create(0); return 0;
        case SFTX_SCENE_CROWD_SHADOWS: sftx_crowd_create(); return 0;
        case SFTX_SCENE_SERVED: sftx_served_create(); return 0;
        case SFTX_SCENE_MAP: sftx_map_create(); return 0;
        case SFTX_SCENE_FEATURES: sftx_features_create(); return 0;
        case SFTX_SCENE_INSTANCED: sftx_instanced_create(); return 0;
        default: return -1;
    }
}

i32 sftx_scene_new(i32 kind, u32 seed, i32 width, i32 height) {
    i32 h = 0;
    while (h < SFTX_MAX_SCENES && sftx_slots[h]) h += 1;
    if (h == SFTX_MAX_SCENES || kind < 0 || kind >= SFTX_SCENE_COUNT) return -1;
    struct sftx_scene_state* st = (struct sftx_scene_state*)calloc(1, sizeof(struct sftx_scene_state));
    sftx_slots[h] = st;
    sftx_cur = st;
    g_scene.kind = kind;
    g_scene.seed = seed;
    g_scene.width = width;
    g_scene.height = height;
    if (0 != sftx_build(kind)) {
        free(st);
        sftx_slots[h] = NULL;
        sftx_cur = NULL;
        return -1;
    }
    return h;
}

i32 sftx_scene_use(i32 handle) {
    if (handle < 0 || handle >= SFTX_MAX_SCENES || !sftx_slots[handle]) return -1;
    sftx_cur = sftx_slots[handle];
    return 0;
}

static void sftx_draw_current(i32 frame) {
    switch (g_scene.kind) {
        case SFTX_SCENE_PLANE_SPHERES: sftx_plane_spheres_render(frame); break;
        case SFTX_SCENE_BOX_ROOM: sftx_box_room_render(frame); break;
        case SFTX_SCENE_CROWD_SHADOWS: sftx_crowd_render(frame); break;
        case SFTX_SCENE_SERVED: sftx_served_render(frame); break;
        case SFTX_SCENE_MAP: sftx_map_render(frame); break;
        case SFTX_SCENE_FEATURES: sftx_features_render(frame); break;
        case SFTX_SCENE_INSTANCED: sftx_instanced_render(frame); break;
        default: break;
    }
}

void sftx_scene_draw(i32 handle, i32 frame, i32 width, i32 height) {
    if (0 != sftx_scene_use(handle)) return;
    const i32 w = g_scene.width, h = g_scene.height;
    g_scene.width = width;
    g_scene.height = height;
    sftx_draw_current(frame);
    g_scene.width = w;
    g_scene.height = h;
}

void sftx_scene_camera(i32 handle, i32 frame, i32 width, i32 height, sftvec* pos, sftmat* view, sftmat* proj) {
    if (handle < 0 || handle >= SFTX_MAX_SCENES || !sftx_slots[handle]) return;
    i32 pose, cam;
    sftx_split_frame(frame, &pose, &cam);
    sftx_camera_for(sftx_slots[handle]->scene.kind, cam, width, height, pos, view, proj);
}

const sft_model* sftx_scene_room_model(void) { return sftx_cur ? g_room.model : NULL; }

i32 sftx_scene_edit_room(i32 mode, f32 dy) {
    if (!sftx_cur || !g_room.model || (mode != 1 && mode != 2)) return -1;
    const sft_model* m = g_room.model;
    i32 moved = 0;
    for (u32 i = 0; i < m->mesh_count; i += 1) {
        sft_mesh* s = &m->meshes[i];
        for (u32 k = 0; k < s->meshlet_count; k += 1) {
            if (2 == mode && (0u == k || k + 1u == s->meshlet_count)) continue;
            for (i32 v = 0; v < SFT_VERTS_PER_MESHLET; v += 1) s->meshlets[k].y[v] += dy;
            s->bounds_center_y[k] += dy;
            s->bounds_min_y[k] += dy;
            s->bounds_max_y[k] += dy;
            moved += 1;
        }
    }
    return moved;
}

void sftx_scene_free(i32 handle) {
    if (handle < 0 || handle >= SFTX_MAX_SCENES || !sftx_slots[handle]) return;
    if (sftx_cur == sftx_slots[handle]) sftx_cur = NULL;
    free(sftx_slots[handle]);
    sftx_slots[handle] = NULL;
}

i32 sftx_scene_create(i32 kind, u32 seed, i32 width, i32 height) {
    for (i32 h = 0; h < SFTX_MAX_SCENES; h += 1) sftx_scene_free(h);
    return sftx_scene_new(kind, seed, width, height) >= 0 ? 0 : -1;
}

void sftx_scene_record(i32 frame, u32* color, i32 pitch_bytes) {
    sft_set_color_buffer(color, pitch_bytes);
    sft_frame_begin();
    if (sftx_cur) sftx_draw_current(frame);
}

void sftx_scene_render(i32 frame, u32* color, i32 pitch_bytes) {
    sftx_scene_record(frame, color, pitch_bytes);
    sft_present();
    sft_frame_end();
}

i32 sftx_scene_characters(i32 handle) {
    if (handle < 0 || handle >= SFTX_MAX_SCENES || !sftx_slots[handle]) return -1;
    return sftx_slots[handle]->scene.n_chars;
}

i32 sftx_scene_shadow_lights(i32 handle) {
    if (handle < 0 || handle >= SFTX_MAX_SCENES || !sftx_slots[handle]) return -1;
    return sftx_slots[handle]->scene.n_lights;
}

u32 sftx_scene_triangle_count(void) { return sftx_cur ? g_scene.tri_count : 0u; }
u32 sftx_scene_meshlet_count(void) { return sftx_cur ? g_scene.meshlet_count : 0u; }

void sftx_scene_destroy(void) {
    for (i32 h = 0; h < SFTX_MAX_SCENES; h += 1) sftx_scene_free(h);
}
