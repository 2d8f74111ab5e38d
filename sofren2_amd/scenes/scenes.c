/*
 * scenes.c — seeded synthetic workloads for the sofren2 API (see scenes.h).
 *
 * Only public API calls and the public model structs are used, so this file
 * compiles against any implementation of the API.  It must be compiled with
 * -ffp-contract=off so the host-side float math is identical for every build.
 */
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
        default: return -1;
    }
}

void sftx_scene_record(i32 frame, u32* color, i32 pitch_bytes) {
    sft_set_color_buffer(color, pitch_bytes);
    sft_frame_begin();
    switch (g_scene.kind) {
        case SFTX_SCENE_PLANE_SPHERES: sftx_plane_spheres_render(frame); break;
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
