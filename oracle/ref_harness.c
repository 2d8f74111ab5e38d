/*
 * ref_harness.c — builds the reference renderer, unmodified, into a command-line
 * checker and CPU baseline (test infrastructure only; nothing in the product
 * links it).
 *
 * The reference is a single C header; this file includes it with SFT_IMPL from
 * the path the Makefile passes in SFT_REF_SOURCE.  No line of it is changed.
 * Intermediates are captured with preprocessor hooks only:
 *   - the tile resolve call (sofren2.c:6811) is wrapped so each tile's
 *     visibility IDs and 1/w depth are copied out before shading
 *     (sofren2.c:5809 defines the function; a token probe tells the definition
 *     from the call and only the call gets the hook);
 *   - the screen-space meshlet cache is read from the frame state after
 *     sft_present (sofren2.c:8154-8158); culled meshlets are found by
 *     pre-filling the scratch arena region the cache will occupy with 0xFF;
 *   - clipped-triangle records are read from the clip pool, pre-filled the same
 *     way so unused slots are recognisable (sofren2.c:4351-4381).
 *
 * Usage:
 *   sofren2_ref dump  <scene> <seed> <w> <h> <threads> <frame> <outdir> [pad prefill]
 *   sofren2_ref bench <scene> <seed> <w> <h> <threads> <warmup> <frames>
 *   sofren2_ref pose  <seed> <times.bin> <out.bin>
 *   sofren2_ref ray   <scene> <seed> <rays> <out.bin>
 *   sofren2_ref build <seed> <kind> <out.bin>
 * dump writes color.bin (u32 ARGB, h rows of w + pad/4 words), vid.bin /
 * depth.bin (per sample over the tile grid), ss.bin, ssvis.bin, clip.bin and
 * meta.txt into outdir; with prefill != 0 the colour buffer (padding
 * included) starts as sfth_prefill's pattern instead of zeros.  bench prints
 * one JSON line with median / p99 frame time.  pose and ray run the shared
 * host-query checks of the scene generator (sftx_pose_check,
 * sftx_raycast_check) and write their floats; build runs the generator's
 * builder check (sftx_builder_check) on the reference's own static builders,
 * which this file exposes under the names the check calls.
 */
#define _GNU_SOURCE
#include <time.h>
#include <sys/stat.h>
#include <sys/mman.h>
#include <unistd.h>
#include <semaphore.h>
#include <pthread.h>
#include <errno.h>

#define SFT_IMPL
#define SFT_NO_DEBUG

/* ---- token probe: 1 if the first token of the argument is `const` ---- */
#define SFTH_CAT(a, b) SFTH_CAT_(a, b)
#define SFTH_CAT_(a, b) a ## b
#define SFTH_SECOND(a, b, ...) b
#define SFTH_IS_PROBE(...) SFTH_SECOND(__VA_ARGS__, 0, 0)
#define SFTH_PROBE_const ~, 1,
#define SFTH_IS_DECL(first) SFTH_IS_PROBE(SFTH_CAT_(SFTH_PROBE_, first))

/* The definition keeps its body under a new name; the call site captures the
   tile first and then calls it. */
#define sft__resolve_tile(a, b, c, d) SFTH_CAT(SFTH_RESOLVE_, SFTH_IS_DECL(a))(a, b, c, d)
#define SFTH_RESOLVE_1(a, b, c, d) sfth_resolve_tile_impl(a, b, c, d)
#define SFTH_RESOLVE_0(a, b, c, d) (sfth_capture_tile(a, b, c, d), sfth_resolve_tile_impl(a, b, c, d))

static void sfth_capture_tile(const float* depth, const unsigned* vid, int tile_min_x, int tile_min_y);

#include SFT_REF_SOURCE

#undef sft__resolve_tile

/* ---------------- capture state ---------------- */
static struct {
    i32 cols, rows;
    u32* vid;           /* [S][rows*64][cols*64] */
    f32* depth;
    u32 ss_count;
    u8* ss;             /* ss_count * 2304 */
    u8* ss_vis;
    u32 clip_count;
    u8* clips;          /* clip_count * 164 */
    u32 proxy_count;
    u8* proxies;        /* proxy_count * (posa, posb, radius, light_vis[64]) */
} g_cap;

static void sfth_capture_tile(const float* depth, const unsigned* vid, int tile_min_x, int tile_min_y) {
    if (!g_cap.vid) {
        return;
    }
    const i32 gw = g_cap.cols * SFT_TILE_SIZE, gh = g_cap.rows * SFT_TILE_SIZE;
    for (i32 s = 0; s < SFT_AA_SAMPLES; s += 1) {
        for (i32 y = 0; y < SFT_TILE_SIZE; y += 1) {
            for (i32 x = 0; x < SFT_TILE_SIZE; x += 1) {
                const u64 dst = (u64)s * gw * gh + (u64)(tile_min_y + y) * gw + (tile_min_x + x);
                const i32 src = s * SFT_TILE_PIXELS + y * SFT_TILE_SIZE + x;
                g_cap.vid[dst] = vid[src];
                g_cap.depth[dst] = depth[src];
            }
        }
    }
}

/* sft_present with the scratch and clip regions pre-filled so that culled
   meshlets and unused clip slots can be told apart afterwards. */
static void sfth_present(void) {
    const u32 total_ss = (u32)sft_state->cmd_queue.total_ss_meshlets.count;
    const u64 ss_off = (sft_state->scratch_arena.offset + 63ULL) & ~63ULL;
    const u64 ss_bytes = (u64)total_ss * sizeof(struct sft_ss_meshlet);
    if (total_ss > 0) {
        memset(sft_state->scratch_arena.base + ss_off, 0xFF, ss_bytes);
    }
    const u64 clip_cap = 1u << 16;
    memset(sft_state->clip_arena.base, 0xFF, clip_cap * sizeof(struct sft_clipped_triangle));

    g_cap.cols = sft_state->display.tile_cols;
    g_cap.rows = sft_state->display.tile_rows;
    const u64 grid = (u64)g_cap.cols * 64 * g_cap.rows * 64 * SFT_AA_SAMPLES;
    free(g_cap.vid);
    free(g_cap.depth);
    g_cap.vid = (u32*)malloc(grid * sizeof(u32));
    g_cap.depth = (f32*)malloc(grid * sizeof(f32));

    sft_present();

    g_cap.ss_count = total_ss;
    free(g_cap.ss);
    free(g_cap.ss_vis);
    g_cap.ss = (u8*)malloc(ss_bytes + 1);
    g_cap.ss_vis = (u8*)malloc(total_ss + 1);
    if (total_ss > 0) {
        memcpy(g_cap.ss, sft_state->ss_cache, ss_bytes);
    }
    for (u32 i = 0; i < total_ss; i += 1) {
        const u32* w = (const u32*)(g_cap.ss + (u64)i * sizeof(struct sft_ss_meshlet));
        u32 all_ff = 1;
        for (u32 k = 0; k < sizeof(struct sft_ss_meshlet) / 4; k += 1) {
            if (0xFFFFFFFFu != w[k]) { all_ff = 0; break; }
        }
        g_cap.ss_vis[i] = (u8)!all_ff;
    }

    u32 used = 0;
    while (used < clip_cap && 0xFFFFFFFFu != sft_state->clip_cache.pool[used].inst_ind) {
        used += 1;
    }
    g_cap.proxy_count = (u32)sft_state->shadows.proxy_count.count;
    free(g_cap.proxies);
    g_cap.proxies = (u8*)malloc((u64)g_cap.proxy_count * sizeof(struct sft_world_shadow_proxy) + 1);
    memcpy(g_cap.proxies, sft_state->shadows.proxies, (u64)g_cap.proxy_count * sizeof(struct sft_world_shadow_proxy));
    g_cap.clip_count = used;
    free(g_cap.clips);
    g_cap.clips = (u8*)malloc((u64)used * sizeof(struct sft_clipped_triangle) + 1);
    memcpy(g_cap.clips, sft_state->clip_cache.pool, (u64)used * sizeof(struct sft_clipped_triangle));
}

/* The scene generator calls sft_present; route it through the capture. */
#define sft_present sfth_present
#include "scenes.c"
#undef sft_present

/* The builder check calls the additive builder names of the GPU library;
   here they are the reference's static builders, called unchanged. */
void sftx_mesh_build(sft_mesh* out_mesh, sft_arena* arena, const f32* vx, const f32* vy, const f32* vz,
                     const i32* inds, i32 tri_count) {
    sft__mesh_build(out_mesh, arena, vx, vy, vz, inds, tri_count);
}
void sftx_blas_build(sft_model* model, sft_arena* arena) { sft__blas_build_model(model, arena); }
void sftx_shadow_proxies_build(sft_model* model, sft_arena* arena, i32 budget) {
    sft__generate_shadow_proxies(model, arena, budget);
}

/* writes dir/name, or the path `dir` itself when name is empty */
static void sfth_write(const char* dir, const char* name, const void* data, u64 bytes) {
    char path[1024];
    snprintf(path, sizeof(path), name[0] ? "%s/%s" : "%s%s", dir, name);
    FILE* f = fopen(path, "wb");
    if (!f) {
        fprintf(stderr, "cannot write %s\n", path);
        exit(2);
    }
    if (bytes > 0 && fwrite(data, 1, bytes, f) != bytes) {
        fprintf(stderr, "short write %s\n", path);
        exit(2);
    }
    fclose(f);
}

static f64 sfth_now_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (f64)ts.tv_sec * 1000.0 + (f64)ts.tv_nsec * 1e-6;
}

static int sfth_cmp_f64(const void* a, const void* b) {
    const f64 x = *(const f64*)a, y = *(const f64*)b;
    return (x > y) - (x < y);
}

/* the pre-filled colour buffer of a dump: a hash of (x, y, seed) per word */
static u32 sfth_prefill(u32 x, u32 y, u32 seed) {
    return (x * 73856093u ^ y * 19349663u ^ seed * 83492791u) * 2654435761u;
}

static void* sfth_read_all(const char* path, u64* bytes) {
    FILE* f = fopen(path, "rb");
    if (!f) {
        fprintf(stderr, "cannot read %s\n", path);
        exit(2);
    }
    fseek(f, 0, SEEK_END);
    *bytes = (u64)ftell(f);
    fseek(f, 0, SEEK_SET);
    void* p = malloc(*bytes + 1);
    if (*bytes && 1 != fread(p, *bytes, 1, f)) {
        fprintf(stderr, "short read %s\n", path);
        exit(2);
    }
    fclose(f);
    return p;
}

int main(int argc, char** argv) {
    if (argc >= 5 && 0 == strcmp(argv[1], "pose")) {
        u64 bytes = 0;
        const f32* times = (const f32*)sfth_read_all(argv[3], &bytes);
        const i32 n = (i32)(bytes / 4);
        f32* out = (f32*)malloc((u64)(n + 1) * 2 * 12 * 16 * sizeof(f32));
        const i32 got = sftx_pose_check((u32)strtoul(argv[2], NULL, 10), n, times, out);
        sfth_write(argv[4], "", out, (u64)got * 4);
        return 0;
    }
    if (argc >= 5 && 0 == strcmp(argv[1], "build")) {
        const u32 seed = (u32)strtoul(argv[2], NULL, 10);
        const i32 kind = atoi(argv[3]);
        const i64 need = sftx_builder_check(seed, kind, NULL, 0);
        u8* out = (u8*)malloc((u64)need + 1);
        const i64 got = sftx_builder_check(seed, kind, out, (u64)need);
        sfth_write(argv[4], "", out, (u64)got);
        return 0;
    }
    if (argc >= 6 && 0 == strcmp(argv[1], "ray")) {
        const i32 n = atoi(argv[4]);
        sft_init(64, 64, 1);
        if (0 != sftx_scene_create(atoi(argv[2]), 0, 64, 64)) {
            fprintf(stderr, "unknown scene %s\n", argv[2]);
            return 1;
        }
        f32* out = (f32*)calloc((u64)n * 8 + 1, sizeof(f32));
        const i32 got = sftx_raycast_check((u32)strtoul(argv[3], NULL, 10), n, out);
        sfth_write(argv[5], "", out, (u64)got * 8 * 4);
        return 0;
    }
    if (argc < 8) {
        fprintf(stderr, "usage: %s dump|bench scene seed w h threads ...\n", argv[0]);
        return 1;
    }
    const char* mode = argv[1];
    const i32 scene = atoi(argv[2]);
    const u32 seed = (u32)strtoul(argv[3], NULL, 10);
    const i32 w = atoi(argv[4]), h = atoi(argv[5]);
    const i32 threads = atoi(argv[6]);

    sft_init(w, h, threads);
    if (0 != sftx_scene_create(scene, seed, w, h)) {
        fprintf(stderr, "unknown scene %d\n", scene);
        return 1;
    }
    /* dump: optional row padding (bytes) and pre-fill */
    const i32 pad = (0 == strcmp(mode, "dump") && argc >= 11) ? atoi(argv[9]) : 0;
    const u32 prefill = (0 == strcmp(mode, "dump") && argc >= 11) ? (u32)strtoul(argv[10], NULL, 10) : 0u;
    const i32 pitch = w + pad / 4;
    u32* color = (u32*)calloc((u64)pitch * h, sizeof(u32));
    if (prefill) {
        for (i32 y = 0; y < h; y += 1) {
            for (i32 x = 0; x < pitch; x += 1) color[(u64)y * pitch + x] = sfth_prefill((u32)x, (u32)y, prefill);
        }
    }

    if (0 == strcmp(mode, "dump")) {
        if (argc < 9) {
            fprintf(stderr, "dump needs frame and outdir\n");
            return 1;
        }
        const i32 frame = atoi(argv[7]);
        const char* dir = argv[8];
        mkdir(dir, 0755);
        sftx_scene_render(frame, color, pitch * 4);
        const u64 grid = (u64)g_cap.cols * 64 * g_cap.rows * 64 * SFT_AA_SAMPLES;
        sfth_write(dir, "color.bin", color, (u64)pitch * h * 4);
        sfth_write(dir, "vid.bin", g_cap.vid, grid * 4);
        sfth_write(dir, "depth.bin", g_cap.depth, grid * 4);
        sfth_write(dir, "ss.bin", g_cap.ss, (u64)g_cap.ss_count * sizeof(struct sft_ss_meshlet));
        sfth_write(dir, "ssvis.bin", g_cap.ss_vis, g_cap.ss_count);
        sfth_write(dir, "clip.bin", g_cap.clips, (u64)g_cap.clip_count * sizeof(struct sft_clipped_triangle));
        sfth_write(dir, "proxies.bin", g_cap.proxies, (u64)g_cap.proxy_count * sizeof(struct sft_world_shadow_proxy));
        char meta[512];
        const int n = snprintf(meta, sizeof(meta),
            "width %d\nheight %d\ncols %d\nrows %d\nsamples %d\nss_count %u\nclip_count %u\nthreads %d\n",
            w, h, g_cap.cols, g_cap.rows, SFT_AA_SAMPLES, g_cap.ss_count, g_cap.clip_count,
            sft_state->thread_count);
        sfth_write(dir, "meta.txt", meta, (u64)n);
        return 0;
    }

    if (0 == strcmp(mode, "bench")) {
        if (argc < 9) {
            fprintf(stderr, "bench needs warmup and frames\n");
            return 1;
        }
        const i32 warmup = atoi(argv[7]);
        const i32 frames = atoi(argv[8]);
        for (i32 i = 0; i < warmup; i += 1) {
            sftx_scene_render(i, color, w * 4);
        }
        f64* t = (f64*)malloc(sizeof(f64) * (frames > 0 ? frames : 1));
        const f64 t0 = sfth_now_ms();
        for (i32 i = 0; i < frames; i += 1) {
            const f64 a = sfth_now_ms();
            sftx_scene_render(i, color, w * 4);
            t[i] = sfth_now_ms() - a;
        }
        const f64 total = sfth_now_ms() - t0;
        qsort(t, frames, sizeof(f64), sfth_cmp_f64);
        const f64 med = t[frames / 2];
        i32 p99i = (i32)(0.99 * (frames - 1) + 0.5);
        const f64 p99 = t[p99i];
        const f64 fps = frames / (total / 1000.0);
        printf("{\"frames\": %d, \"threads\": %d, \"ms_median\": %.4f, \"ms_p99\": %.4f, \"fps\": %.3f, \"mpix_s\": %.3f, \"tris\": %u}\n",
            frames, sft_state->thread_count, med, p99, fps, fps * w * h / 1e6, sftx_scene_triangle_count());
        return 0;
    }

    fprintf(stderr, "unknown mode %s\n", mode);
    return 1;
}
