/*
 * frontend.cpp — host side of the public sofren2 API for the HIP backend.
 *
 * Everything here records the frame (camera, lights, instances, transient
 * dmesh geometry) or manages resources (arenas, textures, materials, models);
 * sft_present hands the recorded frame to the HIP backend (backend.hip), which
 * does all per-pixel and per-vertex work on the GPU.  Behaviour follows the
 * reference API function by function (citations per function); host float math
 * that feeds the GPU (frustum planes, dmesh vertices, mips) keeps the
 * reference's operation order and is compiled with -ffp-contract=off and
 * -fno-builtin so both builds produce the same bits.
 */
#include <sys/mman.h>
#include <math.h>
#include <string.h>
#include "sft_internal.h"

#define SFT_ERR_EXIT(...) do { \
    fprintf(stderr, "fatal sofren2 error in '%s':\n\t", __func__); \
    fprintf(stderr, __VA_ARGS__); \
    exit(1); \
} while (0)
#define SFT_ERR_RET(_r, ...) do { \
    fprintf(stderr, "sofren2 error in '%s':\n\t", __func__); \
    fprintf(stderr, __VA_ARGS__); \
    return _r; \
} while (0)

namespace sft {
Context* g_ctx = nullptr;
}
using namespace sft;

static Context& C() {
    if (!g_ctx) {
        SFT_ERR_EXIT("sofren2 used before sft_init\n");
    }
    return *g_ctx;
}

/* ================================================================ arenas
   mmap-reserved bump allocators with 64-byte alignment (sofren2.c:9458-9554) */
extern "C" void sft_arena_init(sft_arena* arena, u64 reserve_size) {
    arena->max_size = reserve_size;
    arena->committed_size = 0;
    arena->offset = 0;
    arena->commit_lock = 0;
    arena->base = (u8*)mmap(NULL, reserve_size, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
    if (MAP_FAILED == (void*)arena->base) {
        SFT_ERR_EXIT("mmap reserve failed for size %zu\n", (size_t)reserve_size);
    }
    arena->committed_size = reserve_size;
}

extern "C" void* sft_arena_alloc(sft_arena* arena, u64 size) {
    if (!arena) {
        SFT_ERR_RET(NULL, "arena pointer is NULL\n");
    }
    const u64 aligned = (arena->offset + 63ULL) & ~63ULL;
    const u64 next = aligned + size;
    if (next > arena->max_size) {
        SFT_ERR_EXIT("arena out of memory, requested %zu KB, max %zu, offset %zu\n",
            (size_t)(size / 1024), (size_t)(arena->max_size / 1024), (size_t)aligned);
    }
    arena->offset = next;
    return arena->base + aligned;
}

extern "C" void sft_arena_commit_if_needed(sft_arena* arena, u64 required_bytes) {
    (void)arena;
    (void)required_bytes;   /* pages commit on first touch on Linux */
}

extern "C" void sft_arena_frame_begin(sft_arena* arena) {
    arena->offset = 0;
}

/* ================================================================ textures
   4x4-block swizzled mip chains with an integer tent filter
   (sofren2.c:9181-9310).  The host keeps the level-0 pixels and the chain's
   layout; the chain itself is built in HBM by the backend's texture kernels
   when the texture is first made resident. */
static void tex_pool_grow(Context& c) {
    const i32 old_cap = (i32)c.tex.size();
    const i32 new_cap = (0 == old_cap) ? 1024 : old_cap * 2;
    c.tex.resize(new_cap);
    for (i32 i = old_cap; i < new_cap; i += 1) {
        c.tex[i].generation = 1;
        c.tex[i].next_free = i + 1;
    }
    c.tex[new_cap - 1].next_free = c.tex_free_head;
    c.tex_free_head = old_cap;
}

static i32 ilog2(i32 v) { return 31 - __builtin_clz((u32)v); }

extern "C" sft_tex_handle sft_texture_create(i32 width, i32 height, const u32* pixels) {
    Context& c = C();
    if (width <= 0 || height <= 0 || !pixels) {
        return SFT_INVALID_HANDLE;
    }
    const u8 is_pot = (0 == (width & (width - 1))) && (0 == (height & (height - 1)));

    /* level sizes, each padded to a multiple of 4 in both axes */
    i32 lw[14], lh[14];
    i32 levels = 0;
    u64 total = 0;
    {
        i32 cw = width, ch = height;
        while (levels < 14) {
            lw[levels] = cw;
            lh[levels] = ch;
            total += (u64)(((cw + 3) & ~3) * ((ch + 3) & ~3));
            levels += 1;
            if (1 == cw && 1 == ch) break;
            cw = (cw > 1) ? cw / 2 : 1;
            ch = (ch > 1) ? ch / 2 : 1;
        }
    }

    if (-1 == c.tex_free_head) {
        tex_pool_grow(c);
    }
    const i32 t = c.tex_free_head;
    HostTexture& ht = c.tex[t];
    c.tex_free_head = ht.next_free;
    ht.live = true;
    ht.w = width;
    ht.h = height;
    ht.mip_levels = levels;
    ht.is_pot = is_pot;
    ht.level0.assign(pixels, pixels + (u64)width * height);
    ht.total = total;
    ht.uploaded = false;
    u64 off = 0;
    for (i32 l = 0; l < levels; l += 1) {
        const i32 pw = (lw[l] + 3) & ~3, ph = (lh[l] + 3) & ~3;
        ht.mip_offset[l] = (u32)off;
        ht.blocks_w[l] = pw / 4;
        ht.stride_shift[l] = is_pot ? ilog2(pw / 4) : 0;
        off += (u64)pw * ph;
    }
    return (ht.generation << SFT_HANDLE_GEN_SHIFT) | (u32)t;
}

extern "C" sft_tex_handle sft_texture_create_solid(u32 color) {
    return sft_texture_create(1, 1, &color);
}

static HostTexture* tex_lookup(sft_tex_handle h) {
    Context& c = C();
    if (SFT_INVALID_HANDLE == h) return nullptr;
    const u32 i = SFT_HANDLE_INDEX(h);
    if (i >= c.tex.size()) return nullptr;
    return &c.tex[i];
}

extern "C" i32 sft_texture_get_width(sft_tex_handle tex) {
    HostTexture* t = tex_lookup(tex);
    return t ? t->w : 0;
}

extern "C" i32 sft_texture_get_height(sft_tex_handle tex) {
    HostTexture* t = tex_lookup(tex);
    return t ? t->h : 0;
}

extern "C" u32* sft_texture_get_linear_pixels(sft_tex_handle tex) {
    HostTexture* t = tex_lookup(tex);
    if (!t) return NULL;
    u32* out = (u32*)malloc((u64)t->w * t->h * sizeof(u32));
    memcpy(out, t->level0.data(), (u64)t->w * t->h * sizeof(u32));
    return out;
}

extern "C" void sft_texture_release(sft_tex_handle tex) {
    Context& c = C();
    if (SFT_INVALID_HANDLE == tex) return;
    const u32 i = SFT_HANDLE_INDEX(tex);
    const u32 gen = tex >> SFT_HANDLE_GEN_SHIFT;
    if (i >= c.tex.size() || c.tex[i].generation != gen) return;
    HostTexture& t = c.tex[i];
    backend_release_texture(c.backend, i);
    t.level0.clear();
    t.level0.shrink_to_fit();
    t.live = false;
    t.uploaded = false;
    t.generation += 1;
    t.next_free = c.tex_free_head;
    c.tex_free_head = (i32)i;
}

/* ================================================================ materials
   static pool with generation-checked handles, transient pool per frame
   (sofren2.c:9393-9454) */
static void mat_pool_grow(Context& c) {
    const i32 old_cap = (i32)c.mat_gen.size();
    const i32 new_cap = (0 == old_cap) ? 32 : old_cap * 2;
    c.mat_gen.resize(new_cap);
    c.mat_next_free.resize(new_cap);
    c.mat_tex.resize(new_cap, SFT_INVALID_HANDLE);
    c.mat_color.resize(new_cap, 0xFFFFFFFFu);
    c.mat_shininess.resize(new_cap, 0.f);
    for (i32 i = old_cap; i < new_cap; i += 1) {
        c.mat_gen[i] = 1;
        c.mat_next_free[i] = i + 1;
    }
    c.mat_next_free[new_cap - 1] = c.mat_free_head;
    c.mat_free_head = old_cap;
}

extern "C" sft_mat_handle sft_material_create(sft_tex_handle albedo_tex, u32 base_color, f32 shininess) {
    Context& c = C();
    if (-1 == c.mat_free_head) mat_pool_grow(c);
    const i32 i = c.mat_free_head;
    c.mat_free_head = c.mat_next_free[i];
    c.mat_tex[i] = albedo_tex;
    c.mat_color[i] = base_color;
    c.mat_shininess[i] = shininess;
    return (c.mat_gen[i] << SFT_MAT_GEN_SHIFT) | (u32)i;
}

extern "C" void sft_material_release(sft_mat_handle mat) {
    Context& c = C();
    if (SFT_INVALID_HANDLE == mat || SFT_MAT_IS_TRANSIENT(mat)) return;
    const u32 i = SFT_MAT_INDEX(mat);
    const u32 gen = (mat & SFT_MAT_GEN_MASK) >> SFT_MAT_GEN_SHIFT;
    if (i >= c.mat_gen.size() || c.mat_gen[i] != gen) return;
    c.mat_gen[i] += 1;
    c.mat_next_free[i] = c.mat_free_head;
    c.mat_free_head = (i32)i;
}

extern "C" sft_mat_handle sft_material_clone(sft_mat_handle base_mat) {
    Context& c = C();
    const u32 n = (u32)c.tmat_tex.size();
    if (SFT_INVALID_HANDLE != base_mat) {
        const u32 b = SFT_MAT_INDEX(base_mat);
        if (SFT_MAT_IS_TRANSIENT(base_mat)) {
            c.tmat_tex.push_back(c.tmat_tex[b]);
            c.tmat_color.push_back(c.tmat_color[b]);
            c.tmat_shininess.push_back(c.tmat_shininess[b]);
        }
        else {
            c.tmat_tex.push_back(c.mat_tex[b]);
            c.tmat_color.push_back(c.mat_color[b]);
            c.tmat_shininess.push_back(c.mat_shininess[b]);
        }
    }
    else {
        c.tmat_tex.push_back(SFT_INVALID_HANDLE);
        c.tmat_color.push_back(0xFFFFFFFFu);
        c.tmat_shininess.push_back(0.f);
    }
    return SFT_MAT_TRANSIENT_BIT | n;
}

/* ================================================================ lifecycle */
static void set_display(Context& c, i32 width, i32 height) {
    c.width = width;
    c.height = height;
    c.half_width = (f32)width * 0.5f;
    c.half_height = (f32)height * 0.5f;
    c.vp_x = 0.f;
    c.vp_y = 0.f;
    c.tile_cols = (width + SFT_TILE_SIZE - 1) / SFT_TILE_SIZE;
    c.tile_rows = (height + SFT_TILE_SIZE - 1) / SFT_TILE_SIZE;
}

/* sofren2.c:7974-8123 */
extern "C" void sft_init(i32 width, i32 height, i32 requested_threads) {
    (void)requested_threads;   /* all per-frame work runs on the GPU */
    if (g_ctx) {
        SFT_ERR_EXIT("sft_init called twice without sft_release\n");
    }
    if (!view_size_ok(width, height)) {
        SFT_ERR_EXIT("display %dx%d is outside 1..%d pixels a side\n", (int)width, (int)height, (int)kMaxViewSide);
    }
    g_ctx = new Context();
    Context& c = *g_ctx;
    sft_arena_init(&c.scratch_arena, 2ULL * 1024 * 1024 * 1024);
    set_display(c, width, height);

    probe_estimate_tables(&c.est);

    /* 256x256 checkerboard of 32x32 squares and the base material */
    std::vector<u32> checker(256 * 256);
    for (i32 y = 0, i = 0; y < 256; y += 1) {
        for (i32 x = 0; x < 256; x += 1, i += 1) {
            checker[i] = (((x / 32) % 2) ^ ((y / 32) % 2)) ? 0xFFAFFFAFu : 0xFF555555u;
        }
    }
    c.base_tex = sft_texture_create(256, 256, checker.data());
    c.base_mat = sft_material_create(c.base_tex, 0xFFFFFFFFu, 2.f);
    c.lightmap_brightness = 1.5f;
    c.clear_color = 0xFF222222u;

    /* dmesh index templates: 16 quads (tris (0,1,2), (0,2,3)) or 21 tris */
    memset(&c.quad_inds, 0, sizeof(c.quad_inds));
    memset(&c.tri_inds, 0, sizeof(c.tri_inds));
    for (u32 tri = 0; tri < 128; tri += 1) {
        const u32 off = (tri / 8) * 32 + (tri % 8);
        if (tri < 32) {
            const u8 bv = (u8)((tri / 2) * 4);
            c.quad_inds.inds[off + 0] = bv;
            c.quad_inds.inds[off + 8] = (u8)(bv + ((tri % 2) ? 2 : 1));
            c.quad_inds.inds[off + 16] = (u8)(bv + ((tri % 2) ? 3 : 2));
        }
        if (tri < 21) {
            const u8 bv = (u8)(tri * 3);
            c.tri_inds.inds[off + 0] = bv;
            c.tri_inds.inds[off + 8] = (u8)(bv + 1);
            c.tri_inds.inds[off + 16] = (u8)(bv + 2);
        }
    }
}

extern "C" void sft_release(void) {
    if (!g_ctx) return;
    sftx_server_release();
    drain_async_batches(*g_ctx);
    backend_destroy(g_ctx->backend);
    munmap(g_ctx->scratch_arena.base, g_ctx->scratch_arena.max_size);
    delete g_ctx;
    g_ctx = nullptr;
}

extern "C" void sft_resize(i32 width, i32 height) {
    if (!g_ctx) return;
    if (!view_size_ok(width, height)) {
        SFT_ERR_RET(, "display %dx%d is outside 1..%d pixels a side; size unchanged\n", (int)width, (int)height,
                    (int)kMaxViewSide);
    }
    set_display(*g_ctx, width, height);
}

extern "C" void sft_frame_begin(void) {
    Context& c = C();
    sft_arena_frame_begin(&c.scratch_arena);
    c.transient_models.clear();
}

/* sofren2.c:8301-8312 */
extern "C" void sft_frame_end(void) {
    Context& c = C();
    c.instance_count = 0;
    c.instances.clear();
    c.total_ss = 0;
    c.tmat_tex.clear();
    c.tmat_color.clear();
    c.tmat_shininess.clear();
    c.dir_lights.clear();
    c.point_lights.clear();
    c.proxies.clear();
}

/* the lowest frame set among the first `n` that no batch in flight uses */
static int free_frame_set(const Context& c, int n) {
    for (int k = 0; k < n; k += 1) {
        bool used = false;
        for (const Context::AsyncBatch& ab : c.async_batches) used = used || ab.set == k;
        if (!used) return k;
    }
    return 0;
}

/* the frame recorded since sft_frame_begin (or sftx_view_begin) as one view */
void sft::snapshot_view(const Context& c, ViewRecord& v) {
    v.width = c.width;
    v.height = c.height;
    v.tile_cols = c.tile_cols;
    v.tile_rows = c.tile_rows;
    v.samples = c.samples;
    v.half_width = c.half_width;
    v.half_height = c.half_height;
    v.vp_x = c.vp_x;
    v.vp_y = c.vp_y;
    v.color_buffer = c.color_buffer;
    v.color_pitch = c.color_pitch;
    v.vis64 = nullptr;
    v.should_clear = c.should_clear;
    v.clear_color = c.clear_color;
    v.debug_mode = (i32)c.debug_mode;
    v.camera_pos = c.camera_pos;
    v.view_proj = c.view_proj;
    memcpy(v.planes, c.planes, sizeof(v.planes));
    v.ambient = c.ambient;
    v.dir_lights = c.dir_lights;
    v.point_lights = c.point_lights;
    v.lightmap_brightness = c.lightmap_brightness;
    v.shadows_enabled = c.shadows_enabled;
    v.proxies = c.proxies;
    v.instances = c.instances;
    v.dropped = (u32)std::max<i32>(0, c.instance_count - (i32)c.instances.size());
    v.instance_bits = c.instance_bits;
    v.total_ss = c.total_ss;
    v.transient_models = c.transient_models;
    v.tmat_tex = c.tmat_tex;
    v.tmat_color = c.tmat_color;
    v.tmat_shininess = c.tmat_shininess;
    v.capture_flags = c.capture_flags;
}

extern "C" void sft_present(void) {
    Context& c = C();
    if (c.server_running) {
        fprintf(stderr, "sofren2: sft_present while the request server runs; submit requests instead\n");
        return;
    }
    if (c.batch_open) {
        fprintf(stderr, "sofren2: sft_present inside a batch; use sftx_view_end and sftx_batch_submit\n");
        return;
    }
    if (!c.backend) c.backend = backend_create();   /* exits if there is no GPU */
    if (backend_async_present(c.backend) && 0u == c.capture_flags) {
        /* device output: a batch of one in flight, completed later */
        while (c.async_batches.size() >= (size_t)kFrameSets) {
            Context::AsyncBatch& old = c.async_batches.front();
            backend_complete(c.backend, &c, old.views, old.set);
            c.async_batches.pop_front();
        }
        Context::AsyncBatch ab;
        ab.ticket = 0;
        ab.set = free_frame_set(c, kFrameSets);
        ab.views.resize(1);
        /* its inputs are packed by backend_submit before sft_present returns */
        snapshot_view(c, ab.views[0]);
        c.async_batches.push_back(std::move(ab));
        Context::AsyncBatch& cur = c.async_batches.back();
        backend_submit(c.backend, &c, cur.views, cur.set);
        return;
    }
    drain_async_batches(c);
    std::vector<ViewRecord> one(1);
    snapshot_view(c, one[0]);
    backend_render(c.backend, &c, one);
}

/* ================================================================ batches
   Several views recorded with the ordinary calls and rendered by one pass of
   the kernels over all of them.  Transient storage (dmesh geometry, transient
   materials) of every view lives until sftx_batch_end. */
extern "C" void sftx_batch_begin(void) {
    Context& c = C();
    sft_frame_begin();
    c.batch_views.clear();
    c.batch_open = true;
}

extern "C" void sftx_view_begin(i32 width, i32 height, i32 samples) {
    Context& c = C();
    if (!c.batch_open) {
        fprintf(stderr, "sofren2: sftx_view_begin outside sftx_batch_begin/sftx_batch_end\n");
        return;
    }
    c.view_rejected = !view_size_ok(width, height);
    if (c.view_rejected) {
        fprintf(stderr, "sofren2: view %dx%d is outside 1..%d pixels a side; sftx_view_end will drop it\n", (int)width,
                (int)height, (int)kMaxViewSide);
    }
    else {
        set_display(c, width, height);
    }
    sftx_set_msaa(samples);
    c.instance_count = 0;
    c.instances.clear();
    c.total_ss = 0;
    c.dir_lights.clear();
    c.point_lights.clear();
    c.proxies.clear();
    c.transient_models.clear();
    c.tmat_tex.clear();
    c.tmat_color.clear();
    c.tmat_shininess.clear();
}

extern "C" i32 sftx_view_end(u32* color, i32 pitch_bytes, u64* vis64) {
    Context& c = C();
    if (!c.batch_open) {
        fprintf(stderr, "sofren2: sftx_view_end outside sftx_batch_begin/sftx_batch_end\n");
        return -1;
    }
    if (c.view_rejected) {
        c.view_rejected = false;
        return -1;
    }
    c.batch_views.emplace_back();
    ViewRecord& v = c.batch_views.back();
    snapshot_view(c, v);
    v.color_buffer = color;
    v.color_pitch = pitch_bytes / 4;
    v.vis64 = vis64;
    v.capture_flags = 0;
    /* the view is rendered later: keep the poses as they are now */
    size_t nb = 0;
    for (const sft_instance& in : v.instances) {
        if (in.posed_bone_matrices) nb += (size_t)std::max(in.model->skeleton.bone_count, 1);
    }
    v.bone_store.resize(nb);
    v.bones_kept.assign(v.instances.size(), 0);
    size_t at = 0;
    for (size_t i = 0; i < v.instances.size(); i += 1) {
        sft_instance& in = v.instances[i];
        if (!in.posed_bone_matrices) continue;
        const size_t n = (size_t)std::max(in.model->skeleton.bone_count, 1);
        memcpy(&v.bone_store[at], in.posed_bone_matrices, n * sizeof(sftmat));
        in.posed_bone_matrices = &v.bone_store[at];
        v.bones_kept[i] = (u32)n;
        at += n;
    }
    return (i32)c.batch_views.size() - 1;
}

extern "C" void sftx_batch_submit(void) {
    Context& c = C();
    if (!c.batch_open || c.batch_views.empty()) return;
    if (c.server_running) {
        fprintf(stderr, "sofren2: sftx_batch_submit while the request server runs; submit requests instead\n");
        return;
    }
    if (!c.backend) c.backend = backend_create();
    drain_async_batches(c);
    backend_render(c.backend, &c, c.batch_views);
}

/* sftx_batch_submit after the autotuner has timed candidate settings on the
   recorded views (backend_tune) */
extern "C" i32 sftx_batch_tune(i32 repeats, sftx_tuning* report) {
    Context& c = C();
    if (!c.batch_open || c.batch_views.empty()) return -1;
    if (c.server_running) {
        fprintf(stderr, "sofren2: sftx_batch_tune while the request server runs\n");
        return -1;
    }
    if (!c.backend) c.backend = backend_create();
    drain_async_batches(c);
    sftx_tuning local{};
    backend_tune(c.backend, &c, c.batch_views, repeats, report ? report : &local);
    return 0;
}

extern "C" void sftx_get_tuning(sftx_tuning* out) {
    Context& c = C();
    if (!c.backend) c.backend = backend_create();
    *out = sftx_tuning{};
    backend_get_tuning(c.backend, out);
}

extern "C" void sftx_set_tuning(const sftx_tuning* in) {
    Context& c = C();
    if (!c.backend) c.backend = backend_create();
    if (c.server_running) {
        fprintf(stderr, "sofren2: sftx_set_tuning while the request server runs\n");
        return;
    }
    drain_async_batches(c);
    backend_set_tuning(c.backend, in);
}

/* The recorded views as a batch on a free frame set, without waiting: the
   host can record the next batches while this one runs.  kBatchesInFlight
   batches can be in flight; one more submission completes the oldest first.
   Output buffers belong to the renderer until sftx_batch_wait returns for the
   ticket. */
extern "C" i32 sftx_batch_submit_async(void) {
    Context& c = C();
    if (!c.batch_open || c.batch_views.empty()) return 0;
    if (c.server_running) {
        fprintf(stderr, "sofren2: sftx_batch_submit_async while the request server runs; submit requests instead\n");
        return 0;
    }
    if (!c.backend) c.backend = backend_create();
    while (c.async_batches.size() >= (size_t)kBatchesInFlight) {
        Context::AsyncBatch& old = c.async_batches.front();
        backend_complete(c.backend, &c, old.views, old.set);
        c.async_batches.pop_front();
    }
    Context::AsyncBatch ab;
    ab.ticket = c.next_ticket;
    c.next_ticket = (c.next_ticket == 0x7FFFFFFF) ? 1 : c.next_ticket + 1;
    ab.set = free_frame_set(c, kBatchesInFlight);
    ab.views.swap(c.batch_views);
    c.async_batches.push_back(std::move(ab));
    Context::AsyncBatch& cur = c.async_batches.back();
    backend_submit(c.backend, &c, cur.views, cur.set, true);
    return cur.ticket;
}

/* completes batches in submission order up to the ticket's; 0 when done,
   -1 for a ticket that is not in flight */
extern "C" i32 sftx_batch_wait(i32 ticket) {
    if (!g_ctx) return -1;
    Context& c = *g_ctx;
    bool found = false;
    for (const Context::AsyncBatch& ab : c.async_batches) found = found || ab.ticket == ticket;
    if (!found) return -1;
    while (!c.async_batches.empty()) {
        Context::AsyncBatch& ab = c.async_batches.front();
        const i32 t = ab.ticket;
        backend_complete(c.backend, &c, ab.views, ab.set);
        c.async_batches.pop_front();
        if (t == ticket) break;
    }
    return 0;
}

void sft::drain_async_batches(Context& c) {
    while (!c.async_batches.empty()) {
        Context::AsyncBatch& ab = c.async_batches.front();
        backend_complete(c.backend, &c, ab.views, ab.set);
        c.async_batches.pop_front();
    }
}

extern "C" void sftx_batch_end(void) {
    Context& c = C();
    c.batch_views.clear();
    c.batch_open = false;
    sft_frame_end();
}

/* ================================================================ misc setters */
/* view_proj = view * proj and six normalised planes (sofren2.c:8345-8376) */
void sft::camera_setup(const sftmat& view, const sftmat& proj, sftmat* view_proj, f32 out[6][4]) {
    *view_proj = sft_m_mul(view, proj);
    const sftmat& vp = *view_proj;
    f32 planes[6][4];
    for (i32 i = 0; i < 4; i += 1) {
        planes[0][i] = vp.m[i][3] + vp.m[i][0];
        planes[1][i] = vp.m[i][3] - vp.m[i][0];
        planes[2][i] = vp.m[i][3] + vp.m[i][1];
        planes[3][i] = vp.m[i][3] - vp.m[i][1];
        planes[4][i] = vp.m[i][3] + vp.m[i][2];
        planes[5][i] = vp.m[i][3] - vp.m[i][2];
    }
    for (i32 p = 0; p < 6; p += 1) {
        const f32 len = sqrtf(planes[p][0] * planes[p][0] + planes[p][1] * planes[p][1] + planes[p][2] * planes[p][2]);
        const f32 inv_len = (len < SFT_EPSILON) ? SFT_EPSILON : 1.f / len;
        out[p][0] = planes[p][0] * inv_len;
        out[p][1] = planes[p][1] * inv_len;
        out[p][2] = planes[p][2] * inv_len;
        out[p][3] = planes[p][3] * inv_len;
    }
}

extern "C" void sft_set_camera(sftvec cam_pos, sftmat view, sftmat proj) {
    Context& c = C();
    c.camera_pos = cam_pos;
    c.view = view;
    c.proj = proj;
    camera_setup(view, proj, &c.view_proj, c.planes);
}

extern "C" void sft_set_debug_mode(enum sft_debug_mode mode) {
    if (mode < 0 || mode >= SFT_DEBUG_COUNT) {
        SFT_ERR_RET(, "invalid debug mode '%d', expected [0, %d)\n", (int)mode, (int)SFT_DEBUG_COUNT);
    }
    C().debug_mode = mode;
}

extern "C" const char* sft_get_debug_name(enum sft_debug_mode mode) {
    switch (mode) {
        case SFT_DEBUG_NONE: return "None";
        case SFT_DEBUG_WIREFRAME: return "Wireframe";
        case SFT_DEBUG_TILES: return "Tiles";
        case SFT_DEBUG_DEPTH: return "Depth";
        case SFT_DEBUG_LIGHT_COUNT: return "Light Count";
        case SFT_DEBUG_CHECKERBOARD: return "Checkerboard";
        default: return "INVALID MODE";
    }
}

extern "C" void sft_set_clear(u8 enabled) { C().should_clear = enabled; }

extern "C" void sftx_set_msaa(i32 samples) {
    if (1 != samples && 2 != samples && 4 != samples) {
        fprintf(stderr, "sofren2: %d MSAA samples per pixel not supported (1, 2 or 4)\n", samples);
        return;
    }
    C().samples = samples;
}
/* the reference's SFT_INSTANCE_BITS / SFT_MESHLET_BITS (sofren2.c:154-165),
   chosen at run time: instance_bits + meshlet bits = 24 */
extern "C" void sftx_set_id_split(i32 instance_bits) {
    if (instance_bits < 1 || instance_bits > 16) {
        fprintf(stderr, "sofren2: %d instance bits not supported (1..16, meshlet bits = 24 - instance bits)\n",
                instance_bits);
        return;
    }
    C().instance_bits = (u32)instance_bits;
}
extern "C" i32 sftx_get_id_split(void) { return (i32)C().instance_bits; }
extern "C" void sft_set_clear_color(u32 color) { C().clear_color = color; }
extern "C" void sft_set_color_buffer(u32* buf, i32 pitch_in_bytes) {
    Context& c = C();
    c.color_buffer = buf;
    c.color_pitch = pitch_in_bytes / (i32)sizeof(u32);
}

/* ================================================================ lights
   (sofren2.c:9109-9149) */
extern "C" void sft_set_ambient_light(f32 r, f32 g, f32 b) {
    Context& c = C();
    c.ambient.x = r;
    c.ambient.y = g;
    c.ambient.z = b;
}

extern "C" void sft_draw_directional_light(sftvec dir, f32 r, f32 g, f32 b, f32 intensity) {
    sft_directional_light l;
    l.dir = sft_v_norm(dir);
    l.color.x = r * intensity;
    l.color.y = g * intensity;
    l.color.z = b * intensity;
    C().dir_lights.push_back(l);
}

extern "C" void sft_draw_point_light(sftvec pos, f32 r, f32 g, f32 b, f32 intensity, f32 radius, u32 light_id) {
    (void)light_id;
    sft_point_light l;
    l.pos = pos;
    l.color.x = r * intensity;
    l.color.y = g * intensity;
    l.color.z = b * intensity;
    l.radius = radius;
    l.inv_radius_sqr = 1.f / (radius * radius);
    C().point_lights.push_back(l);
}

extern "C" void sft_set_shadows_enabled(u8 enabled) { C().shadows_enabled = enabled; }
extern "C" void sft_set_lightmap_brightness(f32 b) { C().lightmap_brightness = b; }

/* ================================================================ instances
   (sofren2.c:8410-8478): slot claim with the instance cap of the ID split
   (64 by default), screen-space cache offset, world-space shadow proxies.
   Geometry jobs are implicit: the backend walks instances, meshes and
   meshlets in this same order. */
extern "C" void sft_draw_instance(const sft_instance* inst) {
    Context& c = C();
    const i32 slot = c.instance_count++;
    if (slot >= (i32)(1u << c.instance_bits)) {
        return;
    }
    sft_instance copy = *inst;
    copy.ss_meshlet_offset = c.total_ss;
    c.total_ss += inst->model->total_meshlets;
    c.instances.push_back(copy);

    if (c.shadows_enabled && inst->model && inst->model->shadow_proxy_count > 0) {
        for (u32 i = 0; i < inst->model->shadow_proxy_count; i += 1) {
            const sft_shadow_proxy* local = &inst->model->shadow_proxies[i];
            sftmat m = inst->transform;
            if (local->bone_ind != 0xFFFFFFFFu && inst->posed_bone_matrices) {
                m = sft_m_mul(inst->posed_bone_matrices[local->bone_ind], inst->transform);
            }
            WorldProxy w;
            memset(&w, 0, sizeof(w));
            w.posa.x = local->posa.x * m.m[0][0] + local->posa.y * m.m[1][0] + local->posa.z * m.m[2][0] + m.m[3][0];
            w.posa.y = local->posa.x * m.m[0][1] + local->posa.y * m.m[1][1] + local->posa.z * m.m[2][1] + m.m[3][1];
            w.posa.z = local->posa.x * m.m[0][2] + local->posa.y * m.m[1][2] + local->posa.z * m.m[2][2] + m.m[3][2];
            w.posb.x = local->posb.x * m.m[0][0] + local->posb.y * m.m[1][0] + local->posb.z * m.m[2][0] + m.m[3][0];
            w.posb.y = local->posb.x * m.m[0][1] + local->posb.y * m.m[1][1] + local->posb.z * m.m[2][1] + m.m[3][1];
            w.posb.z = local->posb.x * m.m[0][2] + local->posb.y * m.m[1][2] + local->posb.z * m.m[2][2] + m.m[3][2];
            const f32 sx = sqrtf(m.m[0][0] * m.m[0][0] + m.m[0][1] * m.m[0][1] + m.m[0][2] * m.m[0][2]);
            w.radius = local->radius * sx;
            c.proxies.push_back(w);
        }
    }
}

/* ================================================================ dmesh
   transient meshlets built on the host each frame (sofren2.c:2816-2883,
   8480-8938) */
static void dmesh_finalize_chunk(sft_dmesh* dm) {
    Context& c = C();
    if (0 == dm->curr_vert) return;
    const u32 mi = dm->curr_meshlet;
    if (mi >= dm->max_meshlets) return;
    sft_meshlet* m = &dm->meshlets[mi];
    sft_meshlet_attribs* a = &dm->attribs[mi];

    /* unused trailing vertices collapse onto vertex 0 so template triangles
       that reference them are degenerate */
    for (u32 i = dm->curr_vert; i < SFT_VERTS_PER_MESHLET; i += 1) {
        m->x[i] = m->x[0]; m->y[i] = m->y[0]; m->z[i] = m->z[0];
        a->nx[i] = a->nx[0]; a->ny[i] = a->ny[0]; a->nz[i] = a->nz[0];
        a->u[i] = a->u[0]; a->v[i] = a->v[0];
        a->lu[i] = a->lu[0]; a->lv[i] = a->lv[0];
    }

    f32 mnx = 1e30f, mxx = -1e30f, mny = 1e30f, mxy = -1e30f, mnz = 1e30f, mxz = -1e30f;
    for (u32 i = 0; i < dm->curr_vert; i += 1) {
        const f32 x = m->x[i], y = m->y[i], z = m->z[i];
        if (x < mnx) mnx = x;
        if (x > mxx) mxx = x;
        if (y < mny) mny = y;
        if (y > mxy) mxy = y;
        if (z < mnz) mnz = z;
        if (z > mxz) mxz = z;
    }
    const f32 cx = (mnx + mxx) * 0.5f, cy = (mny + mxy) * 0.5f, cz = (mnz + mxz) * 0.5f;
    const f32 dx = mxx - cx, dy = mxy - cy, dz = mxz - cz;
    dm->bounds_min_x[mi] = mnx; dm->bounds_max_x[mi] = mxx;
    dm->bounds_min_y[mi] = mny; dm->bounds_max_y[mi] = mxy;
    dm->bounds_min_z[mi] = mnz; dm->bounds_max_z[mi] = mxz;
    dm->bounds_center_x[mi] = cx;
    dm->bounds_center_y[mi] = cy;
    dm->bounds_center_z[mi] = cz;
    dm->bounds_radius[mi] = sqrtf(dx * dx + dy * dy + dz * dz);
    /* cutoff 2 disables the normal-cone test */
    dm->cone_axis_x[mi] = 0.f;
    dm->cone_axis_y[mi] = 0.f;
    dm->cone_axis_z[mi] = 1.f;
    dm->cone_cutoff[mi] = 2.f;
    memcpy(&dm->inds[mi], (SFT_DMESH_QUADS == dm->type) ? &c.quad_inds : &c.tri_inds, sizeof(sft_meshlet_inds));
    dm->curr_meshlet += 1;
    dm->curr_vert = 0;
}

extern "C" sft_dmesh* sft_dmesh_begin(enum sft_dmesh_type type, sft_mat_handle material, u32 max_prims) {
    Context& c = C();
    if (0 == max_prims) return NULL;
    const u32 per_chunk = (SFT_DMESH_QUADS == type) ? 64u : 63u;
    const u32 per_prim = (SFT_DMESH_QUADS == type) ? 4u : 3u;
    const u32 total = max_prims * per_prim;
    const u32 max_meshlets = (((total + per_chunk - 1) / per_chunk) + 7u) & ~7u;
    sft_arena* ar = &c.scratch_arena;
    sft_dmesh* dm = (sft_dmesh*)sft_arena_alloc(ar, sizeof(sft_dmesh));
    dm->type = type;
    dm->material = material;
    dm->max_meshlets = max_meshlets;
    dm->curr_meshlet = 0;
    dm->curr_vert = 0;
    dm->max_verts_per_chunk = per_chunk;
    dm->meshlets = (sft_meshlet*)sft_arena_alloc(ar, max_meshlets * sizeof(sft_meshlet));
    dm->attribs = (sft_meshlet_attribs*)sft_arena_alloc(ar, max_meshlets * sizeof(sft_meshlet_attribs));
    dm->inds = (sft_meshlet_inds*)sft_arena_alloc(ar, max_meshlets * sizeof(sft_meshlet_inds));
    f32** arrays[] = {
        &dm->bounds_center_x, &dm->bounds_center_y, &dm->bounds_center_z, &dm->bounds_radius,
        &dm->bounds_min_x, &dm->bounds_min_y, &dm->bounds_min_z,
        &dm->bounds_max_x, &dm->bounds_max_y, &dm->bounds_max_z,
        &dm->cone_axis_x, &dm->cone_axis_y, &dm->cone_axis_z, &dm->cone_cutoff };
    for (f32** a : arrays) {
        *a = (f32*)sft_arena_alloc(ar, max_meshlets * sizeof(f32));
    }
    return dm;
}

extern "C" void sft_dmesh_push_quad_4n(sft_dmesh* dm, sftvec p0, sftvec p1, sftvec p2, sftvec p3,
                                       f32 u0, f32 v0, f32 u1, f32 v1,
                                       sftvec n0, sftvec n1, sftvec n2, sftvec n3) {
    if (dm->curr_vert >= dm->max_verts_per_chunk) dmesh_finalize_chunk(dm);
    if (dm->curr_meshlet >= dm->max_meshlets) return;
    const u32 v = dm->curr_vert;
    sft_meshlet* m = &dm->meshlets[dm->curr_meshlet];
    sft_meshlet_attribs* a = &dm->attribs[dm->curr_meshlet];
    const sftvec p[4] = { p0, p1, p2, p3 };
    const sftvec n[4] = { n0, n1, n2, n3 };
    const f32 us[4] = { u0, u1, u1, u0 };
    const f32 vs[4] = { v0, v0, v1, v1 };
    for (u32 i = 0; i < 4; i += 1) {
        m->x[v + i] = p[i].x; m->y[v + i] = p[i].y; m->z[v + i] = p[i].z;
        a->u[v + i] = us[i]; a->v[v + i] = vs[i];
        a->nx[v + i] = n[i].x; a->ny[v + i] = n[i].y; a->nz[v + i] = n[i].z;
        a->lu[v + i] = 0.f; a->lv[v + i] = 0.f;
    }
    dm->curr_vert += 4;
}

extern "C" void sft_dmesh_push_quad(sft_dmesh* dm, sftvec p0, sftvec p1, sftvec p2, sftvec p3,
                                    f32 u0, f32 v0, f32 u1, f32 v1, sftvec normal) {
    sft_dmesh_push_quad_4n(dm, p0, p1, p2, p3, u0, v0, u1, v1, normal, normal, normal, normal);
}

extern "C" void sft_dmesh_push_tri(sft_dmesh* dm, sftvec p0, sftvec p1, sftvec p2,
                                   f32 u0, f32 v0, f32 u1, f32 v1, f32 u2, f32 v2, sftvec normal) {
    if (dm->curr_vert >= dm->max_verts_per_chunk) dmesh_finalize_chunk(dm);
    if (dm->curr_meshlet >= dm->max_meshlets) return;
    const u32 v = dm->curr_vert;
    sft_meshlet* m = &dm->meshlets[dm->curr_meshlet];
    sft_meshlet_attribs* a = &dm->attribs[dm->curr_meshlet];
    const sftvec p[3] = { p0, p1, p2 };
    const f32 us[3] = { u0, u1, u2 };
    const f32 vs[3] = { v0, v1, v2 };
    for (u32 i = 0; i < 3; i += 1) {
        m->x[v + i] = p[i].x; m->y[v + i] = p[i].y; m->z[v + i] = p[i].z;
        a->u[v + i] = us[i]; a->v[v + i] = vs[i];
        a->nx[v + i] = normal.x; a->ny[v + i] = normal.y; a->nz[v + i] = normal.z;
        a->lu[v + i] = 0.f; a->lv[v + i] = 0.f;
    }
    dm->curr_vert += 3;
}

extern "C" u8 sft_dmesh_alloc_quads(sft_dmesh* dm, u32 count, sft_dmesh_qb* out) {
    u32 space = (dm->max_verts_per_chunk - dm->curr_vert) / 4;
    if (0 == space) {
        dmesh_finalize_chunk(dm);
        space = dm->max_verts_per_chunk / 4;
    }
    if (dm->curr_meshlet >= dm->max_meshlets) return 0;
    const u32 n = (count < space) ? count : space;
    sft_meshlet* m = &dm->meshlets[dm->curr_meshlet];
    sft_meshlet_attribs* a = &dm->attribs[dm->curr_meshlet];
    out->x = m->x; out->y = m->y; out->z = m->z;
    out->nx = a->nx; out->ny = a->ny; out->nz = a->nz;
    out->u = a->u; out->v = a->v;
    out->vert_offset = dm->curr_vert;
    out->allocated_count = n;
    dm->curr_vert += n * 4;
    return 1;
}

/* wraps the dmesh in a one-mesh transient model and draws it (8881-8938) */
extern "C" void sft_dmesh_end(sft_dmesh* dm, const sftmat* opt_transform) {
    Context& c = C();
    if (!dm) return;
    if (dm->curr_vert > 0) dmesh_finalize_chunk(dm);
    if (0 == dm->curr_meshlet) return;
    sft_arena* ar = &c.scratch_arena;
    sft_mesh* mesh = (sft_mesh*)sft_arena_alloc(ar, sizeof(sft_mesh));
    memset(mesh, 0, sizeof(sft_mesh));
    mesh->meshlets = dm->meshlets;
    mesh->meshlet_attribs = dm->attribs;
    mesh->inds = dm->inds;
    mesh->meshlet_count = dm->curr_meshlet;
    mesh->material = dm->material;
    mesh->lightmap_page = 0xFFFFFFFFu;
    mesh->bounds_center_x = dm->bounds_center_x;
    mesh->bounds_center_y = dm->bounds_center_y;
    mesh->bounds_center_z = dm->bounds_center_z;
    mesh->bounds_radius = dm->bounds_radius;
    mesh->bounds_min_x = dm->bounds_min_x;
    mesh->bounds_min_y = dm->bounds_min_y;
    mesh->bounds_min_z = dm->bounds_min_z;
    mesh->bounds_max_x = dm->bounds_max_x;
    mesh->bounds_max_y = dm->bounds_max_y;
    mesh->bounds_max_z = dm->bounds_max_z;
    mesh->cone_axis_x = dm->cone_axis_x;
    mesh->cone_axis_y = dm->cone_axis_y;
    mesh->cone_axis_z = dm->cone_axis_z;
    mesh->cone_cutoff = dm->cone_cutoff;
    sft_model* model = (sft_model*)sft_arena_alloc(ar, sizeof(sft_model));
    memset(model, 0, sizeof(sft_model));
    model->meshes = mesh;
    model->mesh_count = 1;
    model->total_meshlets = dm->curr_meshlet;
    model->meshlet_mesh_inds = (u16*)sft_arena_alloc(ar, dm->curr_meshlet * sizeof(u16));
    memset(model->meshlet_mesh_inds, 0, dm->curr_meshlet * sizeof(u16));
    c.transient_models.push_back(model);

    sft_instance inst;
    memset(&inst, 0, sizeof(inst));
    inst.model = model;
    inst.transform = opt_transform ? *opt_transform : sft_m_identity();
    sft_draw_instance(&inst);
}

/* ---- primitive generators (same vertex math as sofren2.c:8480-8879, which
   follows raylib's DrawCube / DrawSphereEx / DrawCapsule / DrawCylinderEx) */
extern "C" void sft_draw_cube(sft_dmesh* dm) {
    const f32 h = 0.5f;
    sft_dmesh_push_quad(dm, SFT_V(-h,-h,-h), SFT_V(-h, h,-h), SFT_V( h, h,-h), SFT_V( h,-h,-h), 0.f,0.f, 1.f,1.f, SFT_V(0.f, 0.f, -1.f));
    sft_dmesh_push_quad(dm, SFT_V( h,-h,-h), SFT_V( h, h,-h), SFT_V( h, h, h), SFT_V( h,-h, h), 0.f,0.f, 1.f,1.f, SFT_V(1.f, 0.f, 0.f));
    sft_dmesh_push_quad(dm, SFT_V( h,-h, h), SFT_V( h, h, h), SFT_V(-h, h, h), SFT_V(-h,-h, h), 0.f,0.f, 1.f,1.f, SFT_V(0.f, 0.f, 1.f));
    sft_dmesh_push_quad(dm, SFT_V(-h,-h, h), SFT_V(-h, h, h), SFT_V(-h, h,-h), SFT_V(-h,-h,-h), 0.f,0.f, 1.f,1.f, SFT_V(-1.f, 0.f, 0.f));
    sft_dmesh_push_quad(dm, SFT_V(-h, h,-h), SFT_V(-h, h, h), SFT_V( h, h, h), SFT_V( h, h,-h), 0.f,0.f, 1.f,1.f, SFT_V(0.f, 1.f, 0.f));
    sft_dmesh_push_quad(dm, SFT_V( h,-h, h), SFT_V(-h,-h, h), SFT_V(-h,-h,-h), SFT_V( h,-h,-h), 0.f,0.f, 1.f,1.f, SFT_V(0.f, -1.f, 0.f));
}

extern "C" void sft_draw_cube_inv(sft_dmesh* dm) {
    const f32 h = 0.5f;
    sft_dmesh_push_quad(dm, SFT_V( h, h,-h), SFT_V(-h, h,-h), SFT_V(-h,-h,-h), SFT_V( h,-h,-h), 0.f,0.f, 1.f,1.f, SFT_V(0.f, 0.f, 1.f));
    sft_dmesh_push_quad(dm, SFT_V( h, h, h), SFT_V( h, h,-h), SFT_V( h,-h,-h), SFT_V( h,-h, h), 0.f,0.f, 1.f,1.f, SFT_V(-1.f, 0.f, 0.f));
    sft_dmesh_push_quad(dm, SFT_V(-h, h, h), SFT_V( h, h, h), SFT_V( h,-h, h), SFT_V(-h,-h, h), 0.f,0.f, 1.f,1.f, SFT_V(0.f, 0.f, -1.f));
    sft_dmesh_push_quad(dm, SFT_V(-h, h,-h), SFT_V(-h, h, h), SFT_V(-h,-h, h), SFT_V(-h,-h,-h), 0.f,0.f, 1.f,1.f, SFT_V(1.f, 0.f, 0.f));
    sft_dmesh_push_quad(dm, SFT_V( h, h, h), SFT_V(-h, h, h), SFT_V(-h, h,-h), SFT_V( h, h,-h), 0.f,0.f, 1.f,1.f, SFT_V(0.f, -1.f, 0.f));
    sft_dmesh_push_quad(dm, SFT_V(-h,-h,-h), SFT_V(-h,-h, h), SFT_V( h,-h, h), SFT_V( h,-h,-h), 0.f,0.f, 1.f,1.f, SFT_V(0.f, 1.f, 0.f));
}

/* The dmesh state a sphere leaves depends only on (rings, slices) and the
   state it starts from.  Started on an empty quad dmesh, the resulting
   meshlet blocks, bounds and counters are kept per thread and copied into the
   next empty dmesh asked for the same sphere: the same bytes the pushes would
   write. */
namespace {
struct SphereSnapshot {
    i32 rings = -1, slices = -1;
    u32 meshlets = 0, curr_meshlet = 0, curr_vert = 0;
    std::vector<sft_meshlet> m;
    std::vector<sft_meshlet_attribs> a;
    std::vector<sft_meshlet_inds> ind;
    std::vector<f32> bounds;    /* 14 arrays of `meshlets` floats */
};

f32** dmesh_bound_arrays(sft_dmesh* dm, f32** out) {
    f32* const arrays[14] = { dm->bounds_center_x, dm->bounds_center_y, dm->bounds_center_z, dm->bounds_radius,
                              dm->bounds_min_x, dm->bounds_min_y, dm->bounds_min_z, dm->bounds_max_x,
                              dm->bounds_max_y, dm->bounds_max_z, dm->cone_axis_x, dm->cone_axis_y,
                              dm->cone_axis_z, dm->cone_cutoff };
    for (int k = 0; k < 14; k += 1) out[k] = arrays[k];
    return out;
}
}  // namespace

static void push_sphere_quads(sft_dmesh* dm, i32 rings, i32 slices) {
    const f32 ring_angle = (f32)(SFT_PI / (rings + 1));
    const f32 slice_angle = (f32)(2.f * SFT_PI / slices);
    const f32 cr = cosf(ring_angle), sr = sinf(ring_angle);
    const f32 cs = cosf(slice_angle), ss = sinf(slice_angle);
    sftvec v[4] = { {0, 0, 0}, {0, 0, 0}, {0.f, 1.f, 0.f}, {sr, cr, 0.f} };
    for (i32 i = 0; i <= rings; i += 1) {
        for (i32 j = 0; j < slices; j += 1) {
            v[0] = v[2];
            v[1] = v[3];
            const sftvec a = v[2], b = v[3];
            v[2] = SFT_V(cs * a.x - ss * a.z, a.y, ss * a.x + cs * a.z);
            v[3] = SFT_V(cs * b.x - ss * b.z, b.y, ss * b.x + cs * b.z);
            sft_dmesh_push_quad_4n(dm, v[0], v[2], v[3], v[1], 0.f, 0.f, 1.f, 1.f,
                sft_v_norm(v[0]), sft_v_norm(v[2]), sft_v_norm(v[3]), sft_v_norm(v[1]));
        }
        v[2] = v[3];
        const sftvec b = v[3];
        v[3] = SFT_V(cr * b.x + sr * b.y, -sr * b.x + cr * b.y, b.z);
    }
}

/* UV sphere: rings+1 latitude bands of `slices` quads, starting at the pole
   and rotating the generating edge about y then z */
extern "C" void sft_draw_sphere(sft_dmesh* dm, i32 rings, i32 slices) {
    thread_local SphereSnapshot snap;
    const bool fresh = dm && SFT_DMESH_QUADS == dm->type && 0u == dm->curr_meshlet && 0u == dm->curr_vert &&
                       64u == dm->max_verts_per_chunk;
    f32* arr[14];
    if (fresh && snap.rings == rings && snap.slices == slices && snap.meshlets <= dm->max_meshlets) {
        const u32 n = snap.meshlets;
        memcpy(dm->meshlets, snap.m.data(), n * sizeof(sft_meshlet));
        memcpy(dm->attribs, snap.a.data(), n * sizeof(sft_meshlet_attribs));
        memcpy(dm->inds, snap.ind.data(), n * sizeof(sft_meshlet_inds));
        dmesh_bound_arrays(dm, arr);
        for (int k = 0; k < 14; k += 1) memcpy(arr[k], snap.bounds.data() + (size_t)k * n, n * sizeof(f32));
        dm->curr_meshlet = snap.curr_meshlet;
        dm->curr_vert = snap.curr_vert;
        return;
    }
    push_sphere_quads(dm, rings, slices);
    /* kept only when no push was dropped for want of meshlets */
    if (!fresh || dm->curr_meshlet >= dm->max_meshlets) return;
    /* meshlets written so far: the finalized ones plus the open chunk */
    const u32 n = std::min(dm->curr_meshlet + (dm->curr_vert > 0 ? 1u : 0u), dm->max_meshlets);
    snap.rings = rings;
    snap.slices = slices;
    snap.meshlets = n;
    snap.curr_meshlet = dm->curr_meshlet;
    snap.curr_vert = dm->curr_vert;
    snap.m.assign(dm->meshlets, dm->meshlets + n);
    snap.a.assign(dm->attribs, dm->attribs + n);
    snap.ind.assign(dm->inds, dm->inds + n);
    snap.bounds.resize((size_t)14 * n);
    dmesh_bound_arrays(dm, arr);
    for (int k = 0; k < 14; k += 1) memcpy(snap.bounds.data() + (size_t)k * n, arr[k], n * sizeof(f32));
}

static sftvec capsule_ring_point(sftvec c, sftvec b0, sftvec b1, sftvec b2, f32 sl, f32 rg, f32 radius) {
    const f32 rs = sinf(sl) * cosf(rg);
    const f32 rc = cosf(sl) * cosf(rg);
    const f32 sg = sinf(rg);
    return SFT_V(c.x + (sg * b0.x + rs * b1.x + rc * b2.x) * radius,
                 c.y + (sg * b0.y + rs * b1.y + rc * b2.y) * radius,
                 c.z + (sg * b0.z + rs * b1.z + rc * b2.z) * radius);
}

static sftvec ring_point(sftvec c, sftvec b1, sftvec b2, f32 s, f32 k) {
    return SFT_V(c.x + s * b1.x + k * b2.x, c.y + s * b1.y + k * b2.y, c.z + s * b1.z + k * b2.z);
}

extern "C" void sft_draw_capsule(sft_dmesh* dm, sftvec start_pos, sftvec end_pos, f32 radius, i32 slices, i32 rings) {
    if (slices < 3) slices = 3;
    sftvec dir = sft_v_sub(end_pos, start_pos);
    const u8 sphere_case = (0.f == dir.x && 0.f == dir.y && 0.f == dir.z);
    if (sphere_case) dir = SFT_V(0.f, 1.f, 0.f);
    sftvec b0 = sft_v_norm(dir);
    const sftvec b1 = sft_v_norm(sft_v_perpendicular(dir));
    const sftvec b2 = sft_v_norm(sft_v_cross(b1, dir));
    sftvec cap = end_pos;
    const f32 sa = (f32)(2.f * SFT_PI / slices);
    const f32 ra = (f32)(0.5f * SFT_PI / rings);
    for (i32 c = 0; c < 2; c += 1) {
        for (i32 i = 0; i < rings; i += 1) {
            for (i32 j = 0; j < slices; j += 1) {
                const sftvec w1 = capsule_ring_point(cap, b0, b1, b2, sa * (j + 0), ra * (i + 0), radius);
                const sftvec w2 = capsule_ring_point(cap, b0, b1, b2, sa * (j + 1), ra * (i + 0), radius);
                const sftvec w3 = capsule_ring_point(cap, b0, b1, b2, sa * (j + 0), ra * (i + 1), radius);
                const sftvec w4 = capsule_ring_point(cap, b0, b1, b2, sa * (j + 1), ra * (i + 1), radius);
                const sftvec centre = sft_v_mul(sft_v_add(sft_v_add(w1, w2), sft_v_add(w3, w4)), 0.25f);
                const sftvec n = sft_v_norm(sft_v_sub(centre, cap));
                if (0 == c) sft_dmesh_push_quad(dm, w3, w1, w2, w4, 0.f, 0.f, 1.f, 1.f, n);
                else        sft_dmesh_push_quad(dm, w3, w4, w2, w1, 0.f, 0.f, 1.f, 1.f, n);
            }
        }
        cap = start_pos;
        b0 = sft_v_mul(b0, -1.f);
    }
    if (!sphere_case) {
        for (i32 j = 0; j < slices; j += 1) {
            const f32 s1 = sinf(sa * (j + 0)) * radius, c1 = cosf(sa * (j + 0)) * radius;
            const f32 s2 = sinf(sa * (j + 1)) * radius, c2 = cosf(sa * (j + 1)) * radius;
            const sftvec w1 = ring_point(start_pos, b1, b2, s1, c1);
            const sftvec w2 = ring_point(start_pos, b1, b2, s2, c2);
            const sftvec w3 = ring_point(end_pos, b1, b2, s1, c1);
            const sftvec w4 = ring_point(end_pos, b1, b2, s2, c2);
            sft_dmesh_push_quad(dm, w3, w1, w2, w4, 0.f, 0.f, 1.f, 1.f, sft_v_face_normal(w3, w1, w2));
        }
    }
}

extern "C" void sft_draw_cylinder(sft_dmesh* dm, sftvec start_pos, sftvec end_pos, f32 start_radius, f32 end_radius, i32 sides) {
    if (sides < 3) sides = 3;
    const sftvec dir = sft_v_sub(end_pos, start_pos);
    if (0 == dir.x && 0 == dir.y && 0 == dir.z) return;
    const sftvec b1 = sft_v_norm(sft_v_perpendicular(dir));
    const sftvec b2 = sft_v_norm(sft_v_cross(b1, dir));
    const f32 a = (f32)((2.f * SFT_PI) / sides);
    for (i32 i = 0; i < sides; i += 1) {
        const f32 s1 = sinf(a * (i + 0)) * start_radius, c1 = cosf(a * (i + 0)) * start_radius;
        const f32 s2 = sinf(a * (i + 1)) * start_radius, c2 = cosf(a * (i + 1)) * start_radius;
        const f32 s3 = sinf(a * (i + 0)) * end_radius, c3 = cosf(a * (i + 0)) * end_radius;
        const f32 s4 = sinf(a * (i + 1)) * end_radius, c4 = cosf(a * (i + 1)) * end_radius;
        const sftvec w1 = ring_point(start_pos, b1, b2, s1, c1);
        const sftvec w2 = ring_point(start_pos, b1, b2, s2, c2);
        const sftvec w3 = ring_point(end_pos, b1, b2, s3, c3);
        const sftvec w4 = ring_point(end_pos, b1, b2, s4, c4);
        if (start_radius > 0) {
            sft_dmesh_push_quad(dm, start_pos, w2, w1, w1, 0.f, 0.f, 1.f, 1.f, sft_v_face_normal(start_pos, w2, w1));
        }
        sft_dmesh_push_quad(dm, w3, w1, w2, w4, 0.f, 0.f, 1.f, 1.f, sft_v_face_normal(w3, w1, w2));
        if (end_radius > 0) {
            sft_dmesh_push_quad(dm, end_pos, w3, w4, w4, 0.f, 0.f, 1.f, 1.f, sft_v_face_normal(end_pos, w3, w4));
        }
    }
}

/* ================================================================ QSF2 loader
   Binary scene format (sofren2.c:7816-7967): header, textures (each becomes a
   material), lightmap pages, meshes with SoA meshlets, flat triangle indices
   re-laid into 8-wide batches, bounds, cones, meshlet->mesh map, BLAS nodes. */
/* A QSF2 file that ends early, or whose counts or indices do not fit its
   own contents, is rejected as a whole: the reference reads such files
   without checks (sofren2.c:7816-7967); here the load returns NULL with a
   message, its recoverable-error convention. */
struct QsfBad { const char* what; };

struct QsfReader {
    FILE* f;
    u64 left;       /* bytes not yet read */
    void read(void* dst, u64 bytes) {
        if (bytes > left || (bytes && 1 != fread(dst, bytes, 1, f))) throw QsfBad{"file ends early"};
        left -= bytes;
    }
    /* a count of `n` records of `each` bytes must fit in what is left */
    u64 fits(u64 n, u64 each) {
        if (each && n > left / each) throw QsfBad{"a count exceeds the file"};
        return n;
    }
};

static sft_model* load_qsf(Context& c, QsfReader& r, sft_arena* arena);

extern "C" sft_model* sft_model_load_qsf(const char* path, sft_arena* arena) {
    Context& c = C();
    FILE* f = fopen(path, "rb");
    if (!f) SFT_ERR_RET(NULL, "failed to open '%s'\n", path);
    fseek(f, 0, SEEK_END);
    const long size = ftell(f);
    fseek(f, 0, SEEK_SET);
    char magic[4];
    if (size < 4 || 4 != fread(magic, 1, 4, f) || 0 != memcmp(magic, "QSF2", 4)) {
        fclose(f);
        SFT_ERR_RET(NULL, "invalid QSF2 magic in '%s'\n", path);
    }
    QsfReader r{f, (u64)size - 4u};
    sft_model* model = nullptr;
    try {
        u32 version = 0;
        r.read(&version, 4);
        if (2 != version) {
            fclose(f);
            SFT_ERR_RET(NULL, "unsupported QSF2 version\n");
        }
        model = load_qsf(c, r, arena);
    }
    catch (const QsfBad& e) {
        fclose(f);
        SFT_ERR_RET(NULL, "corrupt QSF2 file '%s': %s\n", path, e.what);
    }
    fclose(f);
    return model;
}

static sft_model* load_qsf(Context& c, QsfReader& r, sft_arena* arena) {
    u64 game_data_offset = 0;
    r.read(&game_data_offset, 8);

    u32 tex_count = 0;
    r.read(&tex_count, 4);
    r.fits(tex_count, 8);
    sft_mat_handle* mats = (sft_mat_handle*)sft_arena_alloc(arena, (u64)tex_count * sizeof(sft_mat_handle) + 1);
    std::vector<u32> px;
    for (u32 i = 0; i < tex_count; i += 1) {
        u32 w = 0, h = 0;
        r.read(&w, 4);
        r.read(&h, 4);
        px.resize(r.fits((u64)w * h, 4));
        r.read(px.data(), (u64)w * h * 4);
        mats[i] = sft_material_create(sft_texture_create((i32)w, (i32)h, px.data()), 0xFFFFFFFFu, 2.f);
    }

    sft_model* model = (sft_model*)sft_arena_alloc(arena, sizeof(sft_model));
    memset(model, 0, sizeof(sft_model));
    r.read(&model->lightmap_count, 4);
    if (model->lightmap_count > 0) {
        r.fits(model->lightmap_count, 8);
        model->lightmaps = (sft_tex_handle*)sft_arena_alloc(arena, model->lightmap_count * sizeof(sft_tex_handle));
        for (u32 i = 0; i < model->lightmap_count; i += 1) {
            u32 w = 0, h = 0;
            r.read(&w, 4);
            r.read(&h, 4);
            px.resize(r.fits((u64)w * h, 4));
            r.read(px.data(), (u64)w * h * 4);
            model->lightmaps[i] = sft_texture_create((i32)w, (i32)h, px.data());
        }
    }
    r.read(&model->mesh_count, 4);
    r.read(&model->total_meshlets, 4);
    r.read(&model->blas_node_count, 4);
    r.read(&model->blas_root_ind, 4);
    r.fits(model->mesh_count, 16);
    r.fits(model->total_meshlets, 2);
    r.fits(model->blas_node_count, sizeof(sft_bvh_node));
    if (model->blas_node_count > 0 && model->blas_root_ind >= model->blas_node_count) throw QsfBad{"BLAS root outside its nodes"};

    model->meshes = (sft_mesh*)sft_arena_alloc(arena, model->mesh_count * sizeof(sft_mesh));
    memset(model->meshes, 0, model->mesh_count * sizeof(sft_mesh));
    for (u32 i = 0; i < model->mesh_count; i += 1) {
        sft_mesh* mesh = &model->meshes[i];
        r.read(&mesh->meshlet_count, 4);
        r.read(&mesh->model_meshlet_offset, 4);
        u32 tex_ind = 0;
        r.read(&tex_ind, 4);
        mesh->material = (tex_ind < tex_count) ? mats[tex_ind] : c.base_mat;
        r.read(&mesh->lightmap_page, 4);
        if (0xFFFFFFFFu != mesh->lightmap_page && mesh->lightmap_page >= model->lightmap_count) {
            throw QsfBad{"a mesh names a lightmap page the file lacks"};
        }
        const u32 mc = (u32)r.fits(mesh->meshlet_count, sizeof(sft_meshlet) + sizeof(sft_meshlet_attribs) + 384);
        if ((u64)mesh->model_meshlet_offset + mc > model->total_meshlets) throw QsfBad{"a mesh's meshlets lie outside the model"};
        mesh->meshlets = (sft_meshlet*)sft_arena_alloc(arena, (u64)mc * sizeof(sft_meshlet));
        r.read(mesh->meshlets, (u64)mc * sizeof(sft_meshlet));
        mesh->meshlet_attribs = (sft_meshlet_attribs*)sft_arena_alloc(arena, (u64)mc * sizeof(sft_meshlet_attribs));
        r.read(mesh->meshlet_attribs, (u64)mc * sizeof(sft_meshlet_attribs));
        mesh->inds = (sft_meshlet_inds*)sft_arena_alloc(arena, (u64)mc * sizeof(sft_meshlet_inds));
        u8 flat[384];
        for (u32 m = 0; m < mc; m += 1) {
            r.read(flat, 384);
            sft_meshlet_inds* mi = &mesh->inds[m];
            memset(mi, 0, sizeof(*mi));
            for (u32 t = 0; t < 128; t += 1) {
                const u32 off = (t / 8) * 32 + (t % 8);
                mi->inds[off + 0] = flat[t * 3 + 0];
                mi->inds[off + 8] = flat[t * 3 + 1];
                mi->inds[off + 16] = flat[t * 3 + 2];
            }
        }
        f32** groups[] = {
            &mesh->bounds_center_x, &mesh->bounds_center_y, &mesh->bounds_center_z, &mesh->bounds_radius,
            &mesh->bounds_min_x, &mesh->bounds_min_y, &mesh->bounds_min_z,
            &mesh->bounds_max_x, &mesh->bounds_max_y, &mesh->bounds_max_z,
            &mesh->cone_axis_x, &mesh->cone_axis_y, &mesh->cone_axis_z, &mesh->cone_cutoff };
        for (f32** a : groups) {
            *a = (f32*)sft_arena_alloc(arena, (u64)mc * sizeof(f32));
            r.read(*a, (u64)mc * sizeof(f32));
        }
    }
    if (model->total_meshlets > 0) {
        model->meshlet_mesh_inds = (u16*)sft_arena_alloc(arena, model->total_meshlets * sizeof(u16));
        r.read(model->meshlet_mesh_inds, model->total_meshlets * sizeof(u16));
        for (u32 k = 0; k < model->total_meshlets; k += 1) {
            if (model->meshlet_mesh_inds[k] >= model->mesh_count) throw QsfBad{"a meshlet names a mesh the model lacks"};
        }
    }
    if (model->blas_node_count > 0) {
        model->blas_nodes = (sft_bvh_node*)sft_arena_alloc(arena, model->blas_node_count * sizeof(sft_bvh_node));
        r.read(model->blas_nodes, model->blas_node_count * sizeof(sft_bvh_node));
    }
    return model;
}

/* ================================================================ capture + stats */
extern "C" void sftx_set_capture(u32 flags) { C().capture_flags = flags; }
extern "C" void sftx_get_stats(sftx_frame_stats* out) {
    Context& c = C();
    drain_async_batches(c);
    *out = c.stats;
}

extern "C" void sftx_export_frame(sftx_frame_desc* f) {
    Context& c = C();
    memset(f, 0, sizeof(*f));
    f->width = c.width;
    f->height = c.height;
    f->tile_cols = c.tile_cols;
    f->tile_rows = c.tile_rows;
    f->half_width = c.half_width;
    f->half_height = c.half_height;
    f->vp_x = c.vp_x;
    f->vp_y = c.vp_y;
    f->camera_pos = c.camera_pos;
    f->view_proj = c.view_proj;
    memcpy(f->planes, c.planes, sizeof(f->planes));
    f->ambient = c.ambient;
    f->dir_lights = c.dir_lights.data();
    f->n_dir = (i32)c.dir_lights.size();
    f->point_lights = c.point_lights.data();
    f->n_point = (i32)c.point_lights.size();
    f->lightmap_brightness = c.lightmap_brightness;
    f->shadows_enabled = c.shadows_enabled;
    f->should_clear = c.should_clear;
    f->samples = c.samples;
    f->clear_color = c.clear_color;
    f->debug_mode = (i32)c.debug_mode;
    f->instances = c.instances.data();
    f->n_instances = (i32)c.instances.size();
    f->total_ss = c.total_ss;
    c.export_proxies.resize(c.proxies.size());
    for (size_t i = 0; i < c.proxies.size(); i += 1) {
        c.export_proxies[i].posa = c.proxies[i].posa;
        c.export_proxies[i].posb = c.proxies[i].posb;
        c.export_proxies[i].radius = c.proxies[i].radius;
    }
    f->proxies = c.export_proxies.data();
    f->n_proxies = (i32)c.export_proxies.size();
    c.export_tex.resize(c.tex.size());
    for (size_t i = 0; i < c.tex.size(); i += 1) {
        const HostTexture& t = c.tex[i];
        sftx_texture_desc& d = c.export_tex[i];
        memset(&d, 0, sizeof(d));
        d.texels = nullptr;
        d.level0 = t.level0.empty() ? nullptr : t.level0.data();
        d.total = (u32)t.total;
        d.w = t.w;
        d.h = t.h;
        d.mip_levels = t.mip_levels;
        d.is_pot = t.is_pot;
        memcpy(d.mip_offset, t.mip_offset, sizeof(d.mip_offset));
        memcpy(d.blocks_w, t.blocks_w, sizeof(d.blocks_w));
        memcpy(d.stride_shift, t.stride_shift, sizeof(d.stride_shift));
    }
    f->textures = c.export_tex.data();
    f->n_textures = (i32)c.export_tex.size();
    f->mat_tex = c.mat_tex.data();
    f->mat_color = c.mat_color.data();
    f->mat_shininess = c.mat_shininess.data();
    f->n_mat = (i32)c.mat_tex.size();
    f->tmat_tex = c.tmat_tex.data();
    f->tmat_color = c.tmat_color.data();
    f->tmat_shininess = c.tmat_shininess.data();
    f->n_tmat = (i32)c.tmat_tex.size();
    f->base_tex = (c.base_mat != kInvalid) ? c.mat_tex[SFT_MAT_INDEX(c.base_mat)] : SFT_INVALID_HANDLE;
    f->color_buffer = c.color_buffer;
    f->color_pitch = c.color_pitch;
}
