/*
 * backend.cpp — HBM residency and per-frame orchestration of the HIP render
 * path.  sft_present (sofren2.c:8139-8289) becomes: make new models and
 * textures resident, build the frame's instance / work / light / material
 * tables, upload them, run the kernels of kernels.hip on one stream, and copy
 * the ARGB frame into the caller's buffer.  The reference's semaphore hand-off
 * to its worker threads is the host->device boundary here.
 *
 * Capacities of the variable-size intermediates (clip records, tile lists) are
 * not known before the geometry pass; the kernels flag an overflow instead of
 * writing past them and the frame is re-run with larger buffers, so a frame
 * needs no host round trip between passes.
 */
#include <hip/hip_runtime.h>
#include <string.h>
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include "kernels.h"

#define HIP_CHECK(x) do { \
    const hipError_t e_ = (x); \
    if (hipSuccess != e_) { \
        fprintf(stderr, "fatal sofren2 HIP error %s at %s:%d: %s\n", hipGetErrorName(e_), __FILE__, __LINE__, #x); \
        exit(1); \
    } \
} while (0)

namespace sft {

/* Device memory of the no-allocation entry point (sftx_present_noalloc):
   every per-frame intermediate of its frames is carved from this array,
   which each GPU the library runs on holds until the process ends.  Host
   code reaches it only through the address hipGetSymbolAddress returns for
   the device (Backend::scratch), never through the array's name. */
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) u8 scratch[kScratchBytes];

struct DBuf {
    void* p = nullptr;
    size_t cap = 0;
};

/* grows a device buffer to at least `bytes`; `keep` preserves its contents */
static void dgrow(DBuf& b, size_t bytes, bool keep, hipStream_t s) {
    if (bytes <= b.cap) return;
    size_t nc = std::max<size_t>(bytes, b.cap + b.cap / 2);
    nc = std::max<size_t>(nc, 4096);
    void* np = nullptr;
    HIP_CHECK(hipMalloc(&np, nc));
    if (keep && b.p && b.cap) {
        HIP_CHECK(hipMemcpyAsync(np, b.p, b.cap, hipMemcpyDeviceToDevice, s));
        HIP_CHECK(hipStreamSynchronize(s));
    }
    if (b.p) HIP_CHECK(hipFree(b.p));
    b.p = np;
    b.cap = nc;
}

template <class T>
static void upload(DBuf& b, const std::vector<T>& v, hipStream_t s) {
    const size_t bytes = std::max<size_t>(v.size() * sizeof(T), sizeof(T));
    dgrow(b, bytes, false, s);
    if (!v.empty()) HIP_CHECK(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
}

/* pinned host staging for the frame's small tables: one H2D copy per frame */
struct Packer {
    u8* host = nullptr;
    size_t cap = 0, used = 0;
    bool pinned = true;             /* false: ordinary host memory (the no-allocation path) */
    std::vector<std::pair<size_t, size_t>> pending;   /* (offset, bytes) */
    size_t add(const void* src, size_t bytes) {
        const size_t off = (used + 255) & ~(size_t)255;
        used = off + std::max<size_t>(bytes, 4);
        if (used > cap) grow(used);
        if (bytes) memcpy(host + off, src, bytes);
        return off;
    }
    template <class T> size_t add(const std::vector<T>& v) { return add(v.data(), v.size() * sizeof(T)); }
    /* a zeroed section to be filled through host + offset once no later add
       can move the staging memory */
    size_t reserve(size_t bytes) {
        const size_t off = (used + 255) & ~(size_t)255;
        used = off + std::max<size_t>(bytes, 4);
        if (used > cap) grow(used);
        memset(host + off, 0, std::max<size_t>(bytes, 4));
        return off;
    }
    void grow(size_t need) {
        size_t nc = std::max<size_t>(need * 2, 1 << 20);
        u8* nh = nullptr;
        if (pinned) HIP_CHECK(hipHostMalloc((void**)&nh, nc, hipHostMallocDefault));
        else if (!(nh = (u8*)malloc(nc))) {
            fprintf(stderr, "fatal sofren2 error: out of host memory for %zu bytes of batch tables\n", nc);
            exit(1);
        }
        if (host) {
            memcpy(nh, host, cap);
            release();
        }
        host = nh;
        cap = nc;
    }
    void release() {
        if (host && pinned) HIP_CHECK(hipHostFree(host));
        else if (host) free(host);
        host = nullptr;
    }
};

struct ModelRes {
    u32 base = 0;                   /* first slot */
    u32 total = 0;                  /* total_meshlets */
    u32 blas_base = kInvalid;       /* first node in the BLAS pool; kInvalid: none or skinned */
    u32 blas_count = 0, blas_root = 0;
    u32 mesh_rec_base = 0;
    u32 max_bone = 0;               /* bone matrices a pose must provide */
    std::vector<u32> skin_of;       /* skin record per model meshlet, kInvalid if none */
    u64 fp = 0;                     /* model_fingerprint when the rows were uploaded */
    u64 checked = 0;                /* Backend::check_serial of the last comparison */
    bool edited = false;            /* that comparison found another fingerprint */
};

/* Everything one batch in flight owns: its stream and events, the per-batch
   device buffers, the pinned staging, the captured graph, and what the host
   prepared so the batch can be completed (or re-run after an overflow). */
struct FrameSet {
    hipStream_t stream = nullptr;
    hipEvent_t ev[10] = {};
    bool in_flight = false;
    u32 region = 0;                 /* transient region of the pools this set's batches use */
    /* non-null: the per-batch device buffers are carved from this device
       memory (the no-allocation path) instead of being allocated */
    u8* carve_base = nullptr;
    size_t carve_bytes = 0;

    DBuf ss, ss_vis, clip_count, clip_base, clip_mask, batch_pl, bin_m;
    DBuf tile_off, tile_count, tile_list, clip_rec, clip_key, clip_rect, color, vid, tile_lights, cap_depth;
    DBuf proxy_vis, light_proxy_count, light_proxies, tile_z, vis64, scan_sums, raster_items, tile_nch, merge;
    Packer pack;                    /* per-batch tables, staged in pinned memory */
    DBuf pack_d;
    DBuf zero_d;                    /* counters, tile counts, chunk completions */
    u32* counters_h = nullptr;      /* pinned copy of the narrow counters */
    u8* out_h = nullptr;            /* pinned staging of host outputs (colour, 64-bit visibility) */
    size_t out_h_cap = 0;
    u32* band_flag_h = nullptr;     /* pinned per-band flags of a streamed colour download */
    u32 band_flag_cap = 0;
    u32 n_bands = 0;                /* bands (tile rows) of the views with host colour */
    bool streamed = false;          /* the last launch streamed its colour to out_h */
    bool host_copied = false;       /* the copier delivered every host output */
    std::thread copier;             /* an asynchronous batch's host outputs, copied once it is done */
    SideStream side{};              /* shadow-proxy occlusion beside the main passes (eager launches) */
    bool side_used = false;         /* the last launch ran occlusion on the side stream */
    u32 clip_cap = 1u << 16;
    u32 tile_list_cap = 1u << 20;
    hipGraphExec_t graph_exec = nullptr;
    std::vector<u8> graph_key, last_key;

    /* prepared batch */
    BatchGpu bg{};
    std::vector<ViewGpu> vgpu;
    struct ViewPlan {
        u32 inst_base = 0, work_base = 0, n_work = 0, pl_base = 0, dl_base = 0, tmat_base = 0;
        u32 tile_base = 0, proxy_base = 0, tlas_base = 0, occ_base = 0, occ_count = 0;
        u64 vid_base = 0, color_base = 0, grid = 0;
        u64 vis_stage = 0;          /* its 64-bit visibility in the host staging (views that ask for it) */
        u32 n_tiles = 0;
        u32 band_base = kInvalid;   /* its first band when its colour goes to host memory */
    };
    std::vector<ViewPlan> plan;
    u32 nt = 0, nw = 0, max_samples = 1, ss_total = 0, n_proxies = 0, n_lights = 0, t_slots = 0, t_skins = 0;
    u32 t_base = 0, t_skin_base = 0;    /* this set's transient region */
    u64 layout_epoch = 0;           /* Backend::layout_epoch the transient rows are placed for */
    u32 n_insts = 0;                /* instances in the staged tables */
    u64 bin_m_words = 0;
    u64 vid_total = 0, color_total = 0, vis_stage_total = 0;
    size_t zero_words = 0;
    bool any_vis64 = false, cap_vis = false, host_color = false;
    size_t o_batch = 0, o_views = 0, o_tv = 0, o_st = 0, o_insts = 0, o_work = 0, o_bones = 0, o_pl = 0, o_dl = 0;
    size_t o_mt = 0, o_mc = 0, o_ms = 0, o_tt = 0, o_tc = 0, o_ts = 0, o_mr = 0, o_px = 0, o_tlas = 0, o_seg = 0;
    size_t o_clp = 0, o_clpc = 0;
    size_t o_tpos = 0, o_tattr = 0, o_tinds = 0, o_tcull = 0, o_tsm = 0, o_tskin = 0;
    u32 counters[CNT_COUNT] = {};
};

struct Backend {
    /* resident pools */
    DBuf pos, attr, inds, cull, slot_mesh, skin;
    u32 resident_slots = 0, resident_skins = 0, resident_mesh_recs = 0;
    /* pool layout: resident rows in [0, res_cap), then one transient region
       of t_cap rows per frame set (dmesh geometry, rewritten every batch), so
       a batch can stage its transient rows while the other sets' batches run;
       skins alike */
    u32 res_cap = 0, res_skin_cap = 0, t_cap = 0, t_skin_cap = 0;
    /* bumped whenever the pools grow or their layout may move: a prepared
       batch whose epoch is older must be prepared again before a re-run */
    u64 layout_epoch = 0;
    std::vector<MeshRec> mesh_recs;
    std::unordered_map<const sft_model*, ModelRes> models;
    u64 check_serial = 0;           /* one per residency check: fingerprints once per model */
    u64 model_edits = 0;              /* resident models uploaded again after an edit */
    DBuf texels, tex_meta_d, tex_stage;
    u64 texels_used = 0;
    std::vector<TexMeta> tex_meta;
    bool tex_meta_dirty = true;
    DBuf est_rcp, est_rsq;
    bool est_uploaded = false;
    DBuf blas;                      /* resident BLAS nodes of static models */
    u32 blas_used = 0;

    /* settings */
    i32 output_mode = SFTX_OUTPUT_HOST;
    /* payloads per visibility work item, by the view's resolution class
       (sftx_batch_tune chooses them) */
    u32 raster_chunk[SFTX_RES_CLASSES] = { 128, 128, 128 };
    /* row-band split threshold of heavy payloads, by class */
    u32 raster_split[SFTX_RES_CLASSES] = { kRasterJobCost, kRasterJobCost, kRasterJobCost };
    u32 occ_stack = kOccStackEntries;    /* occlusion traversal stack entries per wave */
    int last_set = 0;               /* frame set of the last completed batch */
    int ordered_set = -1;           /* frame set of the last ordered submission */
    u32 raster_groups = 512;        /* persistent visibility workgroups */
    bool use_graphs = true;         /* replay the frame's launches from a HIP graph */
    bool stream_download = true;    /* host colour streamed band by band while the frame shades */
    bool gpu_tlas = true;           /* frame TLAS built on the device (k_tlas) */
    u64 streamed_batches = 0, stream_fallbacks = 0;
    u64 graph_captures = 0;
    u64 reruns = 0;                 /* batches run again after an intermediate overflowed */

    /* frames and batches in flight, one set each; blocking calls use set 0 */
    FrameSet sets[kFrameSets];
    /* the no-allocation path: its buffers are carved from `scratch` (the
       device address of the array above on this device); it uses set 0's
       transient region, and runs only with no frame or batch in flight */
    FrameSet noalloc;
    u8* scratch = nullptr;

    /* host copies of the last frame's intermediates (capture) */
    i32 cap_cols = 0, cap_rows = 0, cap_samples = 1;
    std::vector<u32> h_vid;
    std::vector<f32> h_depth;
    std::vector<u8> h_ss, h_ss_vis, h_clips, h_proxy_vis, h_tlas;
    u32 h_ss_count = 0, h_clip_count = 0, h_proxy_count = 0, h_tlas_count = 0;
};

Backend* backend_create() {
    Backend* b = new Backend();
    int n = 0;
    if (hipSuccess != hipGetDeviceCount(&n) || n < 1) {
        fprintf(stderr, "fatal sofren2 error: no HIP device; the render path runs only on the GPU\n");
        exit(1);
    }
    HIP_CHECK(hipSetDevice(0));
    auto init_set = [](FrameSet& fs) {
        HIP_CHECK(hipStreamCreateWithFlags(&fs.stream, hipStreamNonBlocking));
        for (hipEvent_t& e : fs.ev) HIP_CHECK(hipEventCreate(&e));
        HIP_CHECK(hipStreamCreateWithFlags(&fs.side.stream, hipStreamNonBlocking));
        for (hipEvent_t* e : { &fs.side.fork, &fs.side.join, &fs.side.start, &fs.side.end }) HIP_CHECK(hipEventCreate(e));
        HIP_CHECK(hipHostMalloc((void**)&fs.counters_h, CNT_COUNT * 4, hipHostMallocDefault));
    };
    for (int k = 0; k < kFrameSets; k += 1) {
        init_set(b->sets[k]);
        b->sets[k].region = (u32)k;
    }
    /* the scratch array's address on this device (the array's name is a
       host-side stand-in) */
    void* sp = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
    b->scratch = (u8*)sp;
    init_set(b->noalloc);
    b->noalloc.region = 0;
    b->noalloc.carve_base = b->scratch;
    b->noalloc.carve_bytes = kScratchBytes;
    b->noalloc.pack.pinned = false;
    b->noalloc.clip_cap = 1u << 12;
    b->noalloc.tile_list_cap = 1u << 17;
    b->raster_groups = raster_groups_per_device();
    return b;
}

static void free_buf(DBuf& d) {
    if (d.p) HIP_CHECK(hipFree(d.p));
    d.p = nullptr;
    d.cap = 0;
}

void backend_destroy(Backend* b) {
    if (!b) return;
    for (FrameSet& fs : b->sets) {
        if (fs.copier.joinable()) fs.copier.join();
        HIP_CHECK(hipStreamSynchronize(fs.stream));
    }
    HIP_CHECK(hipStreamSynchronize(b->noalloc.stream));
    for (DBuf* d : { &b->pos, &b->attr, &b->inds, &b->cull, &b->slot_mesh, &b->skin, &b->texels, &b->tex_meta_d, &b->tex_stage,
                     &b->est_rcp, &b->est_rsq, &b->blas }) {
        free_buf(*d);
    }
    for (FrameSet& fs : b->sets) {
        for (DBuf* d : { &fs.ss, &fs.ss_vis, &fs.clip_count, &fs.clip_base, &fs.clip_mask, &fs.batch_pl, &fs.bin_m, &fs.tile_off, &fs.tile_count,
                         &fs.tile_list, &fs.clip_rec, &fs.clip_key, &fs.clip_rect, &fs.color, &fs.vid, &fs.tile_lights,
                         &fs.cap_depth, &fs.proxy_vis, &fs.light_proxy_count, &fs.light_proxies, &fs.tile_z, &fs.vis64,
                         &fs.scan_sums, &fs.raster_items, &fs.tile_nch, &fs.merge, &fs.pack_d, &fs.zero_d }) {
            free_buf(*d);
        }
    }
    /* the no-allocation set's device buffers lie in the scratch array */
    std::vector<FrameSet*> all;
    for (FrameSet& fs : b->sets) all.push_back(&fs);
    all.push_back(&b->noalloc);
    for (FrameSet* fsp : all) {
        FrameSet& fs = *fsp;
        fs.pack.release();
        if (fs.counters_h) HIP_CHECK(hipHostFree(fs.counters_h));
        if (fs.out_h) HIP_CHECK(hipHostFree(fs.out_h));
        if (fs.band_flag_h) HIP_CHECK(hipHostFree(fs.band_flag_h));
        if (fs.graph_exec) HIP_CHECK(hipGraphExecDestroy(fs.graph_exec));
        for (hipEvent_t& e : fs.ev) HIP_CHECK(hipEventDestroy(e));
        HIP_CHECK(hipStreamDestroy(fs.stream));
        for (hipEvent_t e : { fs.side.fork, fs.side.join, fs.side.start, fs.side.end }) HIP_CHECK(hipEventDestroy(e));
        HIP_CHECK(hipStreamDestroy(fs.side.stream));
    }
    delete b;
}

/* waits for every batch in flight: shared pools may move or be rewritten next */
static void quiesce(Backend* b) {
    for (FrameSet& fs : b->sets) HIP_CHECK(hipStreamSynchronize(fs.stream));
}

void backend_release_texture(Backend* b, u32 index) {
    (void)index;
    if (b) b->tex_meta_dirty = true;
}

/* Slots, mesh records and skin records of one model whose meshlets go into
   slots [base, base + total_meshlets) (sofren2.c:8457-8477 order); the rows
   themselves are written by fill_model_rows. */
static void plan_model(const sft_model* m, ModelRes& r, u32 base, u32 mesh_rec_base, std::vector<MeshRec>& recs,
                       u32 skin_base, u32* n_skin_out) {
    const u32 T = m->total_meshlets;
    r.base = base;
    r.total = T;
    r.mesh_rec_base = mesh_rec_base;
    r.skin_of.assign(T, kInvalid);
    r.max_bone = (u32)std::max(m->skeleton.bone_count, 0);
    r.blas_base = kInvalid;
    r.blas_count = 0;
    r.blas_root = 0;
    u32 n_skin = 0;
    for (u32 i = 0; i < m->mesh_count; i += 1) {
        const sft_mesh* mesh = &m->meshes[i];
        MeshRec rec;
        rec.material = mesh->material;
        rec.lightmapped = (0xFFFFFFFFu != mesh->lightmap_page) ? 1u : 0u;
        rec.lightmap_tex = rec.lightmapped && m->lightmaps ? m->lightmaps[mesh->lightmap_page] : SFT_INVALID_HANDLE;
        rec.pad = 0;
        recs.push_back(rec);
        for (u32 k = 0; k < mesh->meshlet_count; k += 1) {
            const u32 gm = mesh->model_meshlet_offset + k;
            if (gm >= T) {
                fprintf(stderr, "fatal sofren2 error: mesh %u meshlet %u lies outside the model's %u meshlets\n", i, k, T);
                exit(1);
            }
            if (mesh->meshlet_skins) {
                const sft_meshlet_skin* sk = &mesh->meshlet_skins[k];
                r.skin_of[gm] = skin_base + n_skin;
                n_skin += 1;
                for (int v = 0; v < 64; v += 1) {
                    const u32 jm = std::max(std::max(sk->j0[v], sk->j1[v]), std::max(sk->j2[v], sk->j3[v]));
                    r.max_bone = std::max(r.max_bone, jm + 1u);
                }
            }
            if (mesh->culling_bone_inds && mesh->culling_bone_inds[k] >= 0) {
                r.max_bone = std::max(r.max_bone, (u32)mesh->culling_bone_inds[k] + 1u);
            }
        }
    }
    *n_skin_out = n_skin;
}

/* writes a planned model's rows into staging arrays whose row 0 is slot
   `row0` (and skin row 0 is skin record `skin_row0`) */
static void fill_model_rows(const sft_model* m, const ModelRes& r, u32 row0, u32 skin_row0, f32* pos, f32* attr,
                            u8* inds, f32* cull, u32* slot_mesh, u8* skin) {
    for (u32 i = 0; i < m->mesh_count; i += 1) {
        const sft_mesh* mesh = &m->meshes[i];
        for (u32 k = 0; k < mesh->meshlet_count; k += 1) {
            const u32 gm = mesh->model_meshlet_offset + k;
            const size_t row = (size_t)(r.base - row0) + gm;
            memcpy(&pos[row * 192], &mesh->meshlets[k], 768);
            memcpy(&attr[row * 448], &mesh->meshlet_attribs[k], 1792);
            memcpy(&inds[row * 512], &mesh->inds[k], 512);
            f32* c = &cull[row * 8];
            c[0] = mesh->bounds_center_x[k];
            c[1] = mesh->bounds_center_y[k];
            c[2] = mesh->bounds_center_z[k];
            c[3] = mesh->bounds_radius[k];
            c[4] = mesh->cone_axis_x[k];
            c[5] = mesh->cone_axis_y[k];
            c[6] = mesh->cone_axis_z[k];
            c[7] = mesh->cone_cutoff[k];
            slot_mesh[row] = r.mesh_rec_base + i;
            if (mesh->meshlet_skins) memcpy(&skin[(size_t)(r.skin_of[gm] - skin_row0) * 768], &mesh->meshlet_skins[k], 768);
        }
    }
}

/* Worker threads kept for the host copies and row fills of every frame (a
   thread started per call costs more than a small frame's copy).  One caller
   at a time owns the pool; a second, concurrent caller starts threads of its
   own.  run(k, fn) calls fn(0) on the caller and fn(w + 1) on worker w for
   w + 1 < k, and returns when all are done; a worker takes no part in a job
   after its own index, so no job outlives its call. */
class WorkerPool {
public:
    /* never destroyed: a batch's copier may still use it while the process
       exits, and its blocked workers end with the process */
    static WorkerPool& get() {
        static WorkerPool* pool = new WorkerPool();
        return *pool;
    }
    size_t size() const { return workers_.size(); }
    bool try_acquire() { return owner_.try_lock(); }
    void release() { owner_.unlock(); }
    void run(size_t k, const std::function<void(size_t)>& fn) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            job_ = &fn;
            n_ = k;
            pending_ = k - 1;
            gen_ += 1;
        }
        cv_.notify_all();
        fn(0);
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [&] { return 0 == pending_; });
        job_ = nullptr;
    }
private:
    WorkerPool() {
        const size_t hw = std::max<size_t>(1u, std::thread::hardware_concurrency());
        for (size_t t = 0; t + 1 < std::min<size_t>(hw, 16u); t += 1) workers_.emplace_back([this, t] { loop(t); });
    }
    void loop(size_t id) {
        u64 seen = 0;
        for (;;) {
            const std::function<void(size_t)>* job;
            size_t n;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return gen_ != seen; });
                seen = gen_;
                job = job_;
                n = n_;
            }
            if (id + 1 < n) {
                (*job)(id + 1);
                std::lock_guard<std::mutex> lk(mu_);
                if (0 == --pending_) done_.notify_all();
            }
        }
    }
    std::vector<std::thread> workers_;
    std::mutex owner_, mu_;
    std::condition_variable cv_, done_;
    const std::function<void(size_t)>* job_ = nullptr;
    size_t n_ = 0, pending_ = 0;
    u64 gen_ = 0;
};

/* fn(i) for i < n on up to max_threads threads (the calling one among them),
   in contiguous runs of about equal weight, each at least min_weight */
template <class W, class F>
static void parallel_runs(size_t n, W weight, F fn, size_t max_threads, size_t min_weight) {
    size_t total = 0;
    for (size_t i = 0; i < n; i += 1) total += weight(i);
    const size_t hw = std::max<size_t>(1u, std::thread::hardware_concurrency());
    const size_t nt = std::min({max_threads, hw, n, std::max<size_t>(1u, total / std::max<size_t>(min_weight, 1u))});
    if (nt < 2) {
        for (size_t i = 0; i < n; i += 1) fn(i);
        return;
    }
    std::vector<size_t> cut(nt + 1, n);
    cut[0] = 0;
    size_t acc = 0, k = 1;
    for (size_t i = 0; i < n && k < nt; i += 1) {
        acc += weight(i);
        if (acc >= total * k / nt) cut[k++] = i + 1;
    }
    auto run = [&](size_t t) {
        for (size_t i = cut[t]; i < cut[t + 1]; i += 1) fn(i);
    };
    WorkerPool& pool = WorkerPool::get();
    if (nt <= pool.size() + 1 && pool.try_acquire()) {
        pool.run(nt, std::function<void(size_t)>(run));
        pool.release();
        return;
    }
    std::vector<std::thread> workers;
    for (size_t t = 1; t < nt; t += 1) workers.emplace_back(run, t);
    run(0);
    for (std::thread& w : workers) w.join();
}

/* grows the shared pools; no batch may be in flight (quiesce first) */
static void grow_pools(Backend* b, size_t slots, size_t skins) {
    hipStream_t s = b->sets[0].stream;
    b->layout_epoch += 1;
    slots = std::max(slots, (size_t)b->res_cap + kFrameSets * (size_t)b->t_cap);
    skins = std::max(skins, (size_t)b->res_skin_cap + kFrameSets * (size_t)b->t_skin_cap);
    dgrow(b->pos, slots * 768, true, s);
    dgrow(b->attr, slots * 1792, true, s);
    dgrow(b->inds, slots * 512, true, s);
    dgrow(b->cull, slots * 32, true, s);
    dgrow(b->slot_mesh, slots * 4, true, s);
    dgrow(b->skin, std::max<size_t>(skins, 1) * 768, true, s);
}

/* makes a model resident: plans it after the resident slots and copies its
   rows (and, for a static model, its BLAS) to HBM */
static void upload_model(Backend* b, const sft_model* m, ModelRes& r) {
    u32 n_skin = 0;
    plan_model(m, r, b->resident_slots, b->resident_mesh_recs, b->mesh_recs, b->resident_skins, &n_skin);
    const u32 T = m->total_meshlets;
    /* the transient regions move up when the resident range grows (callers
       hold no batch in flight here) */
    if (b->resident_slots + T > b->res_cap) b->res_cap = std::max(b->res_cap * 2u, b->resident_slots + T);
    if (b->resident_skins + n_skin > b->res_skin_cap) b->res_skin_cap = std::max(b->res_skin_cap * 2u, b->resident_skins + n_skin);
    grow_pools(b, (size_t)r.base + T, (size_t)b->resident_skins + n_skin);
    hipStream_t s = b->sets[0].stream;
    if (T) {
        std::vector<f32> pos((size_t)T * 192), attr((size_t)T * 448), cull((size_t)T * 8, 0.f);
        std::vector<u8> inds((size_t)T * 512, 0), skin((size_t)std::max(n_skin, 1u) * 768, 0);
        std::vector<u32> slot_mesh(T, r.mesh_rec_base);
        fill_model_rows(m, r, r.base, b->resident_skins, pos.data(), attr.data(), inds.data(), cull.data(), slot_mesh.data(),
                        skin.data());
        HIP_CHECK(hipMemcpyAsync((u8*)b->pos.p + (size_t)r.base * 768, pos.data(), pos.size() * 4, hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemcpyAsync((u8*)b->attr.p + (size_t)r.base * 1792, attr.data(), attr.size() * 4, hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemcpyAsync((u8*)b->inds.p + (size_t)r.base * 512, inds.data(), inds.size(), hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemcpyAsync((u8*)b->cull.p + (size_t)r.base * 32, cull.data(), cull.size() * 4, hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemcpyAsync((u8*)b->slot_mesh.p + (size_t)r.base * 4, slot_mesh.data(), slot_mesh.size() * 4,
                                 hipMemcpyHostToDevice, s));
        if (n_skin) {
            HIP_CHECK(hipMemcpyAsync((u8*)b->skin.p + (size_t)b->resident_skins * 768, skin.data(), (size_t)n_skin * 768,
                                     hipMemcpyHostToDevice, s));
        }
        /* the staging vectors die at scope exit: finish the copies first */
        HIP_CHECK(hipStreamSynchronize(s));
    }
    b->resident_slots += T;
    b->resident_skins += n_skin;
    b->resident_mesh_recs = (u32)b->mesh_recs.size();
    /* BLAS of a static resident model: shadow rays test static geometry only
       (sofren2.c:4130), so skinned models keep none */
    if (m->blas_nodes && m->blas_node_count > 0 && m->skeleton.bone_count <= 0) {
        const u32 nn = m->blas_node_count;
        dgrow(b->blas, ((size_t)b->blas_used + nn) * sizeof(sft_bvh_node), true, s);
        HIP_CHECK(hipMemcpyAsync((u8*)b->blas.p + (size_t)b->blas_used * sizeof(sft_bvh_node), m->blas_nodes,
                                 (size_t)nn * sizeof(sft_bvh_node), hipMemcpyHostToDevice, s));
        HIP_CHECK(hipStreamSynchronize(s));
        r.blas_base = b->blas_used;
        r.blas_count = nn;
        r.blas_root = m->blas_root_ind;
        b->blas_used += nn;
    }
}

/* uploads textures created since the last batch; the caller has quiesced
   when any is pending (texture_changes) */
static bool texture_changes(const Backend* b, const Context& c) {
    if (b->tex_meta_dirty || b->tex_meta.size() != c.tex.size()) return true;
    for (const HostTexture& t : c.tex) {
        if (t.live && !t.uploaded) return true;
    }
    return false;
}

static void sync_textures(Backend* b, Context& c) {
    hipStream_t stream = b->sets[0].stream;
    if (b->tex_meta.size() != c.tex.size()) {
        b->tex_meta.resize(c.tex.size());
        b->tex_meta_dirty = true;
    }
    /* new textures: level 0 to HBM, the chain built there (padding texels
       zero); the texel heap only grows, so a heap offset stays 32-bit */
    for (size_t i = 0; i < c.tex.size(); i += 1) {
        HostTexture& t = c.tex[i];
        if (!t.live || t.uploaded) continue;
        const u64 n = t.total;
        if (b->texels_used + n >= (1ull << 32)) {
            fprintf(stderr, "fatal sofren2 error: texel heap exceeds 2^32 texels\n");
            exit(1);
        }
        dgrow(b->texels, (size_t)(b->texels_used + n) * 4, true, stream);
        dgrow(b->tex_stage, t.level0.size() * 4, false, stream);
        HIP_CHECK(hipMemcpyAsync(b->tex_stage.p, t.level0.data(), t.level0.size() * 4, hipMemcpyHostToDevice, stream));
        HIP_CHECK(hipMemsetAsync((u32*)b->texels.p + b->texels_used, 0, n * 4, stream));
        t.dev_offset = b->texels_used;
        t.uploaded = true;
        b->texels_used += n;
        TexMeta& m = b->tex_meta[i];
        memset(&m, 0, sizeof(m));
        m.w = t.w;
        m.h = t.h;
        m.mip_levels = t.mip_levels;
        m.is_pot = t.is_pot;
        for (int l = 0; l < 14; l += 1) {
            m.mip_offset[l] = (u32)(t.dev_offset + t.mip_offset[l]);
            m.blocks_w[l] = t.blocks_w[l];
            m.stride_shift[l] = t.stride_shift[l];
        }
        launch_texture_build((const u32*)b->tex_stage.p, (u32*)b->texels.p, m, stream);
        /* the staging buffer is reused by the next texture */
        HIP_CHECK(hipStreamSynchronize(stream));
        b->tex_meta_dirty = true;
    }
    if (b->tex_meta_dirty) {
        HIP_CHECK(hipStreamSynchronize(stream));
        upload(b->tex_meta_d, b->tex_meta, stream);
        b->tex_meta_dirty = false;
    }
    dgrow(b->texels, 4, false, stream);
}

static f32 ev_ms(hipEvent_t a, hipEvent_t z) {
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, a, z));
    return ms;
}

/* Runs the batch's kernels.  With graphs on, the launch sequence is captured
   into a HIP graph the second time in a row it is seen and replayed while
   the kernel arguments and launch sizes stay the same (per-frame data
   changes only buffer contents).  Per-pass timings need graphs off.  Eager
   launches run the shadow-proxy occlusion on the set's side stream; a
   captured graph keeps every kernel on the main stream. */
static void run_frame(Backend* b, FrameSet& fs, const FramePtrs& p) {
    hipStream_t s = fs.stream;
    fs.side_used = false;
    if (!b->use_graphs) {
        launch_frame(p, fs.bg, fs.zero_words, s, fs.ev, &fs.side);
        HIP_CHECK(hipGetLastError());
        fs.side_used = fs.bg.occ_total > 0u;
        return;
    }
    std::vector<u8> key(sizeof(p) + sizeof(fs.bg) + sizeof(fs.zero_words));
    memcpy(key.data(), &p, sizeof(p));
    memcpy(key.data() + sizeof(p), &fs.bg, sizeof(fs.bg));
    memcpy(key.data() + sizeof(p) + sizeof(fs.bg), &fs.zero_words, sizeof(fs.zero_words));
    if (!fs.graph_exec || key != fs.graph_key) {
        if (key != fs.last_key) {
            fs.last_key = key;
            launch_frame(p, fs.bg, fs.zero_words, s, fs.ev, &fs.side);
            HIP_CHECK(hipGetLastError());
            fs.side_used = fs.bg.occ_total > 0u;
            return;
        }
        if (fs.graph_exec) HIP_CHECK(hipGraphExecDestroy(fs.graph_exec));
        fs.graph_exec = nullptr;
        hipGraph_t g = nullptr;
        HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        launch_frame(p, fs.bg, fs.zero_words, s, nullptr);
        HIP_CHECK(hipStreamEndCapture(s, &g));
        HIP_CHECK(hipGraphInstantiate(&fs.graph_exec, g, nullptr, nullptr, 0));
        HIP_CHECK(hipGraphDestroy(g));
        fs.graph_key = key;
        b->graph_captures += 1;
    }
    /* pass boundaries are not visible inside the graph: the passes read as
       one span between the first and the last mark */
    HIP_CHECK(hipEventRecord(fs.ev[1], s));
    HIP_CHECK(hipGraphLaunch(fs.graph_exec, s));
    for (int k : { 7, 2, 3, 6, 4 }) HIP_CHECK(hipEventRecord(fs.ev[k], s));
}

/* resolution class of a view: below 512x512 pixels, below 1024x1024, larger */
int res_class(i32 width, i32 height) {
    const i64 px = (i64)width * height;
    return (px < (1 << 18)) ? 0 : (px < (1 << 20)) ? 1 : 2;
}

static bool is_transient(const ViewRecord& v, const sft_model* m) {
    return std::find(v.transient_models.begin(), v.transient_models.end(), m) != v.transient_models.end();
}

/* What the draw-time check of a resident model compares: the model header,
   every mesh header except its material and lightmap page (both read at
   every present anyway), and the first eight vertex positions of each
   mesh's first and last meshlet.  The reference reads a model's arrays at
   every present (sofren2.c:8139-8289); rows kept in HBM follow that only as
   far as this check sees edits, and other in-place edits are declared with
   sftx_model_changed. */
static u64 model_fingerprint(const sft_model* m) {
    u64 h = 0x243F6A8885A308D3ull;
    auto mix = [&h](u64 w) {
        h ^= w * 0x9E3779B97F4A7C15ull;
        h = ((h << 27) | (h >> 37)) * 0xC2B2AE3D27D4EB4Full;
    };
    auto ptr = [&mix](const void* p) { mix((u64)(uintptr_t)p); };
    ptr(m->meshes);
    mix(((u64)m->mesh_count << 32) | m->total_meshlets);
    ptr(m->meshlet_mesh_inds);
    ptr(m->blas_nodes);
    mix(((u64)m->blas_node_count << 32) | m->blas_root_ind);
    mix((u64)(u32)m->skeleton.bone_count);
    for (u32 i = 0; i < m->mesh_count && m->meshes; i += 1) {
        const sft_mesh& s = m->meshes[i];
        for (const void* p : { (const void*)s.meshlets, (const void*)s.meshlet_attribs, (const void*)s.meshlet_skins,
                               (const void*)s.inds, (const void*)s.bounds_center_x, (const void*)s.bounds_center_y,
                               (const void*)s.bounds_center_z, (const void*)s.bounds_radius, (const void*)s.bounds_min_x,
                               (const void*)s.bounds_max_x, (const void*)s.bounds_min_y, (const void*)s.bounds_max_y,
                               (const void*)s.bounds_min_z, (const void*)s.bounds_max_z, (const void*)s.cone_axis_x,
                               (const void*)s.cone_axis_y, (const void*)s.cone_axis_z, (const void*)s.cone_cutoff,
                               (const void*)s.culling_bone_inds }) {
            ptr(p);
        }
        mix(((u64)s.meshlet_count << 32) | s.model_meshlet_offset);
        if (!s.meshlets || 0u == s.meshlet_count) continue;
        for (const sft_meshlet* ml : { &s.meshlets[0], &s.meshlets[s.meshlet_count - 1u] }) {
            for (const f32* a : { ml->x, ml->y, ml->z }) {
                u64 w[4];
                memcpy(w, a, sizeof(w));
                for (u64 x : w) mix(x);
            }
        }
    }
    return h ^ (h >> 29);
}

/* whether a model drawn as resident must be uploaded (again): not resident,
   another meshlet count, or another fingerprint (computed once per model
   and check_serial) */
static bool needs_upload(Backend* b, const sft_model* m) {
    auto it = b->models.find(m);
    if (it == b->models.end() || it->second.total != m->total_meshlets) return true;
    ModelRes& r = it->second;
    if (r.checked != b->check_serial) {
        r.checked = b->check_serial;
        r.edited = r.fp != model_fingerprint(m);
    }
    return r.edited;
}

/* Whether preparing these views changes state the frame sets share, and so
   waits for every batch in flight and may allocate: estimate tables or
   textures to upload, models to make resident, or dmesh rows beyond the
   transient regions (their row and skin counts go to need_t, need_tskin). */
static bool shared_state_changes(Backend* b, const Context& c, const std::vector<ViewRecord>& views, u32& need_t,
                                 u32& need_tskin) {
    bool change = !b->est_uploaded || texture_changes(b, c) || b->texels.cap < 4;
    b->check_serial += 1;
    need_t = 0;
    need_tskin = 0;
    for (const ViewRecord& v : views) {
        for (const sft_model* m : v.transient_models) {
            need_t += m->total_meshlets;
            for (u32 i = 0; i < m->mesh_count; i += 1) {
                if (m->meshes[i].meshlet_skins) need_tskin += m->meshes[i].meshlet_count;
            }
        }
    }
    if (need_t > b->t_cap || need_tskin > b->t_skin_cap) change = true;
    for (const ViewRecord& v : views) {
        for (const sft_instance& in : v.instances) {
            if (!is_transient(v, in.model) && needs_upload(b, in.model)) change = true;
        }
    }
    return change;
}

/* Host side of a batch: residency, then every view's rows of the batch
   tables packed into the set's pinned staging. */
static void prepare_batch(Backend* b, Context& c, std::vector<ViewRecord>& views, FrameSet& fs) {
    const u32 nv = (u32)views.size();
    /* shared pools change only with no batch in flight; transient rows need
       no quiesce unless the regions must grow */
    u32 need_t = 0, need_tskin = 0;
    if (shared_state_changes(b, c, views, need_t, need_tskin)) quiesce(b);
    hipStream_t s0 = b->sets[0].stream;
    if (!b->est_uploaded) {
        upload(b->est_rcp, c.est.rcp, s0);
        upload(b->est_rsq, c.est.rsq, s0);
        HIP_CHECK(hipStreamSynchronize(s0));
        b->est_uploaded = true;
    }

    /* ---- residency: new resident models first, then each view's transient ones ---- */
    for (const ViewRecord& v : views) {
        for (const sft_instance& in : v.instances) {
            const sft_model* m = in.model;
            if (is_transient(v, m) || !needs_upload(b, m)) continue;
            /* an edited model gets new rows; its old rows stay unused until
               sft_release (batches in flight were waited for above) */
            if (b->models.count(m)) b->model_edits += 1;
            ModelRes r;
            b->mesh_recs.resize(b->resident_mesh_recs);
            upload_model(b, m, r);
            r.fp = model_fingerprint(m);
            r.checked = b->check_serial;
            b->models[m] = std::move(r);
        }
    }
    b->mesh_recs.resize(b->resident_mesh_recs);
    /* materials and lightmap pages are read at present time, as the reference
       does: refresh the records of the resident models drawn this batch */
    for (const ViewRecord& v : views) {
        for (const sft_instance& in : v.instances) {
            const sft_model* m = in.model;
            if (is_transient(v, m)) continue;
            const ModelRes& r = b->models[m];
            for (u32 i = 0; i < m->mesh_count; i += 1) {
                const sft_mesh* mesh = &m->meshes[i];
                MeshRec& rec = b->mesh_recs[r.mesh_rec_base + i];
                rec.material = mesh->material;
                rec.lightmapped = (0xFFFFFFFFu != mesh->lightmap_page) ? 1u : 0u;
                rec.lightmap_tex = rec.lightmapped && m->lightmaps ? m->lightmaps[mesh->lightmap_page] : SFT_INVALID_HANDLE;
            }
        }
    }
    /* transient models (dmesh geometry) take the slots after the resident
       ones; their rows are staged with the batch tables and copied in bulk */
    std::vector<std::unordered_map<const sft_model*, ModelRes>> tmodels(nv);
    if (need_t > b->t_cap || need_tskin > b->t_skin_cap) {
        b->t_cap = std::max(b->t_cap, std::max(need_t, b->t_cap * 2u));
        b->t_skin_cap = std::max(b->t_skin_cap, std::max(need_tskin, b->t_skin_cap * 2u));
        grow_pools(b, 0, 0);
    }
    fs.t_base = b->res_cap + fs.region * b->t_cap;
    fs.t_skin_base = b->res_skin_cap + fs.region * b->t_skin_cap;
    u32 tslot = fs.t_base, tskin = fs.t_skin_base;
    for (u32 vi = 0; vi < nv; vi += 1) {
        for (const sft_model* m : views[vi].transient_models) {
            ModelRes r;
            u32 n_skin = 0;
            plan_model(m, r, tslot, (u32)b->mesh_recs.size(), b->mesh_recs, tskin, &n_skin);
            tslot += m->total_meshlets;
            tskin += n_skin;
            tmodels[vi][m] = std::move(r);
        }
    }
    fs.t_slots = tslot - fs.t_base;
    fs.t_skins = tskin - fs.t_skin_base;
    if (fs.t_slots > b->t_cap || fs.t_skins > b->t_skin_cap) {
        fprintf(stderr, "fatal sofren2 error: transient rows exceed their region\n");
        exit(1);
    }
    if (texture_changes(b, c)) sync_textures(b, c);
    dgrow(b->texels, 4, false, s0);

    /* ---- batch tables: views' rows back to back ---- */
    std::vector<InstGpu> insts;
    std::vector<WorkItem> work;
    struct WorkFill {
        const ModelRes* r;
        const sft_model* m;
        u32 ss0, inst, view, ginst, first, count;
        bool posed;
    };
    std::vector<WorkFill> work_fills;
    u32 n_work_total = 0;
    std::vector<f32> bones;
    std::vector<PointLightGpu> pl;
    std::vector<DirLightGpu> dl;
    std::vector<u32> tmat_tex, tmat_color;
    std::vector<f32> tmat_shin;
    std::vector<ProxyGpu> proxies;
    std::vector<sft_bvh_node> tlas;
    std::vector<ProxyRec> cached_lp;    /* capsule records of views served from a template's shadow cache */
    std::vector<u32> cached_lpc;
    std::vector<u32> tile_view, shade_tiles;
    std::vector<BinSeg> segs;
    u64 bin_m_words = 0;
    fs.vgpu.assign(nv, ViewGpu{});
    fs.plan.assign(nv, FrameSet::ViewPlan{});
    u32 ss_total = 0, max_samples = 1, occ_total = 0, min_chunk = kRasterMaxChunk;
    u64 vid_total = 0, color_total = 0, vis_stage_total = 0;
    bool any_vis64 = false, host_color = false, any_tlas_dev = false;
    u32 n_bands = 0;
    for (u32 vi = 0; vi < nv; vi += 1) {
        const ViewRecord& v = views[vi];
        FrameSet::ViewPlan& vp = fs.plan[vi];
        vp.inst_base = (u32)insts.size();
        vp.work_base = n_work_total;
        vp.pl_base = (u32)pl.size();
        vp.dl_base = (u32)dl.size();
        vp.tmat_base = (u32)tmat_tex.size();
        vp.proxy_base = (u32)proxies.size();
        vp.tlas_base = (u32)tlas.size();
        vp.tile_base = (u32)tile_view.size();
        vp.n_tiles = (u32)(v.tile_cols * v.tile_rows);
        vp.grid = (u64)v.tile_cols * 64 * v.tile_rows * 64 * (u64)v.samples;
        vp.vid_base = vid_total;
        vp.color_base = color_total;
        vid_total += vp.grid;
        color_total += (u64)v.width * v.height;
        max_samples = std::max<u32>(max_samples, (u32)v.samples);
        any_vis64 = any_vis64 || v.vis64;
        const bool view_host = (SFTX_OUTPUT_HOST == b->output_mode || v.host_output) && v.color_buffer;
        host_color = host_color || view_host;
        if (view_host && v.width > 0 && v.height > 0) {
            vp.band_base = n_bands;
            n_bands += (u32)v.tile_rows;
        }
        for (u32 t = 0; t < vp.n_tiles; t += 1) tile_view.push_back(vi);
        const u32 ss_base = ss_total;
        ss_total += v.total_ss;

        for (u32 ii = 0; ii < v.instances.size(); ii += 1) {
            const sft_instance& in = v.instances[ii];
            const sft_model* m = in.model;
            auto ti = tmodels[vi].find(m);
            const ModelRes& r = (ti != tmodels[vi].end()) ? ti->second : b->models[m];
            InstGpu g;
            memcpy(g.t, &in.transform, 64);
            const sftmat inv = sft_m_qinv(in.transform);
            memcpy(g.inv, &inv, 64);
            g.model_base = r.base;
            g.ss_offset = ss_base + in.ss_meshlet_offset;
            g.bone_base = kInvalid;
            g.blas_base = r.blas_base;
            g.blas_root = r.blas_root;
            g.blas_count = r.blas_count;
            g.n_meshlets = r.total;
            /* the model-space box of the BLAS root's children: the TLAS item
               of a model with a BLAS (sofren2.c:3630-3660) */
            g.tlas_item = (m->blas_node_count > 0 && m->blas_nodes) ? 1u : 0u;
            for (int k = 0; k < 3; k += 1) {
                g.root_lo[k] = g.tlas_item ? 1e30f : 0.f;
                g.root_hi[k] = g.tlas_item ? -1e30f : 0.f;
            }
            if (g.tlas_item) {
                const sft_bvh_node& rn = m->blas_nodes[m->blas_root_ind];
                for (u32 c = 0; c < rn.child_count; c += 1) {
                    const f32 clo[3] = { rn.min_x[c], rn.min_y[c], rn.min_z[c] };
                    const f32 chi[3] = { rn.max_x[c], rn.max_y[c], rn.max_z[c] };
                    for (int k = 0; k < 3; k += 1) {
                        if (clo[k] < g.root_lo[k]) g.root_lo[k] = clo[k];
                        if (chi[k] > g.root_hi[k]) g.root_hi[k] = chi[k];
                    }
                }
            }
            if (in.posed_bone_matrices) {
                g.bone_base = (u32)(bones.size() / 16);
                const u32 n = std::max(r.max_bone, 1u);
                /* a batched view kept the skeleton's matrices; joints beyond
                   them (a model whose skin names more joints than its
                   skeleton has) read identity */
                const u32 kept = v.bones_kept.empty() ? n : std::min(n, v.bones_kept[ii]);
                bones.insert(bones.end(), (const f32*)in.posed_bone_matrices, (const f32*)in.posed_bone_matrices + (size_t)kept * 16);
                const sftmat id = sft_m_identity();
                for (u32 k = kept; k < n; k += 1) bones.insert(bones.end(), (const f32*)&id, (const f32*)&id + 16);
            }
            const u32 ginst = (u32)insts.size();
            insts.push_back(g);
            /* the instance's work items (one per meshlet, mesh by mesh) are
               filled below, in parallel, from this record */
            u32 n = 0;
            for (u32 mi = 0; mi < m->mesh_count; mi += 1) n += m->meshes[mi].meshlet_count;
            work_fills.push_back(WorkFill{&r, m, ss_base + in.ss_meshlet_offset, ii, vi, ginst, n_work_total, n,
                                          nullptr != in.posed_bone_matrices});
            n_work_total += n;
        }
        vp.n_work = n_work_total - vp.work_base;
        for (const sft_point_light& L : v.point_lights) {
            pl.push_back(PointLightGpu{{L.pos.x, L.pos.y, L.pos.z}, {L.color.x, L.color.y, L.color.z}, L.radius, L.inv_radius_sqr});
        }
        for (const sft_directional_light& L : v.dir_lights) {
            dl.push_back(DirLightGpu{{L.dir.x, L.dir.y, L.dir.z}, {L.color.x, L.color.y, L.color.z}, {0.f, 0.f}});
        }
        tmat_tex.insert(tmat_tex.end(), v.tmat_tex.begin(), v.tmat_tex.end());
        tmat_color.insert(tmat_color.end(), v.tmat_color.begin(), v.tmat_color.end());
        tmat_shin.insert(tmat_shin.end(), v.tmat_shininess.begin(), v.tmat_shininess.end());
        for (const WorldProxy& w : v.proxies) {
            proxies.push_back(ProxyGpu{{w.posa.x, w.posa.y, w.posa.z}, {w.posb.x, w.posb.y, w.posb.z}, w.radius, 0.f});
        }
        /* frame TLAS when the occlusion pass runs (sofren2.c:8161-8164); a
           served template's view whose capsule records were read back from
           an earlier batch skips both and uploads them */
        const bool shadowed = v.shadows_enabled && !v.point_lights.empty() && !v.proxies.empty();
        const bool cached = shadowed && v.shadow_cache && v.shadow_cache->filled &&
                            v.shadow_cache->counts.size() == v.point_lights.size();
        const bool occlusion = shadowed && !cached;
        u32 lp_cache = kInvalid;
        if (cached) {
            lp_cache = (u32)cached_lpc.size();
            cached_lp.insert(cached_lp.end(), v.shadow_cache->recs.begin(), v.shadow_cache->recs.end());
            cached_lpc.insert(cached_lpc.end(), v.shadow_cache->counts.begin(), v.shadow_cache->counts.end());
        }
        std::vector<sft_bvh_node> vt;
        u32 tlas_root = 0;
        bool tlas_dev = false;
        if (occlusion && b->gpu_tlas && v.instances.size() <= 64) {
            /* built on the device by k_tlas: reserve its nodes */
            u32 items = 0;
            for (const sft_instance& in : v.instances) items += (in.model->blas_node_count > 0 && in.model->blas_nodes) ? 1u : 0u;
            vt.assign(tlas_node_count(items), sft_bvh_node{});
            tlas_dev = true;
            any_tlas_dev = true;
        }
        else if (occlusion) {
            tlas_root = build_tlas(v.instances, vt);
        }
        tlas.insert(tlas.end(), vt.begin(), vt.end());

        FrameConst& fc = fs.vgpu[vi].fc;
        memset(&fc, 0, sizeof(fc));
        fc.width = v.width;
        fc.height = v.height;
        fc.screen_w = (i32)(v.half_width * 2.f);
        fc.screen_h = (i32)(v.half_height * 2.f);
        fc.tile_cols = v.tile_cols;
        fc.tile_rows = v.tile_rows;
        fc.n_tiles = v.tile_cols * v.tile_rows;
        fc.samples = v.samples;
        fc.hw = v.half_width;
        fc.hh = v.half_height;
        fc.vp_x = v.vp_x;
        fc.vp_y = v.vp_y;
        fc.bias_x = v.half_width + v.vp_x;
        fc.bias_y = v.half_height + v.vp_y;
        fc.neg_hh = -v.half_height;
        memcpy(fc.vp, &v.view_proj, 64);
        memcpy(fc.planes, v.planes, sizeof(fc.planes));
        fc.cam[0] = v.camera_pos.x; fc.cam[1] = v.camera_pos.y; fc.cam[2] = v.camera_pos.z;
        fc.ambient[0] = v.ambient.x; fc.ambient[1] = v.ambient.y; fc.ambient[2] = v.ambient.z;
        fc.clear_color = v.clear_color;
        fc.should_clear = v.should_clear;
        fc.debug_mode = (u32)v.debug_mode;
        fc.shadows_enabled = v.shadows_enabled;
        fc.lm_brightness = v.lightmap_brightness;
        fc.n_work = vp.n_work;
        fc.n_inst = (u32)v.instances.size();
        fc.n_point = (u32)v.point_lights.size();
        fc.n_dir = (u32)v.dir_lights.size();
        fc.n_proxies = (u32)v.proxies.size();
        fc.base_tex = (c.base_mat != kInvalid) ? c.mat_tex[SFT_MAT_INDEX(c.base_mat)] : SFT_INVALID_HANDLE;
        fc.rcp_bits = (u32)c.est.rcp_bits;
        fc.rsq_bits = (u32)c.est.rsq_bits;
        fc.capture = v.capture_flags;
        fc.n_tex = (u32)c.tex.size();
        fc.n_mat = (u32)c.mat_tex.size();
        fc.n_tmat = (u32)v.tmat_tex.size();
        fc.tlas_root = tlas_root;
        fc.tlas_nodes = (u32)vt.size();
        fc.occ_lights = (occlusion && !vt.empty()) ? (u32)std::min<size_t>(v.point_lights.size(), 64) : 0u;
        fc.occ_batches = (u32)((v.proxies.size() + 7) / 8);
        vp.occ_base = occ_total;
        vp.occ_count = fc.occ_lights * fc.occ_batches;
        occ_total += vp.occ_count;

        ViewGpu& g = fs.vgpu[vi];
        g.inst_base = vp.inst_base;
        g.work_base = vp.work_base;
        g.n_work = vp.n_work;
        g.pl_base = vp.pl_base;
        g.dl_base = vp.dl_base;
        g.tmat_base = vp.tmat_base;
        g.tile_base = vp.tile_base;
        g.proxy_base = vp.proxy_base;
        g.tlas_base = vp.tlas_base;
        g.occ_base = vp.occ_base;
        g.occ_count = vp.occ_count;
        /* binning segments: the view's batch slots in runs of kSegItems,
           then its clip slots; their count rows are consecutive */
        const u32 n_items = vp.n_work * 16u;
        const u32 n_bseg = (n_items + kSegItems - 1u) / kSegItems;
        g.seg_base = (u32)segs.size();
        g.n_segs = n_bseg + 1u;
        for (u32 k = 0; k <= n_bseg; k += 1) {
            BinSeg sg;
            sg.view = vi;
            sg.first = vp.work_base * 16u + k * kSegItems;
            sg.count = (k < n_bseg) ? std::min(kSegItems, n_items - k * kSegItems) : kInvalid;
            sg.m_base = (u32)(bin_m_words + (u64)k * vp.n_tiles);
            segs.push_back(sg);
        }
        bin_m_words += (u64)(n_bseg + 1u) * vp.n_tiles;
        g.vis64 = v.vis64 ? 1u : 0u;
        g.raster_chunk = std::max(1u, std::min(b->raster_chunk[res_class(v.width, v.height)], kRasterMaxItem));
        min_chunk = std::min(min_chunk, g.raster_chunk);
        g.band_base = kInvalid;
        g.raster_split = std::max(1u, b->raster_split[res_class(v.width, v.height)]);
        g.tlas_build = tlas_dev ? 1u : 0u;
        g.lp_cache = lp_cache;
        g.id_shift = (u32)SFT_TRIANGLE_BITS + 24u - v.instance_bits;
        if (v.vis64) {
            vp.vis_stage = vis_stage_total;
            vis_stage_total += vp.grid;
        }
        g.vid_base = vp.vid_base;
        g.color_base = vp.color_base;
    }
    /* work items in the reference's job order: views, their instances, each
       instance's meshes and meshlets; instances filled by up to 8 threads */
    work.resize(n_work_total);
    parallel_runs(work_fills.size(), [&](size_t i) { return (size_t)work_fills[i].count; },
                  [&](size_t i) {
                      const WorkFill& f = work_fills[i];
                      WorkItem* out = work.data() + f.first;
                      for (u32 mi = 0; mi < f.m->mesh_count; mi += 1) {
                          const sft_mesh* mesh = &f.m->meshes[mi];
                          for (u32 k = 0; k < mesh->meshlet_count; k += 1) {
                              WorkItem w;
                              w.gm = mesh->model_meshlet_offset + k;
                              w.slot = f.r->base + w.gm;
                              w.ss = f.ss0 + w.gm;
                              w.inst = f.inst;
                              w.cull_bone = (mesh->culling_bone_inds && f.posed) ? mesh->culling_bone_inds[k & ~7u] : -1;
                              w.skin = (mesh->meshlet_skins && f.posed) ? f.r->skin_of[w.gm] : kInvalid;
                              w.view = f.view;
                              w.ginst = f.ginst;
                              *out++ = w;
                          }
                      }
                  }, 8, 16384);
    if (work.size() >= (1u << 23)) {
        fprintf(stderr, "fatal sofren2 error: %zu meshlets in one batch exceed the 2^23 emission keys\n", work.size());
        exit(1);
    }
    if (bin_m_words >= (1ull << 32)) {
        fprintf(stderr, "fatal sofren2 error: binning matrix of %llu words exceeds 32-bit offsets\n",
                (unsigned long long)bin_m_words);
        exit(1);
    }
    const u32 nt = (u32)tile_view.size();
    /* shaded tiles in four runs: without MSAA, then with; within each, views
       without capsule shadows first (the shading kernels are specialised) */
    u32 n_shade_runs[4] = {0, 0, 0, 0};
    for (int run = 0; run < 4; run += 1) {
        for (u32 vi = 0; vi < nv; vi += 1) {
            const FrameConst& f = fs.vgpu[vi].fc;
            const bool ms = f.samples > 1, sh = f.shadows_enabled && f.n_proxies > 0;
            if ((run >> 1) != (ms ? 1 : 0) || (run & 1) != (sh ? 1 : 0)) continue;
            for (u32 t = 0; t < fs.plan[vi].n_tiles; t += 1) shade_tiles.push_back(fs.plan[vi].tile_base + t);
            n_shade_runs[run] += fs.plan[vi].n_tiles;
        }
    }
    /* proxy lists exist for the lights of views with shadows; the kernel skips the others */
    u32 n_lights_cache = 0;
    for (u32 vi = 0; vi < nv; vi += 1) {
        if (fs.vgpu[vi].fc.shadows_enabled && fs.vgpu[vi].fc.n_proxies > 0) n_lights_cache = (u32)pl.size();
    }
    BatchGpu& bg = fs.bg;
    memset(&bg, 0, sizeof(bg));
    bg.n_views = nv;
    bg.n_work = (u32)work.size();
    bg.n_tiles = nt;
    bg.max_samples = max_samples;
    for (u32 vi = 0; vi < nv; vi += 1) bg.n_views_ms += (fs.vgpu[vi].fc.samples > 1) ? 1u : 0u;
    bg.rcp_bits = (u32)c.est.rcp_bits;
    bg.rsq_bits = (u32)c.est.rsq_bits;
    bg.n_shade_tiles = n_shade_runs[0] + n_shade_runs[1];
    bg.n_shade_tiles_ms = n_shade_runs[2] + n_shade_runs[3];
    bg.n_shade_sh = n_shade_runs[1];
    bg.n_shade_ms_sh = n_shade_runs[3];
    bg.occ_total = occ_total;
    bg.n_lights = n_lights_cache;
    bg.vis64 = any_vis64 ? 1u : 0u;
    bg.raster_chunk = min_chunk;   /* sizes the work-item list */
    bg.raster_groups = b->raster_groups;
    bg.occ_stack = b->occ_stack;
    bg.tlas_gpu = any_tlas_dev ? 1u : 0u;
    bg.n_segs = (u32)segs.size();
    fs.bin_m_words = bin_m_words;

    /* ---- staging ---- */
    Packer& pk = fs.pack;
    pk.used = 0;
    fs.o_batch = pk.add(&bg, sizeof(bg));
    fs.o_views = pk.add(fs.vgpu);
    fs.o_tv = pk.add(tile_view);
    fs.o_st = pk.add(shade_tiles);
    fs.o_insts = pk.add(insts);
    fs.o_work = pk.add(work);
    fs.o_bones = pk.add(bones);
    fs.o_pl = pk.add(pl);
    fs.o_dl = pk.add(dl);
    fs.o_mt = pk.add(c.mat_tex);
    fs.o_mc = pk.add(c.mat_color);
    fs.o_ms = pk.add(c.mat_shininess);
    fs.o_tt = pk.add(tmat_tex);
    fs.o_tc = pk.add(tmat_color);
    fs.o_ts = pk.add(tmat_shin);
    fs.o_mr = pk.add(b->mesh_recs);
    fs.o_px = pk.add(proxies);
    fs.o_tlas = pk.add(tlas);
    fs.o_seg = pk.add(segs);
    fs.o_clp = pk.add(cached_lp);
    fs.o_clpc = pk.add(cached_lpc);
    /* transient rows, filled in place once every section is reserved */
    fs.o_tpos = pk.reserve((size_t)fs.t_slots * 768);
    fs.o_tattr = pk.reserve((size_t)fs.t_slots * 1792);
    fs.o_tinds = pk.reserve((size_t)fs.t_slots * 512);
    fs.o_tcull = pk.reserve((size_t)fs.t_slots * 32);
    fs.o_tsm = pk.reserve((size_t)fs.t_slots * 4);
    fs.o_tskin = pk.reserve((size_t)fs.t_skins * 768);
    std::vector<std::pair<const sft_model*, const ModelRes*>> fills;
    for (u32 vi = 0; vi < nv; vi += 1) {
        for (const auto& kv : tmodels[vi]) fills.emplace_back(kv.first, &kv.second);
    }
    /* disjoint rows: filled by up to 8 threads, 512 meshlets (1.5 MB) or more each */
    parallel_runs(fills.size(), [&](size_t i) { return (size_t)fills[i].first->total_meshlets; },
                  [&](size_t i) {
                      fill_model_rows(fills[i].first, *fills[i].second, fs.t_base, fs.t_skin_base,
                                      (f32*)(pk.host + fs.o_tpos), (f32*)(pk.host + fs.o_tattr), pk.host + fs.o_tinds,
                                      (f32*)(pk.host + fs.o_tcull), (u32*)(pk.host + fs.o_tsm), pk.host + fs.o_tskin);
                  }, 8, 512);

    fs.nt = nt;
    fs.nw = bg.n_work;
    fs.n_insts = (u32)insts.size();
    fs.max_samples = max_samples;
    fs.ss_total = ss_total;
    fs.n_proxies = (u32)proxies.size();
    fs.n_lights = (u32)pl.size();
    fs.vid_total = vid_total;
    fs.vis_stage_total = vis_stage_total;
    fs.color_total = color_total;
    fs.any_vis64 = any_vis64;
    fs.host_color = host_color;
    fs.cap_vis = 1 == nv && 0 != (views[0].capture_flags & SFTX_CAPTURE_VISIBILITY);
    fs.n_bands = n_bands;
    /* counters, chunk completions per tile and sample, band completions */
    fs.zero_words = COUNTER_WORDS + (size_t)nt * max_samples + n_bands;
    fs.layout_epoch = b->layout_epoch;
}

/* The per-batch device buffers of a prepared batch, each at least the
   bytes it needs: grown by allocation (the frame sets), or carved from the
   set's carve memory (the no-allocation set), or only summed (`measure`):
   returns the bytes the carve takes. */
static size_t acquire_buffers(FrameSet& fs, bool measure) {
    hipStream_t s = fs.stream;
    const u32 nt = fs.nt, nw = fs.nw;
    size_t off = 0;
    auto get = [&](DBuf& d, size_t bytes) {
        if (!fs.carve_base) {
            dgrow(d, bytes, false, s);
            return;
        }
        if (!measure) {
            d.p = fs.carve_base + off;
            d.cap = bytes;
        }
        off = (off + std::max<size_t>(bytes, 4) + 255) & ~(size_t)255;
    };
    get(fs.ss, std::max<size_t>((size_t)fs.ss_total * 2304, 4));
    get(fs.ss_vis, std::max<size_t>(fs.ss_total, 4));
    get(fs.clip_count, std::max<size_t>((size_t)nw * 4, 4));
    get(fs.clip_base, std::max<size_t>((size_t)nw * 4, 4));
    get(fs.clip_mask, std::max<size_t>((size_t)nw * 16, 16));
    get(fs.batch_pl, std::max<size_t>((size_t)nw * 16 * 8, 8));
    get(fs.bin_m, std::max<size_t>(fs.bin_m_words * 4, 4));
    get(fs.tile_off, (size_t)nt * 4);
    get(fs.tile_count, (size_t)nt * 4);
    get(fs.scan_sums, ((size_t)std::max(nw, nt) / 4096 + 2) * 4);
    get(fs.zero_d, fs.zero_words * 4);
    get(fs.color, std::max<size_t>(fs.color_total * 4, 4));
    get(fs.vid, std::max<size_t>(fs.vid_total * 4, 4));
    get(fs.tile_z, (size_t)nt * fs.max_samples * 8);
    get(fs.tile_lights, (size_t)nt * (kMaxTileLights + 1) * 4);
    get(fs.tile_nch, (size_t)nt * 4);
    get(fs.merge, (size_t)nt * fs.max_samples * 4096 * 8);
    if (fs.cap_vis) get(fs.cap_depth, fs.vid_total * 4);
    if (fs.any_vis64) get(fs.vis64, fs.vid_total * 8);
    get(fs.proxy_vis, std::max<size_t>((size_t)fs.n_proxies * 64, 64));
    get(fs.light_proxy_count, std::max<size_t>((size_t)fs.n_lights * 4, 4));
    get(fs.light_proxies, std::max<size_t>((size_t)fs.n_lights * kMaxLightProxies * sizeof(ProxyRec), sizeof(ProxyRec)));
    get(fs.tile_list, (size_t)fs.tile_list_cap * 8);
    get(fs.clip_rec, (size_t)fs.clip_cap * 164);
    get(fs.clip_key, (size_t)fs.clip_cap * 4);
    get(fs.clip_rect, (size_t)fs.clip_cap * 4);
    /* visibility work items: at most one per tile and sample plus one per
       raster_chunk payloads */
    const size_t max_items = ((size_t)nt + fs.tile_list_cap / fs.bg.raster_chunk + 1) * fs.max_samples;
    /* two lists: the items of views without MSAA, then those of MSAA views */
    fs.bg.items2_base = (u32)max_items;
    get(fs.raster_items, max_items * 2 * 8);
    get(fs.pack_d, fs.pack.used);
    return off;
}

/* Device side of a prepared batch: buffers, one H2D copy of the tables,
   the kernels, and asynchronous copies of the counters and host outputs.
   A set with carve memory runs its kernels eagerly and leaves its colour on
   the device; it returns false, having queued nothing, when its buffers do
   not fit the carve memory. */
static bool launch_batch(Backend* b, std::vector<ViewRecord>& views, FrameSet& fs, const FrameSet* after = nullptr,
                         bool allow_stream = true) {
    hipStream_t s = fs.stream;
    const u32 nv = (u32)views.size();
    if (fs.carve_base) {
        const size_t need = acquire_buffers(fs, true);
        if (need > fs.carve_bytes) {
            fprintf(stderr, "sofren2: the frame needs %zu bytes of intermediates; the no-allocation path has %zu\n", need,
                    fs.carve_bytes);
            return false;
        }
    }
    /* host outputs go through pinned staging: colour (streamed into it by
       the shading kernels band by band, or copied after the frame) and the
       64-bit visibility planes.  An overflow re-run does not stream. */
    const size_t color_bytes = fs.host_color ? (size_t)fs.color_total * 4 : 0;
    const size_t stage = color_bytes + (size_t)fs.vis_stage_total * 8;
    const bool stream = allow_stream && !fs.carve_base && color_bytes && fs.n_bands > 0 && b->stream_download;
    if (stage && !fs.carve_base && fs.out_h_cap < stage) {
        if (fs.out_h) HIP_CHECK(hipHostFree(fs.out_h));
        HIP_CHECK(hipHostMalloc((void**)&fs.out_h, stage, hipHostMallocCoherent));
        fs.out_h_cap = stage;
    }
    if (stream) {
        if (fs.band_flag_cap < fs.n_bands) {
            if (fs.band_flag_h) HIP_CHECK(hipHostFree(fs.band_flag_h));
            fs.band_flag_cap = std::max(fs.n_bands, 256u);
            HIP_CHECK(hipHostMalloc((void**)&fs.band_flag_h, (size_t)fs.band_flag_cap * 4, hipHostMallocCoherent));
        }
        /* the previous batch of this set is complete: nothing writes them now */
        memset(fs.band_flag_h, 0, (size_t)fs.n_bands * 4);
    }
    fs.streamed = stream;
    for (u32 vi = 0; vi < nv; vi += 1) fs.vgpu[vi].band_base = stream ? fs.plan[vi].band_base : kInvalid;
    HIP_CHECK(hipEventRecord(fs.ev[0], s));
    acquire_buffers(fs, false);
    fs.bg.clip_cap = fs.clip_cap;
    fs.bg.tile_list_cap = fs.tile_list_cap;
    for (u32 vi = 0; vi < nv; vi += 1) {
        fs.vgpu[vi].fc.clip_cap = fs.clip_cap;
        fs.vgpu[vi].fc.tile_list_cap = fs.tile_list_cap;
    }
    Packer& pk = fs.pack;
    memcpy(pk.host + fs.o_batch, &fs.bg, sizeof(fs.bg));
    memcpy(pk.host + fs.o_views, fs.vgpu.data(), fs.vgpu.size() * sizeof(ViewGpu));
    HIP_CHECK(hipMemcpyAsync(fs.pack_d.p, pk.host, pk.used, hipMemcpyHostToDevice, s));
    u8* pd = (u8*)fs.pack_d.p;
    if (fs.t_slots) {
        const size_t r0 = fs.t_base;
        HIP_CHECK(hipMemcpyAsync((u8*)b->pos.p + r0 * 768, pd + fs.o_tpos, (size_t)fs.t_slots * 768, hipMemcpyDeviceToDevice, s));
        HIP_CHECK(hipMemcpyAsync((u8*)b->attr.p + r0 * 1792, pd + fs.o_tattr, (size_t)fs.t_slots * 1792, hipMemcpyDeviceToDevice, s));
        HIP_CHECK(hipMemcpyAsync((u8*)b->inds.p + r0 * 512, pd + fs.o_tinds, (size_t)fs.t_slots * 512, hipMemcpyDeviceToDevice, s));
        HIP_CHECK(hipMemcpyAsync((u8*)b->cull.p + r0 * 32, pd + fs.o_tcull, (size_t)fs.t_slots * 32, hipMemcpyDeviceToDevice, s));
        HIP_CHECK(hipMemcpyAsync((u8*)b->slot_mesh.p + r0 * 4, pd + fs.o_tsm, (size_t)fs.t_slots * 4, hipMemcpyDeviceToDevice, s));
        if (fs.t_skins) {
            HIP_CHECK(hipMemcpyAsync((u8*)b->skin.p + (size_t)fs.t_skin_base * 768, pd + fs.o_tskin, (size_t)fs.t_skins * 768,
                                     hipMemcpyDeviceToDevice, s));
        }
    }
    for (u32 vi = 0; vi < nv; vi += 1) {
        const ViewRecord& v = views[vi];
        const bool occlusion = v.shadows_enabled && !v.point_lights.empty() && !v.proxies.empty() &&
                               kInvalid == fs.vgpu[vi].lp_cache;
        if (occlusion && 0 == fs.vgpu[vi].fc.tlas_nodes) {
            /* no instance to occlude: every ray reaches its light */
            HIP_CHECK(hipMemsetAsync((u8*)fs.proxy_vis.p + (size_t)fs.plan[vi].proxy_base * 64, 3, v.proxies.size() * 64, s));
        }
        /* a view without clearing starts from the caller's frame */
        if (!v.should_clear && v.color_buffer && v.width > 0 && v.height > 0) {
            HIP_CHECK(hipMemcpy2DAsync((u32*)fs.color.p + fs.plan[vi].color_base, (size_t)v.width * 4, v.color_buffer,
                                       (size_t)v.color_pitch * 4, (size_t)v.width * 4, v.height, hipMemcpyHostToDevice, s));
        }
    }
    u32* zd = (u32*)fs.zero_d.p;

    FramePtrs p;
    p.pos = (const f32*)b->pos.p;
    p.attr = (const f32*)b->attr.p;
    p.inds = (const u8*)b->inds.p;
    p.cull = (const f32*)b->cull.p;
    p.slot_mesh = (const u32*)b->slot_mesh.p;
    p.skin = (const u8*)b->skin.p;
    p.texels = (const u32*)b->texels.p;
    p.tex_meta = (const TexMeta*)b->tex_meta_d.p;
    p.est_rcp = (const f32*)b->est_rcp.p;
    p.est_rsq = (const f32*)b->est_rsq.p;
    p.batch = (const BatchGpu*)(pd + fs.o_batch);
    p.views = (const ViewGpu*)(pd + fs.o_views);
    p.tile_view = (const u32*)(pd + fs.o_tv);
    p.shade_tiles = (const u32*)(pd + fs.o_st);
    p.insts = (const InstGpu*)(pd + fs.o_insts);
    p.work = (const WorkItem*)(pd + fs.o_work);
    p.bones = (const f32*)(pd + fs.o_bones);
    p.point_lights = (const PointLightGpu*)(pd + fs.o_pl);
    p.dir_lights = (const DirLightGpu*)(pd + fs.o_dl);
    p.mat_tex = (const u32*)(pd + fs.o_mt);
    p.mat_color = (const u32*)(pd + fs.o_mc);
    p.mat_shin = (const f32*)(pd + fs.o_ms);
    p.tmat_tex = (const u32*)(pd + fs.o_tt);
    p.tmat_color = (const u32*)(pd + fs.o_tc);
    p.tmat_shin = (const f32*)(pd + fs.o_ts);
    p.mesh_recs = (const MeshRec*)(pd + fs.o_mr);
    p.ss = (f32*)fs.ss.p;
    p.ss_vis = (u8*)fs.ss_vis.p;
    p.clip_count = (u32*)fs.clip_count.p;
    p.clip_base = (u32*)fs.clip_base.p;
    p.clip_mask = (u64*)fs.clip_mask.p;
    p.batch_pl = (uint2*)fs.batch_pl.p;
    p.segs = (const BinSeg*)(pd + fs.o_seg);
    p.bin_m = (u32*)fs.bin_m.p;
    p.tile_count = (u32*)fs.tile_count.p;
    p.tile_off = (u32*)fs.tile_off.p;
    p.tile_list = (u64*)fs.tile_list.p;
    p.clip_rec = (f32*)fs.clip_rec.p;
    p.clip_key = (u32*)fs.clip_key.p;
    p.clip_rect = (u32*)fs.clip_rect.p;
    p.color = (u32*)fs.color.p;
    p.vid = (u32*)fs.vid.p;
    p.tile_lights = (u32*)fs.tile_lights.p;
    p.tile_z = (f32*)fs.tile_z.p;
    p.cap_depth = fs.cap_vis ? (f32*)fs.cap_depth.p : nullptr;
    p.vis64 = fs.any_vis64 ? (u64*)fs.vis64.p : nullptr;
    p.counters = zd;
    p.scan_sums = (u32*)fs.scan_sums.p;
    p.raster_items = (uint2*)fs.raster_items.p;
    p.tile_nch = (u32*)fs.tile_nch.p;
    p.tile_done = zd + COUNTER_WORDS;
    p.merge = (u64*)fs.merge.p;
    p.blas = (const sft_bvh_node*)b->blas.p;
    p.tlas = (const sft_bvh_node*)(pd + fs.o_tlas);
    p.proxies = (const ProxyGpu*)(pd + fs.o_px);
    p.proxy_vis = (u8*)fs.proxy_vis.p;
    p.light_proxy_count = (u32*)fs.light_proxy_count.p;
    p.light_proxies = (ProxyRec*)fs.light_proxies.p;
    p.cached_lp = (const ProxyRec*)(pd + fs.o_clp);
    p.cached_lpc = (const u32*)(pd + fs.o_clpc);
    p.color_host = stream ? (u32*)fs.out_h : nullptr;
    p.band_count = zd + COUNTER_WORDS + (size_t)fs.nt * fs.max_samples;
    p.band_flag = stream ? fs.band_flag_h : nullptr;

    /* kernels after the previous ordered batch's (ev[4]: its last pass) */
    if (after && after->in_flight) HIP_CHECK(hipStreamWaitEvent(s, after->ev[4], 0));
    if (fs.carve_base) {
        /* eager launches: instantiating a graph may allocate */
        launch_frame(p, fs.bg, fs.zero_words, s, fs.ev, &fs.side);
        HIP_CHECK(hipGetLastError());
        fs.side_used = fs.bg.occ_total > 0u;
    }
    else {
        run_frame(b, fs, p);
    }
    HIP_CHECK(hipMemcpyAsync(fs.counters_h, zd, CNT_COUNT * 4, hipMemcpyDeviceToHost, s));
    if (fs.carve_base) {
        /* the colour goes to the callers' buffers once the batch is known
           not to have overflowed (render_noalloc) */
        HIP_CHECK(hipEventRecord(fs.ev[5], s));
        fs.in_flight = true;
        return true;
    }

    /* host outputs: the colour planes (unless streamed) and 64-bit visibility */
    if (stage) {
        if (color_bytes && !stream) HIP_CHECK(hipMemcpyAsync(fs.out_h, fs.color.p, color_bytes, hipMemcpyDeviceToHost, s));
        /* 64-bit visibility of the views that ask for it, packed */
        for (u32 vi = 0; vi < nv; vi += 1) {
            if (!views[vi].vis64) continue;
            const FrameSet::ViewPlan& vp = fs.plan[vi];
            HIP_CHECK(hipMemcpyAsync(fs.out_h + color_bytes + vp.vis_stage * 8, (const u64*)fs.vis64.p + vp.vid_base,
                                     (size_t)vp.grid * 8, hipMemcpyDeviceToHost, s));
        }
    }
    HIP_CHECK(hipEventRecord(fs.ev[5], s));
    fs.in_flight = true;
    return true;
}

/* the copies of a launched batch's host outputs out of its pinned staging:
   colour rows into the callers' buffers (unless `color` is false),
   64-bit visibility planes */
struct Span { void* dst; const void* src; size_t n; };
static std::vector<Span> host_output_spans(const std::vector<ViewRecord>& views, const FrameSet& fs, bool color = true) {
    std::vector<Span> spans;
    const size_t color_bytes = fs.host_color ? (size_t)fs.color_total * 4 : 0;
    for (u32 vi = 0; vi < (u32)views.size(); vi += 1) {
        const ViewRecord& v = views[vi];
        if (color && color_bytes && v.color_buffer) {
            const u32* src = (const u32*)fs.out_h + fs.plan[vi].color_base;
            for (i32 y = 0; y < v.height; y += 1) {
                spans.push_back({v.color_buffer + (size_t)y * v.color_pitch, src + (size_t)y * v.width, (size_t)v.width * 4});
            }
        }
        if (v.vis64) spans.push_back({v.vis64, fs.out_h + color_bytes + fs.plan[vi].vis_stage * 8, fs.plan[vi].grid * 8});
    }
    return spans;
}

/* up to 16 threads, 1 MB or more each */
static void copy_spans(const std::vector<Span>& spans) {
    parallel_runs(spans.size(), [&](size_t i) { return spans[i].n; },
                  [&](size_t i) { memcpy(spans[i].dst, spans[i].src, spans[i].n); }, 16, (size_t)1 << 20);
}

/* A streamed batch's colour rows by band (a view's tile row, whose flag the
   shading kernels raise once its pixels are in the pinned staging). */
struct StreamCopy {
    std::vector<Span> rows;
    struct Band { u32 flag; size_t first, end; };   /* rows [first, end) */
    std::vector<Band> bands;
    const u32* flags = nullptr;
    hipEvent_t done = nullptr;      /* the batch's end */
};

static StreamCopy stream_plan(const std::vector<ViewRecord>& views, const FrameSet& fs) {
    StreamCopy sc;
    sc.flags = fs.band_flag_h;
    sc.done = fs.ev[5];
    for (u32 vi = 0; vi < (u32)views.size(); vi += 1) {
        const ViewRecord& v = views[vi];
        const u32 base = fs.plan[vi].band_base;
        if (kInvalid == base || !v.color_buffer) continue;
        const u32* src = (const u32*)fs.out_h + fs.plan[vi].color_base;
        for (i32 ty = 0; ty < v.tile_rows; ty += 1) {
            const size_t first = sc.rows.size();
            for (i32 y = ty * 64; y < std::min(v.height, ty * 64 + 64); y += 1) {
                sc.rows.push_back({v.color_buffer + (size_t)y * v.color_pitch, src + (size_t)y * v.width, (size_t)v.width * 4});
            }
            sc.bands.push_back({base + (u32)ty, first, sc.rows.size()});
        }
    }
    return sc;
}

/* Copies a streamed batch's colour out of the pinned staging band by band,
   each band once the shading kernels have raised its flag, while the later
   bands still shade; `threads` threads take the bands in turn.  Returns
   false, leaving further bands uncopied, when a band reports overflowed
   intermediates (the batch runs again) or the batch ends with a band not
   flagged (the caller then downloads the frame after the batch). */
static bool stream_copy(const StreamCopy& sc, size_t threads) {
    std::atomic<bool> fail{false};
    auto run = [&](size_t t) {
        for (size_t i = t; i < sc.bands.size(); i += threads) {
            const StreamCopy::Band& bd = sc.bands[i];
            u32 f = 0;
            for (u32 spin = 1;; spin += 1) {
                f = __atomic_load_n(&sc.flags[bd.flag], __ATOMIC_ACQUIRE);
                if (0u != f || fail.load(std::memory_order_relaxed)) break;
                if (0u == (spin & 4095u) && hipSuccess == hipEventQuery(sc.done)) {
                    f = __atomic_load_n(&sc.flags[bd.flag], __ATOMIC_ACQUIRE);
                    break;
                }
            }
            if (1u != f) {
                fail.store(true, std::memory_order_relaxed);
                return;
            }
            for (size_t k = bd.first; k < bd.end; k += 1) memcpy(sc.rows[k].dst, sc.rows[k].src, sc.rows[k].n);
        }
    };
    threads = std::max<size_t>(1u, std::min(threads, sc.bands.size()));
    WorkerPool& pool = WorkerPool::get();
    if (threads > 1 && threads <= pool.size() + 1 && pool.try_acquire()) {
        pool.run(threads, std::function<void(size_t)>(run));
        pool.release();
    }
    else {
        std::vector<std::thread> helpers;
        for (size_t t = 1; t < threads; t += 1) helpers.emplace_back(run, t);
        run(0);
        for (std::thread& h : helpers) h.join();
    }
    return !fail.load();
}

/* Moves a prepared batch's transient rows to where its frame set's
   transient region is now.  Another set's batch can grow the pools between
   this batch's prepare and its re-run after an overflow (new resident models
   raise res_cap, larger dmesh batches raise t_cap), which moves the region.
   The batch's staged tables already hold its transient rows and its work
   items and instances refer to them by slot: those slot numbers are shifted
   (the dmesh models they came from may be gone by now). */
static void relocate_transient(Backend* b, FrameSet& fs) {
    const u32 nb = b->res_cap + fs.region * b->t_cap, nsb = b->res_skin_cap + fs.region * b->t_skin_cap;
    if (fs.t_slots > b->t_cap || fs.t_skins > b->t_skin_cap) {
        fprintf(stderr, "fatal sofren2 error: transient rows exceed their region\n");
        exit(1);
    }
    if (nb != fs.t_base || nsb != fs.t_skin_base) {
        auto moved = [](u32 v, u32 from, u32 n, u32 to) { return (v >= from && v < from + n) ? v - from + to : v; };
        WorkItem* w = (WorkItem*)(fs.pack.host + fs.o_work);
        for (u32 i = 0; i < fs.nw; i += 1) {
            w[i].slot = moved(w[i].slot, fs.t_base, fs.t_slots, nb);
            if (kInvalid != w[i].skin) w[i].skin = moved(w[i].skin, fs.t_skin_base, fs.t_skins, nsb);
        }
        InstGpu* in = (InstGpu*)(fs.pack.host + fs.o_insts);
        for (u32 i = 0; i < fs.n_insts; i += 1) in[i].model_base = moved(in[i].model_base, fs.t_base, fs.t_slots, nb);
        fs.t_base = nb;
        fs.t_skin_base = nsb;
    }
    fs.layout_epoch = b->layout_epoch;
}

static void write_stats(Context& c, const std::vector<ViewRecord>& views, const FrameSet& fs);
static void capture_outputs(Backend* b, const std::vector<ViewRecord>& views, const FrameSet& fs);
static void fill_shadow_caches(Backend* b, const std::vector<ViewRecord>& views, const FrameSet& fs);

/* Waits for a launched batch; re-runs it with larger intermediates if any
   overflowed; writes the host outputs (unless its copier already did), the
   statistics and the captures. */
static void complete_batch(Backend* b, Context& c, std::vector<ViewRecord>& views, FrameSet& fs) {
    hipStream_t s = fs.stream;
    if (fs.copier.joinable()) fs.copier.join();
    bool copied = fs.host_copied;
    fs.host_copied = false;
    if (fs.streamed) {
        if (copied) b->streamed_batches += 1;
        else b->stream_fallbacks += 1;
    }
    for (int attempt = 0;; attempt += 1) {
        HIP_CHECK(hipStreamSynchronize(s));
        memcpy(fs.counters, fs.counters_h, sizeof(fs.counters));
        if (0u == fs.counters[CNT_OVERFLOW]) break;
        if (attempt > 8) {
            fprintf(stderr, "fatal sofren2 error: frame intermediates keep overflowing\n");
            exit(1);
        }
        if (fs.counters[CNT_OVERFLOW] & 1u) fs.clip_cap = std::max(fs.clip_cap * 2u, fs.counters[CNT_CLIP_TRIS] + 1024u);
        if (fs.counters[CNT_OVERFLOW] & 2u) fs.tile_list_cap = std::max(fs.tile_list_cap * 2u, fs.counters[CNT_TILE_PAIRS] + 4096u);
        /* a smaller occlusion stack than the full one filled up: trace again with the full one */
        if (fs.counters[CNT_OVERFLOW] & 4u) fs.bg.occ_stack = kOccStackEntries;
        /* another set's batch may have grown the pools since this one was
           prepared, moving this set's transient region */
        if (fs.layout_epoch != b->layout_epoch) relocate_transient(b, fs);
        b->reruns += 1;
        launch_batch(b, views, fs, nullptr, false);
        copied = false;
    }
    fs.in_flight = false;
    if (!copied) {
        /* a streamed launch whose bands were not all delivered: download
           the frame now */
        if (fs.streamed && fs.host_color) {
            HIP_CHECK(hipMemcpy(fs.out_h, fs.color.p, (size_t)fs.color_total * 4, hipMemcpyDeviceToHost));
        }
        copy_spans(host_output_spans(views, fs));
    }
    fill_shadow_caches(b, views, fs);
    write_stats(c, views, fs);
    capture_outputs(b, views, fs);
}

/* served templates' shadow caches: the capsule records and counts of each
   view that traced its template's proxies, read back once the batch is done */
static void fill_shadow_caches(Backend* b, const std::vector<ViewRecord>& views, const FrameSet& fs) {
    (void)b;
    for (u32 vi = 0; vi < (u32)views.size(); vi += 1) {
        const ViewRecord& v = views[vi];
        ShadowCache* sc = v.shadow_cache.get();
        if (!sc || sc->filled || kInvalid != fs.vgpu[vi].lp_cache) continue;
        if (!(v.shadows_enabled && !v.point_lights.empty() && !v.proxies.empty())) continue;
        const size_t nl = v.point_lights.size(), l0 = fs.vgpu[vi].pl_base;
        sc->recs.resize(nl * kMaxLightProxies);
        sc->counts.resize(nl);
        HIP_CHECK(hipMemcpy(sc->recs.data(), (const ProxyRec*)fs.light_proxies.p + l0 * kMaxLightProxies,
                            nl * kMaxLightProxies * sizeof(ProxyRec), hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(sc->counts.data(), (const u32*)fs.light_proxy_count.p + l0, nl * 4, hipMemcpyDeviceToHost));
        sc->filled = true;
    }
}

/* the frame statistics of a completed batch (sftx_get_stats) */
static void write_stats(Context& c, const std::vector<ViewRecord>& views, const FrameSet& fs) {
    const u32* counters = fs.counters;
    u64 px = 0, samples = 0;
    for (const ViewRecord& v : views) {
        px += (u64)v.width * v.height;
        samples += (u64)v.width * v.height * (u64)v.samples;
    }
    sftx_frame_stats& st = c.stats;
    memset(&st, 0, sizeof(st));
    st.meshlets_tested = fs.nw;
    st.meshlets_visible = counters[CNT_VISIBLE];
    st.meshlets_skinned = counters[CNT_SKINNED];
    st.batch_payloads = counters[CNT_BATCH_PAYLOADS];
    st.clip_triangles = counters[CNT_CLIP_TRIS];
    st.tile_payloads = counters[CNT_TILE_PAIRS];
    st.tiles = fs.nt;
    st.samples = (u32)samples;
    st.pixels = (u32)px;
    st.shaded_pixels = counters[CNT_SHADED];
    st.shading_groups = counters[CNT_QGROUPS];
    st.tex_bilinear = counters[CNT_TEX_BILINEAR];
    st.tex_nearest = counters[CNT_TEX_NEAREST];
    st.lightmapped_pixels = counters[CNT_LIGHTMAPPED];
    st.occlusion_rays = counters[CNT_OCC_RAYS];
    st.occlusion_nodes = counters[CNT_OCC_NODES];
    st.occlusion_tri_tests = counters[CNT_OCC_TRIS];
    st.occlusion_invalid = counters[CNT_OCC_INVALID];
    st.occlusion_overflow = counters[CNT_OCC_OVERFLOW];
    st.shadow_proxies = fs.n_proxies;
    st.tex_npot = counters[CNT_TEX_NPOT];
    st.unlit_shadings = counters[CNT_UNLIT];
    st.dirlit_shadings = counters[CNT_DIR_LIT];
    st.light_cut_tiles = counters[CNT_LIGHT_CUT];
    st.bg_blend_pixels = counters[CNT_BG_BLEND];
    st.capsule_tests = counters[CNT_CAP_TESTS];
    for (const ViewRecord& v : views) st.instances_dropped += v.dropped;
    st.ms_upload = ev_ms(fs.ev[0], fs.ev[1]);
    st.ms_shadows = fs.side_used ? ev_ms(fs.side.start, fs.side.end) : ev_ms(fs.ev[1], fs.ev[7]);
    st.ms_geometry = ev_ms(fs.ev[7], fs.ev[2]);
    st.ms_binning = ev_ms(fs.ev[2], fs.ev[3]);
    st.ms_raster_shade = ev_ms(fs.ev[3], fs.ev[4]);
    st.ms_raster = ev_ms(fs.ev[3], fs.ev[6]);
    st.ms_shade = ev_ms(fs.ev[6], fs.ev[4]);
    st.ms_download = ev_ms(fs.ev[4], fs.ev[5]);
    st.ms_total = ev_ms(fs.ev[0], fs.ev[5]);
}

/* host copies of a completed single-view batch's intermediates, as its
   capture flags ask */
static void capture_outputs(Backend* b, const std::vector<ViewRecord>& views, const FrameSet& fs) {
    const u32* counters = fs.counters;
    const u32 nv = (u32)views.size();
    if (1 != nv) return;
    const ViewRecord& v0 = views[0];
    const u32 cap_flags = v0.capture_flags;
    b->cap_cols = v0.tile_cols;
    b->cap_rows = v0.tile_rows;
    b->cap_samples = v0.samples;
    const size_t grid = fs.plan[0].grid;
    if (fs.cap_vis) {
        b->h_vid.resize(grid);
        b->h_depth.resize(grid);
        HIP_CHECK(hipMemcpy(b->h_vid.data(), fs.vid.p, grid * 4, hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(b->h_depth.data(), fs.cap_depth.p, grid * 4, hipMemcpyDeviceToHost));
    }
    if (cap_flags & SFTX_CAPTURE_SS_CACHE) {
        b->h_ss_count = v0.total_ss;
        b->h_ss.resize((size_t)v0.total_ss * 2304);
        b->h_ss_vis.assign(v0.total_ss, 0);
        if (v0.total_ss) {
            HIP_CHECK(hipMemcpy(b->h_ss.data(), fs.ss.p, b->h_ss.size(), hipMemcpyDeviceToHost));
            HIP_CHECK(hipMemcpy(b->h_ss_vis.data(), fs.ss_vis.p, v0.total_ss, hipMemcpyDeviceToHost));
        }
    }
    if (cap_flags & SFTX_CAPTURE_SHADOWS) {
        const bool occlusion = v0.shadows_enabled && !v0.point_lights.empty() && !v0.proxies.empty();
        b->h_proxy_count = fs.n_proxies;
        b->h_proxy_vis.assign((size_t)fs.n_proxies * 64, 0);
        /* the view's TLAS nodes as the occlusion pass read them */
        b->h_tlas_count = occlusion ? fs.vgpu[0].fc.tlas_nodes : 0u;
        b->h_tlas.resize((size_t)b->h_tlas_count * sizeof(sft_bvh_node));
        if (b->h_tlas_count) {
            HIP_CHECK(hipMemcpy(b->h_tlas.data(), (const u8*)fs.pack_d.p + fs.o_tlas + (size_t)fs.plan[0].tlas_base * sizeof(sft_bvh_node),
                                b->h_tlas.size(), hipMemcpyDeviceToHost));
        }
        if (occlusion) {
            HIP_CHECK(hipMemcpy(b->h_proxy_vis.data(), fs.proxy_vis.p, b->h_proxy_vis.size(), hipMemcpyDeviceToHost));
            /* counts of lights beyond the traced ones are not kept */
            const size_t lc = std::min<size_t>(v0.point_lights.size(), 64);
            for (size_t i = 0; i < fs.n_proxies; i += 1) memset(&b->h_proxy_vis[i * 64 + lc], 0, 64 - lc);
        }
    }
    if (cap_flags & SFTX_CAPTURE_CLIPS) {
        b->h_clip_count = std::min(counters[CNT_CLIP_TRIS], fs.clip_cap);
        b->h_clips.resize((size_t)b->h_clip_count * 164);
        if (b->h_clip_count) HIP_CHECK(hipMemcpy(b->h_clips.data(), fs.clip_rec.p, b->h_clips.size(), hipMemcpyDeviceToHost));
    }
}

void backend_render(Backend* b, Context* cp, std::vector<ViewRecord>& views) {
    if (views.empty()) return;
    FrameSet& fs = b->sets[0];
    if (fs.in_flight) HIP_CHECK(hipStreamSynchronize(fs.stream));
    prepare_batch(b, *cp, views, fs);
    launch_batch(b, views, fs);
    if (fs.streamed) {
        /* the colour while the frame shades, then the rest after it */
        if (stream_copy(stream_plan(views, fs), 2)) {
            HIP_CHECK(hipEventSynchronize(fs.ev[5]));
            if (0u == fs.counters_h[CNT_OVERFLOW]) {
                copy_spans(host_output_spans(views, fs, false));
                fs.host_copied = true;
            }
        }
    }
    complete_batch(b, *cp, views, fs);
    b->last_set = 0;
}

void backend_submit(Backend* b, Context* cp, std::vector<ViewRecord>& views, int set, bool ordered) {
    FrameSet& fs = b->sets[set];
    prepare_batch(b, *cp, views, fs);
    const FrameSet* after = (ordered && b->ordered_set >= 0 && b->ordered_set != set) ? &b->sets[b->ordered_set] : nullptr;
    launch_batch(b, views, fs, after);
    if (ordered) b->ordered_set = set;
    /* host outputs are copied as the batch delivers them (streamed colour
       band by band, the rest once it is done), while the caller records and
       submits the next one; completing the batch joins the copy */
    std::vector<Span> spans = host_output_spans(views, fs, !fs.streamed);
    if (!spans.empty() || fs.streamed) {
        /* an overflowed batch is re-run by backend_complete, which may read
           the caller's buffers again (views without clearing start from
           them): its outputs are not copied */
        StreamCopy sc;
        if (fs.streamed) sc = stream_plan(views, fs);
        fs.host_copied = false;
        fs.copier = std::thread([streamed = fs.streamed, ev = fs.ev[5], counters = fs.counters_h, copied = &fs.host_copied,
                                 sc = std::move(sc), sp = std::move(spans)]() {
            /* two threads per frame-sized band set, up to eight for batches
               of several megabytes (first-touch page faults in the callers'
               buffers slow a single thread well below the PCIe rate) */
            const size_t threads = std::min<size_t>(8u, std::max<size_t>(2u, sc.rows.size() / 2048u));
            if (streamed && !stream_copy(sc, threads)) return;
            HIP_CHECK(hipEventSynchronize(ev));
            if (0u != counters[CNT_OVERFLOW]) return;
            copy_spans(sp);
            *copied = true;
        });
    }
}

void backend_complete(Backend* b, Context* cp, std::vector<ViewRecord>& views, int set) {
    complete_batch(b, *cp, views, b->sets[set]);
    b->last_set = set;
}

bool backend_async_present(const Backend* b) { return SFTX_OUTPUT_DEVICE == b->output_mode; }

bool backend_batch_done(Backend* b, int set) {
    return hipSuccess == hipEventQuery(b->sets[set].ev[5]);
}

void backend_forget_model(Backend* b, const sft_model* m) { b->models.erase(m); }

void backend_get_tuning(const Backend* b, sftx_tuning* t) {
    for (int k = 0; k < SFTX_RES_CLASSES; k += 1) t->raster_chunk[k] = b->raster_chunk[k];
    t->raster_groups = b->raster_groups;
    t->raster_groups_resident = raster_groups_per_device();
    t->tile_px = SFT_TILE_SIZE;
    for (int k = 0; k < SFTX_RES_CLASSES; k += 1) t->raster_split[k] = b->raster_split[k];
    t->occlusion_stack = b->occ_stack;
}

void backend_set_tuning(Backend* b, const sftx_tuning* t) {
    for (int k = 0; k < SFTX_RES_CLASSES; k += 1) {
        b->raster_chunk[k] = std::max(1u, std::min(t->raster_chunk[k], kRasterMaxItem));
    }
    b->raster_groups = std::max(1u, std::min(t->raster_groups, 4u * raster_groups_per_device()));
    for (int k = 0; k < SFTX_RES_CLASSES; k += 1) b->raster_split[k] = std::max(1u, t->raster_split[k]);
    b->occ_stack = (t->occlusion_stack <= 256u) ? 256u : (t->occlusion_stack <= 512u) ? 512u : kOccStackEntries;
}

/* Autotuner (blocking, on frame set 0).  The knobs leave every output bit
   unchanged and change only how passes are cut into work: for each
   resolution class the payloads per visibility work item (the item's
   triangles are staged in LDS, up to 8 per payload; fewer payloads per item
   spread a dense tile over more workgroups and merge more planes in HBM) and
   the row-band split threshold (a payload whose estimated sweep exceeds it
   is swept as 2, 4 or 8 jobs over interleaved rows, each reaching its rows
   by add-only steps from the tile's origin); for the batch the size of the
   persistent visibility grid and the occlusion traversal stack per wave
   (LDS per workgroup; a stack that fills up re-runs the batch with the full
   one).  Each candidate renders the views being tuned `repeats` times with
   eager launches; the objective is the smallest pass time (HIP events:
   k_raster_plan + k_raster for the visibility knobs, the occlusion pass for
   the stack).  Classes are tuned one after the other, then the batch knobs
   on the whole batch, which is rendered once more with the chosen settings
   so its outputs are written as by a blocking submit.  A candidate replaces
   the current choice only when at least 3% faster. */
void backend_tune(Backend* b, Context* cp, std::vector<ViewRecord>& views, int repeats, sftx_tuning* report) {
    if (views.empty()) return;
    repeats = std::max(1, std::min(repeats, 64));
    const bool graphs = b->use_graphs;
    b->use_graphs = false;
    auto measure_pass = [&](std::vector<ViewRecord>& vs, bool occlusion) {
        f32 best = 3.4e38f;
        for (int r = 0; r < repeats; r += 1) {
            backend_render(b, cp, vs);
            best = std::min(best, occlusion ? cp->stats.ms_shadows : cp->stats.ms_raster);
        }
        return best;
    };
    auto measure = [&](std::vector<ViewRecord>& vs) { return measure_pass(vs, false); };
    /* one knob over candidate values: keeps the first value 3% faster than the best so far */
    auto tune_knob = [&](u32& knob, std::initializer_list<u32> values, std::vector<ViewRecord>& vs, bool occlusion) {
        u32 best_v = knob;
        f32 best_ms = measure_pass(vs, occlusion);
        for (u32 val : values) {
            if (val == best_v) continue;
            const u32 keep = knob;
            knob = val;
            const f32 ms = measure_pass(vs, occlusion);
            knob = keep;
            if (ms < best_ms * 0.97f) {
                best_ms = ms;
                best_v = val;
            }
        }
        knob = best_v;
        return best_ms;
    };
    static const u32 chunks[] = { 64, 96, 128, 256, 384, 512 };
    for (int k = 0; k < SFTX_RES_CLASSES; k += 1) {
        report->ms_class[k] = 0.f;
        std::vector<ViewRecord> sub;
        for (const ViewRecord& v : views) {
            if (res_class(v.width, v.height) == k) sub.push_back(v);
        }
        if (sub.empty()) continue;
        u32 best_chunk = b->raster_chunk[k];
        f32 best_ms = measure(sub);
        for (u32 ch : chunks) {
            if (ch == best_chunk) continue;
            const u32 keep = b->raster_chunk[k];
            b->raster_chunk[k] = ch;
            const f32 ms = measure(sub);
            b->raster_chunk[k] = keep;
            /* a candidate replaces the current choice only when clearly faster */
            if (ms < best_ms * 0.97f) {
                best_ms = ms;
                best_chunk = ch;
            }
        }
        b->raster_chunk[k] = best_chunk;
        best_ms = std::min(best_ms, tune_knob(b->raster_split[k], { 1024u, 2048u, 4096u, 8192u, 0xFFFFFFFFu }, sub, false));
        report->ms_class[k] = best_ms;
    }
    const u32 resident = raster_groups_per_device();
    u32 best_groups = b->raster_groups;
    f32 best_ms = measure(views);
    for (u32 g : { resident / 2u, resident, 2u * resident }) {
        if (g == best_groups || g == 0u) continue;
        const u32 keep = b->raster_groups;
        b->raster_groups = g;
        const f32 ms = measure(views);
        b->raster_groups = keep;
        if (ms < best_ms * 0.97f) {
            best_ms = ms;
            best_groups = g;
        }
    }
    b->raster_groups = best_groups;
    /* the occlusion stack, when the batch traces shadow packets */
    bool occlusion = false;
    for (const ViewRecord& v : views) occlusion = occlusion || (v.shadows_enabled && !v.point_lights.empty() && !v.proxies.empty());
    const f32 ms_occ = occlusion ? tune_knob(b->occ_stack, { 256u, 512u, kOccStackEntries }, views, true) : 0.f;
    b->use_graphs = graphs;
    backend_render(b, cp, views);
    const f32 mc[SFTX_RES_CLASSES] = { report->ms_class[0], report->ms_class[1], report->ms_class[2] };
    backend_get_tuning(b, report);
    for (int k = 0; k < SFTX_RES_CLASSES; k += 1) report->ms_class[k] = mc[k];
    report->ms_batch = best_ms;
    report->ms_occlusion = ms_occ;
}

/* sftx_present_noalloc: the recorded frame rendered on the no-allocation
   set, every intermediate carved from the scratch array, tables staged in
   ordinary host memory, kernels launched eagerly, the colour copied into
   the caller's buffer.  Nothing is allocated on the device or pinned on the
   host: a frame that would need it is refused with an error code. */
static i32 render_noalloc(Backend* b, Context& c) {
    if (!c.async_batches.empty()) {
        fprintf(stderr, "sofren2: sftx_present_noalloc with frames or batches in flight; call sftx_synchronize first\n");
        return SFTX_NOALLOC_IN_FLIGHT;
    }
    std::vector<ViewRecord> views(1);
    snapshot_view(c, views[0]);
    ViewRecord& v = views[0];
    v.capture_flags = 0;
    v.vis64 = nullptr;
    v.host_output = true;
    u32 need_t = 0, need_tskin = 0;
    if (shared_state_changes(b, c, views, need_t, need_tskin)) {
        fprintf(stderr, "sofren2: sftx_present_noalloc needs models, textures or dmesh room made resident first "
                        "(present the scene once with sft_present)\n");
        return SFTX_NOALLOC_NOT_RESIDENT;
    }
    /* the set's transient rows use region 0: nothing else may run */
    quiesce(b);
    FrameSet& fs = b->noalloc;
    prepare_batch(b, c, views, fs);
    for (int attempt = 0;; attempt += 1) {
        if (!launch_batch(b, views, fs)) return SFTX_NOALLOC_TOO_LARGE;
        HIP_CHECK(hipStreamSynchronize(fs.stream));
        fs.in_flight = false;
        memcpy(fs.counters, fs.counters_h, sizeof(fs.counters));
        if (0u == fs.counters[CNT_OVERFLOW]) break;
        if (attempt > 8) {
            fprintf(stderr, "fatal sofren2 error: frame intermediates keep overflowing\n");
            exit(1);
        }
        if (fs.counters[CNT_OVERFLOW] & 1u) fs.clip_cap = std::max(fs.clip_cap * 2u, fs.counters[CNT_CLIP_TRIS] + 1024u);
        if (fs.counters[CNT_OVERFLOW] & 2u) fs.tile_list_cap = std::max(fs.tile_list_cap * 2u, fs.counters[CNT_TILE_PAIRS] + 4096u);
        /* a smaller occlusion stack than the full one filled up: trace again with the full one */
        if (fs.counters[CNT_OVERFLOW] & 4u) fs.bg.occ_stack = kOccStackEntries;
        b->reruns += 1;
    }
    if (v.color_buffer && v.width > 0 && v.height > 0) {
        HIP_CHECK(hipMemcpy2DAsync(v.color_buffer, (size_t)v.color_pitch * 4, (const u32*)fs.color.p + fs.plan[0].color_base,
                                   (size_t)v.width * 4, (size_t)v.width * 4, v.height, hipMemcpyDeviceToHost, fs.stream));
        HIP_CHECK(hipStreamSynchronize(fs.stream));
    }
    write_stats(c, views, fs);
    b->last_set = 0;
    return 0;
}

u64 backend_resident_bytes(const Backend* b) {
    u64 n = 0;
    for (const DBuf* d : { &b->pos, &b->attr, &b->inds, &b->cull, &b->slot_mesh, &b->skin, &b->texels, &b->tex_meta_d,
                           &b->blas }) {
        n += d->cap;
    }
    return n;
}

}  // namespace sft

using namespace sft;

static Backend* ensure_backend();

/* the texture's mip chain as built in HBM (texels of the whole chain, in the
   4x4-block layout); makes pending textures resident first */
extern "C" i64 sftx_cap_texture(sft_tex_handle tex, u32* out, u64 capacity) {
    if (!g_ctx || SFT_INVALID_HANDLE == tex) return -1;
    Context& c = *g_ctx;
    const u32 i = SFT_HANDLE_INDEX(tex);
    if (i >= c.tex.size() || !c.tex[i].live || c.tex[i].generation != (tex >> SFT_HANDLE_GEN_SHIFT)) return -1;
    Backend* b = ensure_backend();
    if (texture_changes(b, c)) {
        quiesce(b);
        sync_textures(b, c);
    }
    const HostTexture& t = c.tex[i];
    if (!out) return (i64)t.total;
    if (capacity < t.total) return -1;
    HIP_CHECK(hipMemcpy(out, (const u32*)b->texels.p + t.dev_offset, t.total * 4, hipMemcpyDeviceToHost));
    return (i64)t.total;
}

extern "C" void sftx_cap_grid(i32* cols, i32* rows, i32* samples) {
    Backend* b = g_ctx ? g_ctx->backend : nullptr;
    if (cols) *cols = b ? b->cap_cols : 0;
    if (rows) *rows = b ? b->cap_rows : 0;
    if (samples) *samples = b ? b->cap_samples : 1;
}

extern "C" void sftx_cap_visibility(u32* vid, f32* depth) {
    Backend* b = g_ctx ? g_ctx->backend : nullptr;
    if (!b) return;
    if (vid && !b->h_vid.empty()) memcpy(vid, b->h_vid.data(), b->h_vid.size() * 4);
    if (depth && !b->h_depth.empty()) memcpy(depth, b->h_depth.data(), b->h_depth.size() * 4);
}

extern "C" u32 sftx_cap_ss_count(void) { return (g_ctx && g_ctx->backend) ? g_ctx->backend->h_ss_count : 0u; }

extern "C" void sftx_cap_ss_cache(void* out) {
    Backend* b = g_ctx ? g_ctx->backend : nullptr;
    if (!b) return;
    if (!b->h_ss.empty()) memcpy(out, b->h_ss.data(), b->h_ss.size());
}

extern "C" void sftx_cap_ss_visible(u8* out) {
    Backend* b = g_ctx ? g_ctx->backend : nullptr;
    if (!b) return;
    if (!b->h_ss_vis.empty()) memcpy(out, b->h_ss_vis.data(), b->h_ss_vis.size());
}

extern "C" u32 sftx_cap_clip_count(void) { return (g_ctx && g_ctx->backend) ? g_ctx->backend->h_clip_count : 0u; }

extern "C" void sftx_cap_clip_records(void* out) {
    Backend* b = g_ctx ? g_ctx->backend : nullptr;
    if (!b) return;
    if (!b->h_clips.empty()) memcpy(out, b->h_clips.data(), b->h_clips.size());
}

extern "C" u32 sftx_cap_proxy_count(void) { return (g_ctx && g_ctx->backend) ? g_ctx->backend->h_proxy_count : 0u; }

extern "C" void sftx_cap_proxy_visibility(u8* out) {
    Backend* b = g_ctx ? g_ctx->backend : nullptr;
    if (!b || !out) return;
    if (!b->h_proxy_vis.empty()) memcpy(out, b->h_proxy_vis.data(), b->h_proxy_vis.size());
}

extern "C" u32 sftx_cap_tlas_count(void) { return (g_ctx && g_ctx->backend) ? g_ctx->backend->h_tlas_count : 0u; }

extern "C" void sftx_cap_tlas(void* out) {
    Backend* b = g_ctx ? g_ctx->backend : nullptr;
    if (!b || !out) return;
    if (!b->h_tlas.empty()) memcpy(out, b->h_tlas.data(), b->h_tlas.size());
}

static Backend* ensure_backend() {
    if (!g_ctx) {
        fprintf(stderr, "fatal sofren2 error: used before sft_init\n");
        exit(1);
    }
    if (!g_ctx->backend) g_ctx->backend = backend_create();
    return g_ctx->backend;
}

extern "C" void sftx_set_output(i32 mode) { ensure_backend()->output_mode = mode; }

extern "C" void sftx_set_graphs(i32 enabled) { ensure_backend()->use_graphs = 0 != enabled; }

extern "C" void sftx_stream_stats(u64* streamed, u64* fallbacks) {
    Backend* b = ensure_backend();
    if (streamed) *streamed = b->streamed_batches;
    if (fallbacks) *fallbacks = b->stream_fallbacks;
}

extern "C" void sftx_set_gpu_tlas(i32 enabled) {
    Backend* b = ensure_backend();
    drain_async_batches(*g_ctx);
    quiesce(b);
    b->gpu_tlas = 0 != enabled;
}

extern "C" void sftx_set_stream_download(i32 enabled) {
    Backend* b = ensure_backend();
    drain_async_batches(*g_ctx);
    quiesce(b);
    b->stream_download = 0 != enabled;
}

extern "C" u64 sftx_graph_captures(void) { return ensure_backend()->graph_captures; }

extern "C" void sftx_set_initial_caps(u32 clip_cap, u32 tile_list_cap) {
    Backend* b = ensure_backend();
    if (g_ctx->server_running) {
        fprintf(stderr, "sofren2: sftx_set_initial_caps while the request server runs\n");
        return;
    }
    drain_async_batches(*g_ctx);
    quiesce(b);
    for (FrameSet& fs : b->sets) {
        fs.clip_cap = std::max(clip_cap, 1u);
        fs.tile_list_cap = std::max(tile_list_cap, 1u);
    }
}

extern "C" u64 sftx_overflow_reruns(void) { return ensure_backend()->reruns; }

extern "C" i32 sftx_model_changed(const sft_model* model) {
    if (!g_ctx || !model) return SFTX_MODEL_UNKNOWN;
    Context& c = *g_ctx;
    if (c.server_running) {
        fprintf(stderr, "sofren2: sftx_model_changed while the request server runs\n");
        return SFTX_MODEL_BUSY;
    }
    if (!c.backend || !c.backend->models.count(model)) return SFTX_MODEL_UNKNOWN;
    drain_async_batches(c);
    quiesce(c.backend);
    backend_forget_model(c.backend, model);
    c.backend->model_edits += 1;
    return 0;
}

extern "C" u64 sftx_model_edits(void) { return ensure_backend()->model_edits; }

extern "C" i32 sftx_present_noalloc(void) {
    if (!g_ctx) {
        fprintf(stderr, "sofren2: sftx_present_noalloc before sft_init\n");
        return SFTX_NOALLOC_UNREADY;
    }
    Context& c = *g_ctx;
    if (!c.backend || c.server_running || c.batch_open) {
        fprintf(stderr, "sofren2: sftx_present_noalloc needs a renderer set up by sft_present, outside batches and "
                        "the request server\n");
        return SFTX_NOALLOC_UNREADY;
    }
    return render_noalloc(c.backend, c);
}

extern "C" u64 sftx_noalloc_scratch_bytes(void) { return kScratchBytes; }

extern "C" const u32* sftx_device_color(void) {
    Backend* b = ensure_backend();
    drain_async_batches(*g_ctx);
    return (const u32*)b->sets[b->last_set].color.p;
}

extern "C" void sftx_synchronize(void) {
    Backend* b = ensure_backend();
    drain_async_batches(*g_ctx);
    quiesce(b);
}

extern "C" f32 sftx_copy_calibration(u64 bytes) {
    Backend* b = ensure_backend();
    bytes = (bytes + 15) & ~15ull;
    void *src = nullptr, *dst = nullptr;
    HIP_CHECK(hipMalloc(&src, bytes));
    HIP_CHECK(hipMalloc(&dst, bytes));
    FrameSet& fs = b->sets[0];
    HIP_CHECK(hipMemsetAsync(src, 1, bytes, fs.stream));
    HIP_CHECK(hipEventRecord(fs.ev[8], fs.stream));
    launch_copy((const uint4*)src, (uint4*)dst, bytes / 16, fs.stream);
    HIP_CHECK(hipEventRecord(fs.ev[9], fs.stream));
    HIP_CHECK(hipStreamSynchronize(fs.stream));
    const f32 ms = ev_ms(fs.ev[8], fs.ev[9]);
    HIP_CHECK(hipFree(src));
    HIP_CHECK(hipFree(dst));
    return ms;
}
