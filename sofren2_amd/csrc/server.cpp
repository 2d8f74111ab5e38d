/*
 * server.cpp — the request server: resident scene templates, render requests
 * admitted under an HBM budget, and continuous batching.
 *
 * A template is a recorded frame (instances, poses, lights, materials, shadow
 * proxies) kept with everything it points at that would not outlive the
 * frame: transient (dmesh) models are deep-copied and become resident
 * models, posed bone matrices are copied.  A request names a template, a
 * camera, a resolution and an MSAA setting; it renders exactly as that frame
 * would through sft_present with that camera and display size.
 *
 * One server thread owns the backend while the server runs.  It keeps up to
 * three batches in flight (on the backend's frame sets, one stream each; each
 * batch's kernels start after the previous batch's): while batches execute,
 * requests that arrived are admitted into the next one as long as the HBM
 * budget allows (resident scene state plus the framebuffer state of the
 * batches in flight plus the candidate), launched on a free set's stream, and
 * the oldest batch is completed when every slot is busy.  Requests are taken
 * in arrival order.
 *
 * A template with capsule shadows keeps the per-light capsule records of its
 * proxies once a batch has traced them (they depend on the template only):
 * later requests upload them instead of tracing the proxy-occlusion rays and
 * building the frame TLAS again.
 */
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string.h>
#include <thread>
#include <unordered_map>
#include <vector>
#include "sft_internal.h"

namespace sft {

namespace {

using Clock = std::chrono::steady_clock;

/* a dmesh model copied out of the frame's scratch arena */
struct OwnedModel {
    sft_model model{};
    sft_mesh mesh{};
    std::vector<sft_meshlet> meshlets;
    std::vector<sft_meshlet_attribs> attribs;
    std::vector<sft_meshlet_inds> inds;
    std::vector<f32> arrays[14];
    std::vector<u16> mesh_inds;
};

struct Template {
    ViewRecord frame;
    std::vector<std::unique_ptr<OwnedModel>> owned;
};

struct Request {
    sftx_request r{};
    u64 id = 0;
    u64 footprint = 0;
    Clock::time_point t_submit, t_done;
    bool done = false;
};

struct InFlight {
    int set = 0;
    std::vector<Request*> reqs;
    std::vector<ViewRecord> views;
    u64 footprint = 0;
};

struct Server {
    std::thread thread;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    bool running = false;
    u64 budget = 0;
    i32 max_views = 64;
    std::vector<std::unique_ptr<Template>> templates;
    std::deque<Request*> pending;
    std::unordered_map<u64, std::unique_ptr<Request>> requests;
    std::deque<InFlight> inflight;
    u64 next_id = 1;
    u64 inflight_bytes = 0;
    sftx_server_stats stats{};
    Clock::time_point t_start;
};

Server* g_server = nullptr;

std::unique_ptr<OwnedModel> copy_dmesh_model(const sft_model* m) {
    auto o = std::make_unique<OwnedModel>();
    const sft_mesh& src = m->meshes[0];
    const u32 n = src.meshlet_count;
    o->meshlets.assign(src.meshlets, src.meshlets + n);
    o->attribs.assign(src.meshlet_attribs, src.meshlet_attribs + n);
    o->inds.assign(src.inds, src.inds + n);
    f32* const in[14] = { src.bounds_center_x, src.bounds_center_y, src.bounds_center_z, src.bounds_radius,
                          src.bounds_min_x, src.bounds_max_x, src.bounds_min_y, src.bounds_max_y,
                          src.bounds_min_z, src.bounds_max_z, src.cone_axis_x, src.cone_axis_y, src.cone_axis_z,
                          src.cone_cutoff };
    for (int k = 0; k < 14; k += 1) {
        if (in[k]) o->arrays[k].assign(in[k], in[k] + n);
        else o->arrays[k].assign(n, 0.f);
    }
    o->mesh_inds.assign(m->meshlet_mesh_inds, m->meshlet_mesh_inds + m->total_meshlets);
    o->mesh = src;
    o->mesh.meshlets = o->meshlets.data();
    o->mesh.meshlet_attribs = o->attribs.data();
    o->mesh.inds = o->inds.data();
    f32** const out[14] = { &o->mesh.bounds_center_x, &o->mesh.bounds_center_y, &o->mesh.bounds_center_z,
                            &o->mesh.bounds_radius, &o->mesh.bounds_min_x, &o->mesh.bounds_max_x,
                            &o->mesh.bounds_min_y, &o->mesh.bounds_max_y, &o->mesh.bounds_min_z,
                            &o->mesh.bounds_max_z, &o->mesh.cone_axis_x, &o->mesh.cone_axis_y,
                            &o->mesh.cone_axis_z, &o->mesh.cone_cutoff };
    for (int k = 0; k < 14; k += 1) *out[k] = o->arrays[k].data();
    o->mesh.meshlet_skins = nullptr;
    o->mesh.culling_bone_inds = nullptr;
    o->model = *m;
    o->model.meshes = &o->mesh;
    o->model.mesh_count = 1;
    o->model.meshlet_mesh_inds = o->mesh_inds.data();
    return o;
}

/* HBM a request's framebuffer state takes while its batch is in flight:
   colour, per-sample IDs (and 64-bit visibility), 1/w range and merge planes
   of its tiles, its screen-space meshlet cache and per-meshlet binning rows */
u64 request_footprint(const sftx_request& r, const Template& t) {
    const u64 cols = ((u64)r.width + 63) / 64, rows = ((u64)r.height + 63) / 64;
    const u64 grid = cols * 64 * rows * 64 * (u64)r.samples;
    u64 meshlets = 0;
    for (const sft_instance& in : t.frame.instances) meshlets += in.model->total_meshlets;
    return (u64)r.width * r.height * 4 + grid * (4 + (r.vis64 ? 8 : 0)) + cols * rows * (u64)r.samples * 4096 * 8 +
           meshlets * (2304 + 16 * 8 + 32) + (1u << 20);
}

ViewRecord make_view(const Template& t, const sftx_request& r) {
    ViewRecord v = t.frame;
    v.width = r.width;
    v.height = r.height;
    v.half_width = (f32)r.width * 0.5f;
    v.half_height = (f32)r.height * 0.5f;
    v.vp_x = 0.f;
    v.vp_y = 0.f;
    v.tile_cols = (r.width + SFT_TILE_SIZE - 1) / SFT_TILE_SIZE;
    v.tile_rows = (r.height + SFT_TILE_SIZE - 1) / SFT_TILE_SIZE;
    v.samples = r.samples;
    v.camera_pos = r.camera_pos;
    camera_setup(r.view, r.proj, &v.view_proj, v.planes);
    v.color_buffer = r.color;
    v.color_pitch = r.width;
    v.host_output = nullptr != r.color;
    v.vis64 = r.vis64;
    v.capture_flags = 0;
    return v;
}

void server_loop(Server* sv) {
    Context& c = *g_ctx;
    std::unique_lock<std::mutex> lk(sv->mu);
    int next_set = 0;
    for (;;) {
        if (!sv->running && sv->pending.empty() && sv->inflight.empty()) break;
        if (sv->inflight.size() < (size_t)kServerBatchesInFlight && !sv->pending.empty()) {
            /* admit: arrival order, while the budget and the batch size allow */
            const u64 resident = backend_resident_bytes(c.backend);
            InFlight f;
            u64 bytes = 0;
            while (!sv->pending.empty() && (i32)f.reqs.size() < sv->max_views) {
                Request* q = sv->pending.front();
                const bool fits = resident + sv->inflight_bytes + bytes + q->footprint <= sv->budget;
                /* a request larger than the whole budget still runs, alone */
                const bool alone = f.reqs.empty() && sv->inflight.empty();
                if (!fits && !alone) break;
                sv->pending.pop_front();
                f.reqs.push_back(q);
                bytes += q->footprint;
            }
            if (!f.reqs.empty()) {
                f.footprint = bytes;
                f.set = next_set;
                next_set = (next_set + 1) % kServerBatchesInFlight;
                sv->inflight_bytes += bytes;
                for (Request* q : f.reqs) {
                    f.views.push_back(make_view(*sv->templates[(size_t)q->r.scene], q->r));
                    if (f.views.back().shadow_cache && f.views.back().shadow_cache->filled) sv->stats.shadow_cache_views += 1;
                }
                sv->stats.batches += 1;
                sv->stats.views_in_batches += f.reqs.size();
                sv->stats.max_batch = std::max<u64>(sv->stats.max_batch, f.reqs.size());
                sv->inflight.push_back(std::move(f));
                InFlight& g = sv->inflight.back();
                lk.unlock();
                const Clock::time_point t0 = Clock::now();
                backend_submit(c.backend, &c, g.views, g.set, true);
                const f64 ms = std::chrono::duration<f64, std::milli>(Clock::now() - t0).count();
                lk.lock();
                sv->stats.host_submit_ms += ms;
                continue;
            }
            sv->stats.admission_waits += 1;
        }
        if (!sv->inflight.empty()) {
            InFlight& f = sv->inflight.front();
            /* complete the oldest batch when it is done, when every slot is
               busy, or when waiting requests need its memory; otherwise keep
               admitting arrivals while it runs */
            if (backend_batch_done(c.backend, f.set) || (size_t)kServerBatchesInFlight == sv->inflight.size() ||
                !sv->pending.empty()) {
                lk.unlock();
                const Clock::time_point t0 = Clock::now();
                backend_complete(c.backend, &c, f.views, f.set);
                const Clock::time_point now = Clock::now();
                lk.lock();
                sv->stats.host_complete_ms += std::chrono::duration<f64, std::milli>(now - t0).count();
                sv->stats.gpu_ms += (f64)(c.stats.ms_total - c.stats.ms_upload - c.stats.ms_download);
                for (Request* q : f.reqs) {
                    q->t_done = now;
                    q->done = true;
                }
                sv->stats.completed += f.reqs.size();
                sv->inflight_bytes -= f.footprint;
                sv->inflight.pop_front();
                sv->cv_done.notify_all();
            }
            else {
                sv->cv_work.wait_for(lk, std::chrono::microseconds(50));
            }
            continue;
        }
        sv->cv_work.wait(lk, [&] { return !sv->pending.empty() || !sv->running; });
    }
}

}  // namespace

}  // namespace sft

using namespace sft;

extern "C" i32 sftx_server_template(void) {
    Context& c = *g_ctx;
    auto t = std::make_unique<Template>();
    snapshot_view(c, t->frame);
    t->frame.capture_flags = 0;
    /* transient models would die with the frame: keep copies, drawn as resident models */
    for (const sft_model* m : t->frame.transient_models) {
        auto o = copy_dmesh_model(m);
        for (sft_instance& in : t->frame.instances) {
            if (in.model == m) in.model = &o->model;
        }
        t->owned.push_back(std::move(o));
    }
    t->frame.transient_models.clear();
    /* capsule shadows: the records every request shares, filled by the first batch that traces them */
    if (t->frame.shadows_enabled && !t->frame.point_lights.empty() && !t->frame.proxies.empty()) {
        t->frame.shadow_cache = std::make_shared<ShadowCache>();
    }
    /* poses as they are now */
    size_t nb = 0;
    for (const sft_instance& in : t->frame.instances) {
        if (in.posed_bone_matrices) nb += (size_t)std::max(in.model->skeleton.bone_count, 1);
    }
    t->frame.bone_store.resize(nb);
    t->frame.bones_kept.assign(t->frame.instances.size(), 0);
    size_t at = 0;
    for (size_t i = 0; i < t->frame.instances.size(); i += 1) {
        sft_instance& in = t->frame.instances[i];
        if (!in.posed_bone_matrices) continue;
        const size_t n = (size_t)std::max(in.model->skeleton.bone_count, 1);
        memcpy(&t->frame.bone_store[at], in.posed_bone_matrices, n * sizeof(sftmat));
        in.posed_bone_matrices = &t->frame.bone_store[at];
        t->frame.bones_kept[i] = (u32)n;
        at += n;
    }
    if (!g_server) g_server = new Server();
    std::lock_guard<std::mutex> lk(g_server->mu);
    g_server->templates.push_back(std::move(t));
    return (i32)g_server->templates.size() - 1;
}

extern "C" i32 sftx_server_start(u64 hbm_budget_bytes, i32 max_views_per_batch) {
    Context& c = *g_ctx;
    if (!g_server) g_server = new Server();
    Server* sv = g_server;
    if (sv->running) return -1;
    if (!c.backend) c.backend = backend_create();
    drain_async_batches(c);
    sv->budget = hbm_budget_bytes;
    sv->max_views = std::max(1, max_views_per_batch);
    memset(&sv->stats, 0, sizeof(sv->stats));
    sv->t_start = Clock::now();
    sv->running = true;
    c.server_running = true;
    sv->thread = std::thread(server_loop, sv);
    return 0;
}

extern "C" u64 sftx_server_submit(const sftx_request* r) {
    Server* sv = g_server;
    if (!sv || !sv->running || !r) return 0;
    std::lock_guard<std::mutex> lk(sv->mu);
    if (r->scene < 0 || (size_t)r->scene >= sv->templates.size() || !view_size_ok(r->width, r->height) ||
        (1 != r->samples && 2 != r->samples && 4 != r->samples)) {
        fprintf(stderr, "sofren2: invalid render request (scene %d, %dx%d, %d samples)\n", r->scene, r->width, r->height,
                r->samples);
        return 0;
    }
    auto q = std::make_unique<Request>();
    q->r = *r;
    q->id = sv->next_id++;
    q->footprint = request_footprint(*r, *sv->templates[(size_t)r->scene]);
    q->t_submit = Clock::now();
    sv->pending.push_back(q.get());
    const u64 id = q->id;
    sv->requests[id] = std::move(q);
    sv->stats.submitted += 1;
    sv->cv_work.notify_one();
    return id;
}

extern "C" i32 sftx_server_poll(u64 id) {
    Server* sv = g_server;
    if (!sv) return -1;
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->requests.find(id);
    if (it == sv->requests.end()) return -1;
    return it->second->done ? 1 : 0;
}

extern "C" i32 sftx_server_wait(u64 id, f64* latency_ms) {
    Server* sv = g_server;
    if (!sv) return -1;
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->requests.find(id);
    if (it == sv->requests.end()) return -1;
    Request* q = it->second.get();
    sv->cv_done.wait(lk, [&] { return q->done; });
    if (latency_ms) *latency_ms = std::chrono::duration<f64, std::milli>(q->t_done - q->t_submit).count();
    sv->requests.erase(it);
    return 0;
}

extern "C" void sftx_server_get_stats(sftx_server_stats* out) {
    Server* sv = g_server;
    if (!out) return;
    memset(out, 0, sizeof(*out));
    if (!sv) return;
    std::lock_guard<std::mutex> lk(sv->mu);
    *out = sv->stats;
    out->resident_bytes = g_ctx && g_ctx->backend ? backend_resident_bytes(g_ctx->backend) : 0;
    out->pending = sv->pending.size();
    out->in_flight = sv->inflight.size();
}

extern "C" void sftx_server_stop(void) {
    Server* sv = g_server;
    if (!sv || !sv->running) return;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->running = false;
        sv->cv_work.notify_all();
    }
    sv->thread.join();
    g_ctx->server_running = false;
}

extern "C" void sftx_server_release(void) {
    sftx_server_stop();
    if (!g_server) return;
    /* the copies' addresses may be reused: drop them from the residency map */
    if (g_ctx && g_ctx->backend) {
        for (const auto& t : g_server->templates) {
            for (const auto& o : t->owned) backend_forget_model(g_ctx->backend, &o->model);
        }
    }
    delete g_server;
    g_server = nullptr;
}
