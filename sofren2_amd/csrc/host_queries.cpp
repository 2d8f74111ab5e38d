/*
 * host_queries.cpp — the reference API's host-side queries around the render
 * path: skeletal pose evaluation (sft_pose_create / sft_pose_update,
 * sofren2.c:8940-9107) and the single-ray pick query sft_raycast_instance
 * (sofren2.c:3705-3838).  Both run on the calling thread, as in the
 * reference: a pose feeds sft_instance::posed_bone_matrices, which the GPU
 * reads at sft_present; a pick is one ray, far too little work for a launch.
 *
 * Compiled with -ffp-contract=off -fno-builtin, like every host file whose
 * floats must equal the reference build's: the scalar math is the header's
 * (include/sofren2.h), in the reference's operation order.
 */
#include <math.h>
#include <string.h>
#include <vector>
#include "sft_internal.h"

using namespace sft;

/* ================================================================ poses */

/* sofren2.c:8940-8950: a pose exists only for a model with a skeleton */
extern "C" sft_pose* sft_pose_create(const sft_model* model, sft_arena* arena) {
    if (!model || model->skeleton.bone_count <= 0 || !arena) return NULL;
    const i32 n = model->skeleton.bone_count;
    sft_pose* p = (sft_pose*)sft_arena_alloc(arena, sizeof(sft_pose));
    p->model = model;
    p->global_matrices = (sftmat*)sft_arena_alloc(arena, (u64)n * sizeof(sftmat));
    p->posed_bone_matrices = (sftmat*)sft_arena_alloc(arena, (u64)n * sizeof(sftmat));
    for (i32 i = 0; i < n; i += 1) {
        p->global_matrices[i] = sft_m_identity();
        p->posed_bone_matrices[i] = sft_m_identity();
    }
    return p;
}

namespace {

/* the keyframe pair around `time` and the blend factor between them
   (sofren2.c:8975-9003).  Keyframe times ascend (glTF requires it); before
   the first or after the last key both indices name that key and t = 0. */
struct KeySpan { i32 f0, f1; f32 t; };

KeySpan key_span(const sft_anim_sampler& s, f32 time) {
    KeySpan k{0, 0, 0.f};
    const i32 last = s.count - 1;
    if (time <= s.inputs[0]) return k;
    if (time >= s.inputs[last]) {
        k.f0 = k.f1 = last;
        return k;
    }
    /* the interval [inputs[lo], inputs[lo + 1]) holding time */
    i32 lo = 0, hi = last;
    while (hi - lo > 1) {
        const i32 mid = lo + (hi - lo) / 2;
        if (time < s.inputs[mid]) hi = mid;
        else lo = mid;
    }
    k.f0 = lo;
    k.f1 = lo + 1;
    const f32 dt = s.inputs[k.f1] - s.inputs[k.f0];
    k.t = (dt > 0.0001f) ? (time - s.inputs[k.f0]) / dt : 0.f;
    return k;
}

sftvec key_vec(const sft_anim_sampler& s, i32 f) {
    return SFT_V(s.outputs[f * 4 + 0], s.outputs[f * 4 + 1], s.outputs[f * 4 + 2]);
}

sftquat key_quat(const sft_anim_sampler& s, i32 f) {
    sftquat q;
    q.x = s.outputs[f * 4 + 0];
    q.y = s.outputs[f * 4 + 1];
    q.z = s.outputs[f * 4 + 2];
    q.w = s.outputs[f * 4 + 3];
    return q;
}

/* the animated parts of one bone's local transform */
struct BoneTrack {
    sftvec t{0, 0, 0}, s{0, 0, 0};
    sftquat r{0, 0, 0, 0};
    bool has_t = false, has_r = false, has_s = false;
};

/* global = local * global(parent), parents first (sofren2.c:9073-9097);
   `state` 0 = not started, 1 = in progress (a cycle), 2 = done */
void flatten(sft_pose* pose, i32 i, std::vector<u8>& state) {
    if (2 == state[i]) return;
    const struct sft_bone* bones = pose->model->skeleton.bones;
    const i32 parent = bones[i].parent_ind;
    if (parent >= 0 && parent < pose->model->skeleton.bone_count && 1 != state[parent]) {
        state[i] = 1;
        flatten(pose, parent, state);
        pose->global_matrices[i] = sft_m_mul(pose->global_matrices[i], pose->global_matrices[parent]);
    }
    state[i] = 2;
}

}  // namespace

/* Samples animation `anim_id` at `time` (wrapped to its duration) into the
   pose's per-bone global matrices and posed_bone_matrices = inverse_bind *
   global (sofren2.c:8952-9107): per channel the keyframe span, then
   translation/scale lerp and rotation slerp (or the span's first key for
   step samplers); a bone with any animated part gets S * R * T from its
   animated and base parts, any other bone its local bind matrix. */
extern "C" void sft_pose_update(sft_pose* pose, i32 anim_id, f32 time) {
    if (!pose || !pose->model || !pose->model->animations || anim_id < 0 || anim_id >= pose->model->anim_count) return;
    const sft_model* model = pose->model;
    const sft_animation& anim = model->animations[anim_id];
    const i32 nb = model->skeleton.bone_count;
    if (anim.duration > SFT_EPSILON) time = fmodf(time, anim.duration);

    std::vector<BoneTrack> track((size_t)(nb > 0 ? nb : 0));
    for (i32 c = 0; c < anim.channel_count; c += 1) {
        const sft_anim_channel& ch = anim.channels[c];
        if (ch.node_ind < 0 || ch.node_ind >= nb) continue;
        const sft_anim_sampler& s = anim.samplers[ch.sampler_ind];
        const KeySpan k = key_span(s, time);
        BoneTrack& b = track[(size_t)ch.node_ind];
        if (SFT_ANIM_PATH_ROTATION == ch.path) {
            b.r = s.is_step ? key_quat(s, k.f0) : sft_q_slerp(key_quat(s, k.f0), key_quat(s, k.f1), k.t);
            b.has_r = true;
            continue;
        }
        sftvec v = key_vec(s, k.f0);
        if (!s.is_step) {
            const sftvec v1 = key_vec(s, k.f1);
            v = sft_v_add(v, sft_v_mul(sft_v_sub(v1, v), k.t));
        }
        if (SFT_ANIM_PATH_TRANSLATION == ch.path) {
            b.t = v;
            b.has_t = true;
        }
        else if (SFT_ANIM_PATH_SCALE == ch.path) {
            b.s = v;
            b.has_s = true;
        }
    }

    for (i32 i = 0; i < nb; i += 1) {
        const BoneTrack& b = track[(size_t)i];
        const struct sft_bone& bone = model->skeleton.bones[i];
        if (!(b.has_t || b.has_r || b.has_s)) {
            pose->global_matrices[i] = bone.local_bind_matrix;
            continue;
        }
        const sftvec t = b.has_t ? b.t : bone.base_t;
        const sftvec s = b.has_s ? b.s : bone.base_s;
        const sftmat T = sft_m_translate(t.x, t.y, t.z);
        const sftmat R = sft_m_from_quat(b.has_r ? b.r : bone.base_r);
        const sftmat S = sft_m_scale(s.x, s.y, s.z);
        pose->global_matrices[i] = sft_m_mul(S, sft_m_mul(R, T));
    }
    std::vector<u8> state((size_t)(nb > 0 ? nb : 0), 0);
    for (i32 i = 0; i < nb; i += 1) flatten(pose, i, state);
    for (i32 i = 0; i < nb; i += 1) {
        pose->posed_bone_matrices[i] = sft_m_mul(model->skeleton.bones[i].inverse_bind_matrix, pose->global_matrices[i]);
    }
}

/* ================================================================ picking */

namespace {

/* Moller-Trumbore with the reference's rejection order and exact division
   (sofren2.c:3705-3731) */
bool ray_triangle(sftvec ro, sftvec rd, sftvec v0, sftvec v1, sftvec v2, f32* t) {
    const sftvec e1 = sft_v_sub(v1, v0);
    const sftvec e2 = sft_v_sub(v2, v0);
    const sftvec p = sft_v_cross(rd, e2);
    const f32 det = sft_v_dot(e1, p);
    if (fabsf(det) < SFT_EPSILON) return false;
    const f32 inv_det = 1.f / det;
    const sftvec s = sft_v_sub(ro, v0);
    const f32 u = sft_v_dot(s, p) * inv_det;
    if (u < 0.f || u > 1.f) return false;
    const sftvec q = sft_v_cross(s, e1);
    const f32 v = sft_v_dot(rd, q) * inv_det;
    if (v < 0.f || u + v > 1.f) return false;
    *t = sft_v_dot(e2, q) * inv_det;
    return *t > SFT_EPSILON;
}

/* 1 / d with |d| raised to at least epsilon, keeping its sign */
f32 safe_inv(f32 d) {
    return 1.f / (fabsf(d) > SFT_EPSILON ? d : SFT_EPSILON * (d < 0 ? -1 : 1));
}

/* row-vector 3x3 product d * M (the linear part of a transform) */
sftvec mul_dir(sftvec d, const sftmat& m) {
    sftvec r;
    r.x = d.x * m.m[0][0] + d.y * m.m[1][0] + d.z * m.m[2][0];
    r.y = d.x * m.m[0][1] + d.y * m.m[1][1] + d.z * m.m[2][1];
    r.z = d.x * m.m[0][2] + d.y * m.m[1][2] + d.z * m.m[2][2];
    return r;
}

}  // namespace

/* Closest hit of the ray origin + s * dir, s < max_dist, with the
   instance's model, through its BLAS (sofren2.c:3733-3838).  The ray goes to
   model space by the rigid inverse of the instance transform; its direction
   is renormalised there and distances scale back.  Nodes are visited from an
   explicit stack in the reference's order (children pushed in index order,
   last pushed first), so equal-distance hits resolve to the same triangle.
   A model without a BLAS returns no hit at max_dist. */
extern "C" sft_ray_hit sft_raycast_instance(const sft_instance* inst, sftvec origin, sftvec dir, f32 max_dist) {
    sft_ray_hit hit;
    memset(&hit, 0, sizeof(hit));
    hit.distance = max_dist;
    if (!inst || !inst->model || 0 == inst->model->blas_node_count || !inst->model->blas_nodes) return hit;
    const sft_model* model = inst->model;

    const sftmat inv = sft_m_qinv(inst->transform);
    const sftvec lo = sft_m_mul_vec(inv, origin);
    sftvec ld = mul_dir(dir, inv);
    const f32 inv_scale = 1.f / sqrtf(ld.x * ld.x + ld.y * ld.y + ld.z * ld.z);
    ld = sft_v_mul(ld, inv_scale);
    const f32 ix = safe_inv(ld.x), iy = safe_inv(ld.y), iz = safe_inv(ld.z);
    f32 best = max_dist / inv_scale;

    std::vector<u32> stack;
    stack.reserve(64);
    stack.push_back(model->blas_root_ind);
    while (!stack.empty()) {
        const u32 ni = stack.back();
        stack.pop_back();
        if (ni >= model->blas_node_count) continue;
        const sft_bvh_node& node = model->blas_nodes[ni];
        for (u32 c = 0; c < node.child_count && c < 8u; c += 1) {
            const u32 child = node.child_indices[c];
            if (0xFFFFFFFFu == child) continue;
            const f32 tx1 = (node.min_x[c] - lo.x) * ix, tx2 = (node.max_x[c] - lo.x) * ix;
            const f32 ty1 = (node.min_y[c] - lo.y) * iy, ty2 = (node.max_y[c] - lo.y) * iy;
            const f32 tz1 = (node.min_z[c] - lo.z) * iz, tz2 = (node.max_z[c] - lo.z) * iz;
            const f32 tmin = SFT_MAX(SFT_MAX(SFT_MIN(tx1, tx2), SFT_MIN(ty1, ty2)), SFT_MIN(tz1, tz2));
            const f32 tmax = SFT_MIN(SFT_MIN(SFT_MAX(tx1, tx2), SFT_MAX(ty1, ty2)), SFT_MAX(tz1, tz2));
            if (!(tmax >= tmin && tmax > 0.f && tmin < best)) continue;
            if (!node.is_leaf_group) {
                stack.push_back(child);
                continue;
            }
            /* leaf payload: model meshlet << 7 | triangle */
            const u32 payload = child & ~0x80000000u;
            const u32 gm = payload >> SFT_TRIANGLE_BITS;
            const u32 tri = payload & ((1u << SFT_TRIANGLE_BITS) - 1u);
            if (gm >= model->total_meshlets) continue;
            const sft_mesh& mesh = model->meshes[model->meshlet_mesh_inds[gm]];
            const u32 k = gm - mesh.model_meshlet_offset;
            const sft_meshlet& mv = mesh.meshlets[k];
            const u8* ind = mesh.inds[k].inds + (tri / 8u) * 32u + (tri % 8u);
            const u8 i0 = ind[0], i1 = ind[8], i2 = ind[16];
            const sftvec v0 = SFT_V(mv.x[i0], mv.y[i0], mv.z[i0]);
            const sftvec v1 = SFT_V(mv.x[i1], mv.y[i1], mv.z[i1]);
            const sftvec v2 = SFT_V(mv.x[i2], mv.y[i2], mv.z[i2]);
            f32 t = 0.f;
            if (ray_triangle(lo, ld, v0, v1, v2, &t) && t < best) {
                best = t;
                hit.hit = 1;
                hit.distance = t * inv_scale;
                hit.pos = sft_m_mul_vec(inst->transform, sft_v_add(lo, sft_v_mul(ld, t)));
                hit.normal = sft_v_norm(mul_dir(sft_v_face_normal(v0, v1, v2), inst->transform));
            }
        }
    }
    return hit;
}
