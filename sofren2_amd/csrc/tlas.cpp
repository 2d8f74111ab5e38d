/*
 * tlas.cpp — the 8-wide BVH builder of the reference (sofren2.c:2895-3062)
 * and the frame's top-level BVH over the drawn instances, built on the
 * host at present time (sofren2.c:3619-3703) and uploaded with the frame's
 * tables when the device build is switched off.  The same builder makes a
 * model's triangle BLAS (builders.cpp).  The shadow-ray traversal on the GPU culls TLAS nodes against the
 * packet frustum, so the node grouping is kept identical to the reference
 * builder (median splits into up to 8 groups, sofren2.c:2895-3062): a different
 * grouping would change the bounds a packet is tested against.
 */
#include <string.h>
#include "sft_internal.h"

namespace sft {

namespace {

using Item = BvhItem;

/* Hoare partition around the middle element's key, recursing into both sides
   (sofren2.c:2895-2924) */
void sort_items(Item* it, i32 n, int axis) {
    while (n > 1) {
        const f32 pivot = it[n / 2].mid[axis];
        i32 i = 0, j = n - 1;
        while (i <= j) {
            while (it[i].mid[axis] < pivot) i += 1;
            while (it[j].mid[axis] > pivot) j -= 1;
            if (i <= j) {
                const Item t = it[i];
                it[i] = it[j];
                it[j] = t;
                i += 1;
                j -= 1;
            }
        }
        if (j > 0) sort_items(it, j + 1, axis);
        if (i >= n) return;
        it += i;
        n -= i;
    }
}

/* three levels of median splits along the widest centroid axis give up to 8
   groups (sofren2.c:2926-2965) */
void split8(Item* it, u32 start, u32 count, u32 level, u32* starts, u32* counts, u32* ngroups) {
    if (0 == count) return;
    if (3 == level || 1 == count) {
        starts[*ngroups] = start;
        counts[*ngroups] = count;
        *ngroups += 1;
        return;
    }
    f32 lo[3] = { 1e30f, 1e30f, 1e30f }, hi[3] = { -1e30f, -1e30f, -1e30f };
    for (u32 i = start; i < start + count; i += 1) {
        for (int k = 0; k < 3; k += 1) {
            if (it[i].mid[k] < lo[k]) lo[k] = it[i].mid[k];
            if (it[i].mid[k] > hi[k]) hi[k] = it[i].mid[k];
        }
    }
    const f32 ext[3] = { hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2] };
    int axis = 0;
    if (ext[1] > ext[0] && ext[1] > ext[2]) axis = 1;
    else if (ext[2] > ext[0] && ext[2] > ext[1]) axis = 2;
    sort_items(it + start, (i32)count, axis);
    const u32 half = count / 2;
    split8(it, start, half, level + 1, starts, counts, ngroups);
    split8(it, start + half, count - half, level + 1, starts, counts, ngroups);
}

void grow(f32* bounds, const f32* lo, const f32* hi) {
    for (int k = 0; k < 3; k += 1) {
        if (lo[k] < bounds[k]) bounds[k] = lo[k];
        if (hi[k] > bounds[3 + k]) bounds[3 + k] = hi[k];
    }
}

/* 8-wide node over items[0, n); returns its index, writes its bounds
   (sofren2.c:2967-3062) */
u32 build_node(Item* it, u32 n, std::vector<sft_bvh_node>& nodes, f32* bounds) {
    const u32 idx = (u32)nodes.size();
    nodes.emplace_back();
    {
        sft_bvh_node& nd = nodes[idx];
        memset(&nd, 0, sizeof(nd));
        for (int c = 0; c < 8; c += 1) {
            nd.min_x[c] = nd.min_y[c] = nd.min_z[c] = 1e30f;
            nd.max_x[c] = nd.max_y[c] = nd.max_z[c] = -1e30f;
            nd.child_indices[c] = kInvalid;
        }
    }
    f32 b[6] = { 1e30f, 1e30f, 1e30f, -1e30f, -1e30f, -1e30f };
    auto set_child = [&](u32 c, const f32* lo, const f32* hi, u32 child) {
        sft_bvh_node& nd = nodes[idx];
        nd.min_x[c] = lo[0]; nd.min_y[c] = lo[1]; nd.min_z[c] = lo[2];
        nd.max_x[c] = hi[0]; nd.max_y[c] = hi[1]; nd.max_z[c] = hi[2];
        nd.child_indices[c] = child;
        grow(b, lo, hi);
    };
    if (n <= 8) {
        nodes[idx].is_leaf_group = 1;
        nodes[idx].child_count = (u8)n;
        for (u32 i = 0; i < n; i += 1) set_child(i, it[i].lo, it[i].hi, 0x80000000u | it[i].id);
    }
    else {
        u32 starts[8] = {0}, counts[8] = {0}, ng = 0;
        split8(it, 0, n, 0, starts, counts, &ng);
        nodes[idx].is_leaf_group = 0;
        nodes[idx].child_count = (u8)ng;
        for (u32 g = 0; g < ng; g += 1) {
            f32 cb[6];
            const u32 child = build_node(it + starts[g], counts[g], nodes, cb);
            set_child(g, cb, cb + 3, child);
        }
    }
    if (bounds) memcpy(bounds, b, sizeof(b));
    return idx;
}

}  // namespace

/* The 8-wide BVH over items[0, n) (reordered in place), appended to `nodes`;
   returns the root's index.  Shared by the frame TLAS and the model BLAS
   builder (sofren2.c:2967-3062). */
u32 build_bvh(BvhItem* items, u32 n, std::vector<sft_bvh_node>& nodes) {
    return build_node(items, n, nodes, nullptr);
}

/* nodes of the TLAS over `items` items: one leaf root up to 8 items, else the
   root and one leaf per group of the three halving levels (every group holds
   at most 8 items up to 64) */
static u32 halving_groups(u32 count, u32 level) {
    if (0 == count) return 0;
    if (3 == level || 1 == count) return 1;
    return halving_groups(count / 2, level + 1) + halving_groups(count - count / 2, level + 1);
}

u32 tlas_node_count(u32 items) { return (items <= 8u) ? 1u : 1u + halving_groups(items, 0); }

/* Items are indexed by draw slot while only instances with a BLAS are counted:
   an instance without one, drawn before one with a BLAS, leaves a zeroed item
   inside the built range (the reference leaves that item unwritten). */
u32 build_tlas(const std::vector<sft_instance>& instances, std::vector<sft_bvh_node>& nodes) {
    nodes.clear();
    const u32 n = (u32)instances.size();
    if (0 == n) return 0;
    std::vector<Item> items(n);
    memset(items.data(), 0, n * sizeof(Item));
    u32 built = 0;
    for (u32 i = 0; i < n; i += 1) {
        const sft_instance& in = instances[i];
        const sft_model* m = in.model;
        if (0 == m->blas_node_count) continue;
        const sft_bvh_node& r = m->blas_nodes[m->blas_root_ind];
        f32 lo[3] = { 1e30f, 1e30f, 1e30f }, hi[3] = { -1e30f, -1e30f, -1e30f };
        for (u32 c = 0; c < r.child_count; c += 1) {
            const f32 clo[3] = { r.min_x[c], r.min_y[c], r.min_z[c] }, chi[3] = { r.max_x[c], r.max_y[c], r.max_z[c] };
            for (int k = 0; k < 3; k += 1) {
                if (clo[k] < lo[k]) lo[k] = clo[k];
                if (chi[k] > hi[k]) hi[k] = chi[k];
            }
        }
        Item& it = items[i];
        for (int k = 0; k < 3; k += 1) {
            it.lo[k] = 1e30f;
            it.hi[k] = -1e30f;
        }
        const sftmat& T = in.transform;
        for (int corner = 0; corner < 8; corner += 1) {
            const f32 x = (corner & 1) ? hi[0] : lo[0];
            const f32 y = (corner & 2) ? hi[1] : lo[1];
            const f32 z = (corner & 4) ? hi[2] : lo[2];
            f32 w[3];
            for (int k = 0; k < 3; k += 1) w[k] = x * T.m[0][k] + y * T.m[1][k] + z * T.m[2][k] + T.m[3][k];
            for (int k = 0; k < 3; k += 1) {
                if (w[k] < it.lo[k]) it.lo[k] = w[k];
                if (w[k] > it.hi[k]) it.hi[k] = w[k];
            }
        }
        it.id = i;
        for (int k = 0; k < 3; k += 1) it.mid[k] = (it.lo[k] + it.hi[k]) * 0.5f;
        built += 1;
    }
    nodes.reserve((size_t)built * 2 + 1);
    return build_node(items.data(), built, nodes, nullptr);
}

}  // namespace sft
