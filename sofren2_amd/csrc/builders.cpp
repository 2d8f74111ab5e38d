/*
 * builders.cpp — the reference's offline asset builders, for applications
 * that construct models in memory instead of loading QSF2 files:
 *   - sftx_mesh_build: triangle soup -> 64-vertex / 128-triangle meshlets with
 *     bounding spheres, boxes and normal cones (sofren2.c:2526-2813, the
 *     builder helpers 2180-2488);
 *   - sftx_mesh_build_attribs: the same for a mesh with normals, UVs and
 *     skin weights, the glTF importer's variant (sofren2.c:7420-7660,
 *     finalisation 6916-7045);
 *   - sftx_model_link: model-relative meshlet offsets and the meshlet -> mesh
 *     table (sofren2.c:7668-7681);
 *   - sftx_blas_build: the model's 8-wide triangle BVH (sofren2.c:3526-3617);
 *   - sftx_shadow_proxies_build: capsule shadow proxies, per bone cluster for
 *     skinned models, per BVH subtree otherwise (sofren2.c:3065-3524).
 * They run on the calling thread, once per asset, like the reference's: the
 * greedy meshlet growth and the Hoare-partition BVH are sequential decisions
 * whose outcome (which triangle joins which meshlet, which items share a BVH
 * group) the visibility IDs, cull results and shadow-ray traversal of the GPU
 * path depend on, so their order is kept exactly.
 *
 * Compiled with -ffp-contract=off -fno-builtin: every float (spheres, cones,
 * capsule fits) equals the reference build's.  The bounding sphere is Welzl's
 * algorithm with the C library's rand() choosing points, as in the reference:
 * the results match the reference's when the process's rand() stream is in
 * the same state.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <vector>
#include "sft_internal.h"

using namespace sft;

namespace {

constexpr u32 kVerts = SFT_VERTS_PER_MESHLET;
constexpr u32 kTris = 128;
constexpr f32 kConeLimit = -0.173648f;  // cos(100 degrees), sofren2.c:152

struct Sphere {
    sftvec c;
    f32 r;
};

/* ---------------------------------------------------- Welzl's sphere */

/* sofren2.c:2180-2185 */
Sphere sphere2(sftvec a, sftvec b) {
    Sphere s;
    s.c = sft_v_mul(sft_v_add(a, b), 0.5f);
    s.r = sft_v_length(sft_v_sub(a, b)) * 0.5f;
    return s;
}

/* circumsphere of a triangle; a degenerate one falls back to its first edge
   (sofren2.c:2187-2204) */
Sphere sphere3(sftvec p0, sftvec p1, sftvec p2) {
    const sftvec a = sft_v_sub(p1, p0), b = sft_v_sub(p2, p0);
    const sftvec n = sft_v_cross(a, b);
    const f32 denom = 2.f * sft_v_length2(n);
    if (denom < SFT_EPSILON) return sphere2(p0, p1);
    const sftvec t = sft_v_sub(sft_v_mul(b, sft_v_length2(a)), sft_v_mul(a, sft_v_length2(b)));
    Sphere s;
    s.c = sft_v_add(p0, sft_v_div(sft_v_cross(t, n), denom));
    s.r = sft_v_length(sft_v_sub(s.c, p0));
    return s;
}

/* circumsphere of a tetrahedron from the 3x3 system's cofactors; coplanar
   points fall back to the first three (sofren2.c:2206-2243) */
Sphere sphere4(sftvec p0, sftvec p1, sftvec p2, sftvec p3) {
    const sftvec r1 = sft_v_sub(p1, p0), r2 = sft_v_sub(p2, p0), r3 = sft_v_sub(p3, p0);
    const f32 q1 = sft_v_length2(r1), q2 = sft_v_length2(r2), q3 = sft_v_length2(r3);
    /* 2x2 minors of the rows r1, r2, r3 */
    const f32 m23x = r2.y*r3.z - r3.y*r2.z, m13x = r1.y*r3.z - r3.y*r1.z, m12x = r1.y*r2.z - r2.y*r1.z;
    const f32 m23y = r2.x*r3.z - r3.x*r2.z, m13y = r1.x*r3.z - r3.x*r1.z, m12y = r1.x*r2.z - r2.x*r1.z;
    const f32 m23z = r2.x*r3.y - r3.x*r2.y, m13z = r1.x*r3.y - r3.x*r1.y, m12z = r1.x*r2.y - r2.x*r1.y;
    const f32 det = r1.x * m23x - r1.y * m23y + r1.z * m23z;
    if (fabsf(det) < SFT_EPSILON) return sphere3(p0, p1, p2);
    const f32 f = 0.5f / det;
    sftvec o;
    o.x = f * (m23x * q1 - m13x * q2 + m12x * q3);
    o.y = f * (-m23y * q1 + m13y * q2 - m12y * q3);
    o.z = f * (m23z * q1 - m13z * q2 + m12z * q3);
    Sphere s;
    s.c = sft_v_add(p0, o);
    s.r = sft_v_length(o);
    return s;
}

/* sofren2.c:2245-2253 */
Sphere sphere_on(const sftvec* b, int nb) {
    switch (nb) {
        case 0: { Sphere s; memset(&s, 0, sizeof(s)); return s; }
        case 1: { Sphere s; s.c = b[0]; s.r = 0.f; return s; }
        case 2: return sphere2(b[0], b[1]);
        case 3: return sphere3(b[0], b[1], b[2]);
        default: return sphere4(b[0], b[1], b[2], b[3]);
    }
}

/* smallest sphere of p[0, n) with the points of b[0, nb) on its boundary;
   the point left out is drawn with rand() (sofren2.c:2255-2278) */
Sphere welzl(sftvec* p, int n, sftvec* b, int nb) {
    if (0 == n || 4 == nb) return sphere_on(b, nb);
    const int k = rand() % n;
    const sftvec q = p[k];
    p[k] = p[n - 1];
    p[n - 1] = q;
    const Sphere s = welzl(p, n - 1, b, nb);
    if (sft_v_length2(sft_v_sub(q, s.c)) <= s.r * s.r + SFT_EPSILON) return s;
    b[nb] = q;
    return welzl(p, n - 1, b, nb + 1);
}

/* sofren2.c:2280-2290: on a copy of the meshlet's vertices */
Sphere bounding_sphere(const sftvec* v, int n) {
    sftvec pts[kVerts], bnd[4];
    for (int i = 0; i < n; i += 1) pts[i] = v[i];
    return welzl(pts, n, bnd, 0);
}

/* ------------------------------------------------- greedy meshlet growth */

struct Growing {
    i32 src[kVerts];     // meshlet vertex -> source vertex
    u32 nv;
    u8 tri[kTris * 3];   // meshlet-local corners
    u32 nt;
    sftvec axis;         // running normal sum, renormalised after each triangle
    void clear() { nv = 0; nt = 0; axis.x = axis.y = axis.z = 0.f; }
};

/* the candidate's score: shared vertices plus half its agreement with the
   running normal, -1 when it would break the 128-triangle, 64-vertex or
   100-degree cone limits (sofren2.c:2323-2362) */
f32 score(const Growing& g, i32 a, i32 b, i32 c, sftvec n) {
    if (g.nt >= kTris) return -1.f;
    if (g.nt > 0 && sft_v_dot(g.axis, n) < kConeLimit) return -1.f;
    u32 shared = 0;
    for (u32 i = 0; i < g.nv; i += 1) {
        const i32 s = g.src[i];
        if (s == a || s == b || s == c) shared += 1;
    }
    if (g.nv + (3 - shared) > kVerts) return -1.f;
    f32 sc = (f32)shared;
    if (g.nt > 0) sc += sft_v_dot(g.axis, n) * 0.5f;
    return sc;
}

/* sofren2.c:2364-2408 */
void add_triangle(Growing& g, i32 a, i32 b, i32 c, sftvec n) {
    const i32 corner[3] = { a, b, c };
    for (int k = 0; k < 3; k += 1) {
        u32 local = g.nv;
        for (u32 i = 0; i < g.nv; i += 1) {
            if (g.src[i] == corner[k]) { local = i; break; }
        }
        if (local == g.nv) g.src[g.nv++] = corner[k];
        g.tri[g.nt * 3 + k] = (u8)local;
    }
    g.axis = sft_v_norm(sft_v_add(g.axis, n));
    g.nt += 1;
}

/* triangles into 16 batches of 8 lanes: v0[8] v1[8] v2[8] pad[8], unused
   lanes all zero (sofren2.c:2418-2440) */
void pack_indices(const Growing& g, sft_meshlet_inds* out) {
    memset(out->inds, 0, sizeof(out->inds));
    for (u32 t = 0; t < g.nt; t += 1) {
        const u32 base = (t / 8) * 32 + (t % 8);
        out->inds[base] = g.tri[t * 3 + 0];
        out->inds[base + 8] = g.tri[t * 3 + 1];
        out->inds[base + 16] = g.tri[t * 3 + 2];
    }
}

/* positions into the SoA block (zeros past the last vertex), their box and
   bounding sphere (sofren2.c:2442-2482, 6949-7043) */
void place_positions(const Growing& g, const f32* vx, const f32* vy, const f32* vz, sft_meshlet* m, Sphere* sphere,
                     sft_aabb* box) {
    sftvec pts[kVerts];
    sft_aabb b;
    b.min.x = b.min.y = b.min.z = 1e30f;
    b.max.x = b.max.y = b.max.z = -1e30f;
    for (u32 i = 0; i < kVerts; i += 1) {
        if (i >= g.nv) {
            m->x[i] = m->y[i] = m->z[i] = 0.f;
            continue;
        }
        const i32 s = g.src[i];
        m->x[i] = vx[s];
        m->y[i] = vy[s];
        m->z[i] = vz[s];
        pts[i].x = vx[s];
        pts[i].y = vy[s];
        pts[i].z = vz[s];
        if (vx[s] < b.min.x) b.min.x = vx[s];
        if (vy[s] < b.min.y) b.min.y = vy[s];
        if (vz[s] < b.min.z) b.min.z = vz[s];
        if (vx[s] > b.max.x) b.max.x = vx[s];
        if (vy[s] > b.max.y) b.max.y = vy[s];
        if (vz[s] > b.max.z) b.max.z = vz[s];
    }
    if (g.nv > 0) {
        *sphere = bounding_sphere(pts, (int)g.nv);
        *box = b;
    }
}

/* the final axis and the sine of the widest face-normal deviation from it,
   plus padding; a cone wider than a hemisphere is disabled with sine 1
   (sofren2.c:2772-2803, 7640-7658) */
void cone(const Growing& g, const sft_meshlet* m, f32 pad, sftvec* axis, f32* cutoff) {
    const sftvec ax = sft_v_norm(g.axis);
    f32 min_dot = 1.f;
    for (u32 t = 0; t < g.nt; t += 1) {
        const u8 a = g.tri[t * 3], b = g.tri[t * 3 + 1], c = g.tri[t * 3 + 2];
        sftvec p0, p1, p2;
        p0.x = m->x[a]; p0.y = m->y[a]; p0.z = m->z[a];
        p1.x = m->x[b]; p1.y = m->y[b]; p1.z = m->z[b];
        p2.x = m->x[c]; p2.y = m->y[c]; p2.z = m->z[c];
        const f32 d = sft_v_dot(sft_v_face_normal(p0, p1, p2), ax);
        if (d < min_dot) min_dot = d;
    }
    *axis = ax;
    *cutoff = ((min_dot < 0.f) ? 1.f : sqrtf(1.f - min_dot * min_dot)) + pad;
}

/* per-vertex lists of the triangles that use it, in triangle order
   (sofren2.c:2541-2574) */
struct Adjacency {
    std::vector<i32> first, tris;
    Adjacency(const i32* inds, u32 tri_count, u32 vertex_count) : first(vertex_count + 1, 0), tris((size_t)tri_count * 3) {
        for (u64 i = 0; i < (u64)tri_count * 3; i += 1) first[inds[i]] += 1;
        i32 sum = 0;
        for (u32 v = 0; v < vertex_count; v += 1) {
            const i32 c = first[v];
            first[v] = sum;
            sum += c;
        }
        first[vertex_count] = sum;
        std::vector<i32> head(first.begin(), first.end() - 1);
        for (u32 t = 0; t < tri_count; t += 1) {
            for (int k = 0; k < 3; k += 1) tris[head[inds[t * 3 + k]]++] = (i32)t;
        }
    }
};

/* the candidate triangles next to a growing meshlet, in discovery order;
   taking entry k moves the head entry into its place (sofren2.c:2652-2728) */
struct Frontier {
    std::vector<i32> q;
    std::vector<u8> queued;
    i32 head = 0, tail = 0;
    explicit Frontier(u32 tri_count) : q(tri_count), queued(tri_count, 0) {}
    void push_neighbours(const Adjacency& adj, const std::vector<u8>& used, const i32* corners) {
        for (int k = 0; k < 3; k += 1) {
            for (i32 a = adj.first[corners[k]]; a < adj.first[corners[k] + 1]; a += 1) {
                const i32 t = adj.tris[a];
                if (!used[t] && !queued[t]) {
                    q[tail++] = t;
                    queued[t] = 1;
                }
            }
        }
    }
    /* unqueue what was left; entries before head were taken */
    void reset() {
        for (i32 i = head; i < tail; i += 1) queued[q[i]] = 0;
        head = tail = 0;
    }
};

sftvec face_normal_of(const f32* vx, const f32* vy, const f32* vz, const i32* c) {
    sftvec p[3];
    for (int k = 0; k < 3; k += 1) {
        p[k].x = vx[c[k]];
        p[k].y = vy[c[k]];
        p[k].z = vz[c[k]];
    }
    return sft_v_face_normal(p[0], p[1], p[2]);
}

/* the best frontier entry (first of the highest score), -1 if none fits */
i32 best_candidate(const Frontier& fr, const std::vector<u8>& used, const Growing& g, const i32* inds, const f32* vx,
                   const f32* vy, const f32* vz, sftvec* best_n) {
    f32 best = -1.f;
    i32 at = -1;
    for (i32 i = fr.head; i < fr.tail; i += 1) {
        const i32 t = fr.q[i];
        if (used[t]) continue;
        const i32* c = inds + (u64)t * 3;
        const sftvec n = face_normal_of(vx, vy, vz, c);
        const f32 s = score(g, c[0], c[1], c[2], n);
        if (s > best) {
            best = s;
            at = i;
            *best_n = n;
        }
    }
    return at;
}

/* the per-meshlet arrays of a mesh, allocated in the reference's order */
void alloc_mesh_arrays(sft_mesh* m, sft_arena* arena, u32 cap, bool attribs) {
    m->meshlets = (sft_meshlet*)sft_arena_alloc(arena, (u64)cap * sizeof(sft_meshlet));
    if (attribs) m->meshlet_attribs = (sft_meshlet_attribs*)sft_arena_alloc(arena, (u64)cap * sizeof(sft_meshlet_attribs));
    m->inds = (sft_meshlet_inds*)sft_arena_alloc(arena, (u64)cap * sizeof(sft_meshlet_inds));
    f32** arrays[] = { &m->bounds_center_x, &m->bounds_center_y, &m->bounds_center_z, &m->bounds_radius,
                       &m->bounds_min_x, &m->bounds_min_y, &m->bounds_min_z, &m->bounds_max_x, &m->bounds_max_y,
                       &m->bounds_max_z, &m->cone_axis_x, &m->cone_axis_y, &m->cone_axis_z, &m->cone_cutoff };
    for (f32** a : arrays) *a = (f32*)sft_arena_alloc(arena, (u64)cap * sizeof(f32));
}

void store_bounds(sft_mesh* m, u32 i, const Sphere& s, const sft_aabb& b, sftvec axis, f32 cutoff) {
    m->bounds_center_x[i] = s.c.x;
    m->bounds_center_y[i] = s.c.y;
    m->bounds_center_z[i] = s.c.z;
    m->bounds_radius[i] = s.r;
    m->bounds_min_x[i] = b.min.x;
    m->bounds_min_y[i] = b.min.y;
    m->bounds_min_z[i] = b.min.z;
    m->bounds_max_x[i] = b.max.x;
    m->bounds_max_y[i] = b.max.y;
    m->bounds_max_z[i] = b.max.z;
    m->cone_axis_x[i] = axis.x;
    m->cone_axis_y[i] = axis.y;
    m->cone_axis_z[i] = axis.z;
    m->cone_cutoff[i] = cutoff;
}

bool valid_soup(const char* who, const void* a, const void* b, const void* c, const i32* inds, i64 tri_count) {
    if (tri_count <= 0 || !a || !b || !c || !inds) {
        fprintf(stderr, "[sofren2] %s: no triangles\n", who);
        return false;
    }
    for (i64 i = 0; i < tri_count * 3; i += 1) {
        if (inds[i] < 0) {
            fprintf(stderr, "[sofren2] %s: negative index at %lld\n", who, (long long)i);
            return false;
        }
    }
    return true;
}

}  // namespace

/* Meshlets of a triangle soup: each meshlet grows from the first unplaced
   triangle through the triangles that share a vertex with it, taking the
   best-scoring candidate each step until no candidate fits; positions,
   indices, bounds and cones are written into arena memory sized for one
   meshlet per triangle, as the reference allocates (sofren2.c:2526-2813). */
extern "C" void sftx_mesh_build(sft_mesh* out_mesh, sft_arena* arena, const f32* vx, const f32* vy, const f32* vz,
                                const i32* inds, i32 tri_count) {
    if (!out_mesh || !arena || !valid_soup("sftx_mesh_build", vx, vy, vz, inds, tri_count)) return;
    i32 max_vertex = 0;
    for (i64 i = 0; i < (i64)tri_count * 3; i += 1) max_vertex = inds[i] > max_vertex ? inds[i] : max_vertex;
    const Adjacency adj(inds, (u32)tri_count, (u32)max_vertex + 1);
    const u32 cap = ((u32)tri_count + 7u) & ~7u;
    alloc_mesh_arrays(out_mesh, arena, cap, false);

    std::vector<u8> used((size_t)tri_count, 0);
    Frontier fr((u32)tri_count);
    Growing g;
    i32 left = tri_count;
    i32 seed = 0;  // every triangle before it is placed
    u32 count = 0;
    while (left > 0) {
        g.clear();
        while (seed < tri_count && used[seed]) seed += 1;
        if (seed >= tri_count) break;
        const i32* c = inds + (u64)seed * 3;
        add_triangle(g, c[0], c[1], c[2], face_normal_of(vx, vy, vz, c));
        used[seed] = 1;
        left -= 1;
        fr.push_neighbours(adj, used, c);
        while (g.nv < kVerts && g.nt < kTris && fr.head < fr.tail) {
            sftvec n;
            const i32 at = best_candidate(fr, used, g, inds, vx, vy, vz, &n);
            if (at < 0) break;
            const i32 t = fr.q[at];
            const i32* tc = inds + (u64)t * 3;
            add_triangle(g, tc[0], tc[1], tc[2], n);
            used[t] = 1;
            left -= 1;
            fr.q[at] = fr.q[fr.head++];
            fr.push_neighbours(adj, used, tc);
        }
        fr.reset();
        Sphere s;
        sft_aabb b;
        place_positions(g, vx, vy, vz, &out_mesh->meshlets[count], &s, &b);
        pack_indices(g, &out_mesh->inds[count]);
        sftvec axis;
        f32 cutoff;
        cone(g, &out_mesh->meshlets[count], 0.01f, &axis, &cutoff);
        store_bounds(out_mesh, count, s, b, axis, cutoff);
        count += 1;
    }
    out_mesh->meshlet_count = count;
}

/* The importer's meshlet build for a mesh with attributes: a meshlet whose
   candidates run out (or all fail) keeps growing from the next unplaced
   triangle if that one fits, instead of closing; normals, UVs and skin data
   travel with their vertex, lightmap UVs are zero, and a skinned meshlet's
   culling bone is the joint with the most accumulated weight (first to reach
   the maximum).  Cones are padded by 0.02 (sofren2.c:7420-7660, 6916-7045).
   Returns the meshlet count, 0 on invalid input. */
extern "C" u32 sftx_mesh_build_attribs(sft_mesh* out_mesh, sft_arena* arena, const sftx_raw_geometry* raw,
                                       u32 vertex_count, const i32* inds, i32 tri_count, u32 bone_count) {
    if (!out_mesh || !arena || !raw || !valid_soup("sftx_mesh_build_attribs", raw->vx, raw->vy, raw->vz, inds, tri_count))
        return 0;
    if (!raw->nx || !raw->ny || !raw->nz || !raw->u || !raw->v) {
        fprintf(stderr, "[sofren2] sftx_mesh_build_attribs: normals and UVs are required\n");
        return 0;
    }
    for (i64 i = 0; i < (i64)tri_count * 3; i += 1) {
        if ((u32)inds[i] >= vertex_count) {
            fprintf(stderr, "[sofren2] sftx_mesh_build_attribs: index %d beyond %u vertices\n", inds[i], vertex_count);
            return 0;
        }
    }
    const bool skinned = raw->j0 && raw->j1 && raw->j2 && raw->j3 && raw->w0 && raw->w1 && raw->w2 && raw->w3;
    const f32 *vx = raw->vx, *vy = raw->vy, *vz = raw->vz;
    const Adjacency adj(inds, (u32)tri_count, vertex_count);
    const u32 cap = ((u32)tri_count + 7u) & ~7u;
    alloc_mesh_arrays(out_mesh, arena, cap, true);
    if (skinned) {
        out_mesh->meshlet_skins = (sft_meshlet_skin*)sft_arena_alloc(arena, (u64)cap * sizeof(sft_meshlet_skin));
        out_mesh->culling_bone_inds = (i32*)sft_arena_alloc(arena, (u64)cap * sizeof(i32));
    }
    else {
        out_mesh->meshlet_skins = NULL;
        out_mesh->culling_bone_inds = NULL;
    }

    std::vector<u8> used((size_t)tri_count, 0);
    std::vector<i32> weight_sum(bone_count ? bone_count : 1);
    Frontier fr((u32)tri_count);
    Growing g;
    i32 left = tri_count;
    i32 seed = 0;
    u32 count = 0;
    while (left > 0) {
        g.clear();
        fr.head = fr.tail = 0;
        while (g.nv < kVerts && g.nt < kTris && left > 0) {
            if (fr.head == fr.tail) {
                while (seed < tri_count && used[seed]) seed += 1;
                if (seed >= tri_count) break;
                const i32* c = inds + (u64)seed * 3;
                const sftvec n = face_normal_of(vx, vy, vz, c);
                if (g.nt > 0 && score(g, c[0], c[1], c[2], n) < 0.f) break;
                add_triangle(g, c[0], c[1], c[2], n);
                used[seed] = 1;
                left -= 1;
                fr.push_neighbours(adj, used, c);
                continue;
            }
            sftvec n;
            const i32 at = best_candidate(fr, used, g, inds, vx, vy, vz, &n);
            if (at < 0) {
                /* no candidate fits: the importer drops them without
                   unqueueing, so they come back only as seeds */
                fr.head = fr.tail = 0;
                continue;
            }
            const i32 t = fr.q[at];
            const i32* tc = inds + (u64)t * 3;
            add_triangle(g, tc[0], tc[1], tc[2], n);
            used[t] = 1;
            left -= 1;
            fr.q[at] = fr.q[fr.head++];
            fr.push_neighbours(adj, used, tc);
        }
        fr.reset();

        const u32 mi = count;
        Sphere s;
        memset(&s, 0, sizeof(s));
        sft_aabb b;
        memset(&b, 0, sizeof(b));
        place_positions(g, vx, vy, vz, &out_mesh->meshlets[mi], &s, &b);
        pack_indices(g, &out_mesh->inds[mi]);
        sft_meshlet_attribs* at = &out_mesh->meshlet_attribs[mi];
        sft_meshlet_skin* sk = skinned ? &out_mesh->meshlet_skins[mi] : NULL;
        if (sk) std::fill(weight_sum.begin(), weight_sum.end(), 0);
        i32 cull_bone = 0, cull_weight = 0;
        for (u32 i = 0; i < kVerts; i += 1) {
            if (i < g.nv) {
                const i32 v = g.src[i];
                at->nx[i] = raw->nx[v];
                at->ny[i] = raw->ny[v];
                at->nz[i] = raw->nz[v];
                at->u[i] = raw->u[v];
                at->v[i] = raw->v[v];
                if (sk) {
                    const u16 j[4] = { raw->j0[v], raw->j1[v], raw->j2[v], raw->j3[v] };
                    const u8 w[4] = { raw->w0[v], raw->w1[v], raw->w2[v], raw->w3[v] };
                    sk->j0[i] = j[0]; sk->j1[i] = j[1]; sk->j2[i] = j[2]; sk->j3[i] = j[3];
                    sk->w0[i] = w[0]; sk->w1[i] = w[1]; sk->w2[i] = w[2]; sk->w3[i] = w[3];
                    for (int k = 0; k < 4; k += 1) {
                        if (j[k] >= bone_count) continue;
                        weight_sum[j[k]] += w[k];
                        if (weight_sum[j[k]] > cull_weight) {
                            cull_weight = weight_sum[j[k]];
                            cull_bone = j[k];
                        }
                    }
                }
            }
            else {
                at->nx[i] = at->ny[i] = at->nz[i] = 0.f;
                at->u[i] = at->v[i] = 0.f;
                if (sk) {
                    sk->j0[i] = sk->j1[i] = sk->j2[i] = sk->j3[i] = 0;
                    sk->w0[i] = sk->w1[i] = sk->w2[i] = sk->w3[i] = 0;
                }
            }
            at->lu[i] = 0.f;
            at->lv[i] = 0.f;
        }
        if (sk) out_mesh->culling_bone_inds[mi] = cull_bone;
        sftvec axis;
        f32 cutoff;
        cone(g, &out_mesh->meshlets[mi], 0.02f, &axis, &cutoff);
        store_bounds(out_mesh, mi, s, b, axis, cutoff);
        count += 1;
    }
    out_mesh->meshlet_count = count;
    return count;
}

/* Model-relative meshlet offsets of the model's meshes in order and the
   meshlet -> mesh table (sofren2.c:7668-7681). */
extern "C" void sftx_model_link(sft_model* model, sft_arena* arena) {
    if (!model || !arena) return;
    model->total_meshlets = 0;
    for (u32 i = 0; i < model->mesh_count; i += 1) {
        model->meshes[i].model_meshlet_offset = model->total_meshlets;
        model->total_meshlets += model->meshes[i].meshlet_count;
    }
    if (0 == model->total_meshlets) return;
    model->meshlet_mesh_inds = (u16*)sft_arena_alloc(arena, (u64)model->total_meshlets * sizeof(u16));
    for (u32 i = 0; i < model->mesh_count; i += 1) {
        for (u32 j = 0; j < model->meshes[i].meshlet_count; j += 1)
            model->meshlet_mesh_inds[model->meshes[i].model_meshlet_offset + j] = (u16)i;
    }
}

namespace {

/* calls f(meshlet, local triangle, corner indices) for every triangle of the
   model that is not padding (all three corners equal), in meshlet and batch
   order (sofren2.c:3531-3555) */
template <class F>
void for_each_triangle(const sft_model* m, F f) {
    for (u32 gi = 0; gi < m->total_meshlets; gi += 1) {
        const sft_mesh* mesh = &m->meshes[m->meshlet_mesh_inds[gi]];
        const u32 local = gi - mesh->model_meshlet_offset;
        const u8* in = mesh->inds[local].inds;
        for (u32 t = 0; t < kTris; t += 1) {
            const u32 base = (t / 8) * 32 + (t % 8);
            const u8 a = in[base], b = in[base + 8], c = in[base + 16];
            if (a == b && b == c) continue;
            f(gi, t, mesh, local, a, b, c);
        }
    }
}

}  // namespace

/* The model's triangle BVH: one item per non-padding triangle (its box,
   centre and meshlet << 7 | triangle payload) in meshlet order, grouped by
   the reference's 8-way median splits; node storage for 2n + 1 nodes comes
   from the arena, as in the reference (sofren2.c:3526-3617). */
extern "C" void sftx_blas_build(sft_model* model, sft_arena* arena) {
    if (!model || !arena || 0 == model->total_meshlets) return;
    if (!model->meshlet_mesh_inds) {
        fprintf(stderr, "[sofren2] sftx_blas_build: model has no meshlet table (sftx_model_link)\n");
        return;
    }
    std::vector<BvhItem> items;
    for_each_triangle(model, [&](u32 gi, u32 t, const sft_mesh* mesh, u32 local, u8 a, u8 b, u8 c) {
        const sft_meshlet& mv = mesh->meshlets[local];
        BvhItem it;
        it.id = (gi << SFT_TRIANGLE_BITS) | t;
        const f32 x[3] = { mv.x[a], mv.x[b], mv.x[c] }, y[3] = { mv.y[a], mv.y[b], mv.y[c] },
                  z[3] = { mv.z[a], mv.z[b], mv.z[c] };
        const f32* axes[3] = { x, y, z };
        for (int k = 0; k < 3; k += 1) {
            const f32* v = axes[k];
            it.lo[k] = SFT_MIN(v[0], SFT_MIN(v[1], v[2]));
            it.hi[k] = SFT_MAX(v[0], SFT_MAX(v[1], v[2]));
        }
        for (int k = 0; k < 3; k += 1) it.mid[k] = (it.lo[k] + it.hi[k]) * 0.5f;
        items.push_back(it);
    });
    if (items.empty()) return;
    const u64 max_nodes = (u64)items.size() * 2 + 1;
    std::vector<sft_bvh_node> nodes;
    nodes.reserve(max_nodes);
    const u32 root = build_bvh(items.data(), (u32)items.size(), nodes);
    model->blas_nodes = (sft_bvh_node*)sft_arena_alloc(arena, max_nodes * sizeof(sft_bvh_node));
    memcpy(model->blas_nodes, nodes.data(), nodes.size() * sizeof(sft_bvh_node));
    model->blas_node_count = (u32)nodes.size();
    model->blas_root_ind = root;
}

namespace {

/* capsule along the principal axis of the points (10 power iterations from
   (1,1,1)/sqrt 3 on the covariance), radius min(1.8 x mean, 0.95 x max)
   distance to the segment, ends pushed out by 0.625 r (sofren2.c:3065-3149) */
sft_shadow_proxy fit_capsule(const sftvec* v, i32 n) {
    sft_shadow_proxy out;
    memset(&out, 0, sizeof(out));
    if (0 == n) return out;
    sftvec centroid = { 0.f, 0.f, 0.f };
    for (i32 i = 0; i < n; i += 1) centroid = sft_v_add(centroid, v[i]);
    centroid = sft_v_mul(centroid, 1.f / n);
    f32 cov[3][3] = { { 0.f } };
    for (i32 k = 0; k < n; k += 1) {
        const sftvec d = sft_v_sub(v[k], centroid);
        const f32 a[3] = { d.x, d.y, d.z };
        for (int i = 0; i < 3; i += 1)
            for (int j = 0; j < 3; j += 1) cov[i][j] += a[i] * a[j];
    }
    for (int i = 0; i < 3; i += 1)
        for (int j = 0; j < 3; j += 1) cov[i][j] /= n;
    sftvec axis = sft_v_normf(1.f, 1.f, 1.f);
    for (int it = 0; it < 10; it += 1) {
        sftvec next;
        next.x = cov[0][0] * axis.x + cov[0][1] * axis.y + cov[0][2] * axis.z;
        next.y = cov[1][0] * axis.x + cov[1][1] * axis.y + cov[1][2] * axis.z;
        next.z = cov[2][0] * axis.x + cov[2][1] * axis.y + cov[2][2] * axis.z;
        const f32 len = sft_v_length(next);
        if (len < SFT_EPSILON) break;
        axis = sft_v_mul(next, 1.f / len);
    }
    f32 t0 = 1e9f, t1 = -1e9f;
    for (i32 i = 0; i < n; i += 1) {
        const f32 t = sft_v_dot(sft_v_sub(v[i], centroid), axis);
        if (t < t0) t0 = t;
        if (t > t1) t1 = t;
    }
    const sftvec a = sft_v_add(centroid, sft_v_mul(axis, t0));
    const sftvec ab = sft_v_sub(sft_v_add(centroid, sft_v_mul(axis, t1)), a);
    const f32 len2 = sft_v_length2(ab);
    f32 sum_r = 0.f, max_r = 0.f;
    for (i32 i = 0; i < n; i += 1) {
        const sftvec ap = sft_v_sub(v[i], a);
        const f32 t = (len2 < SFT_EPSILON) ? 0.f : SFT_CLAMP(sft_v_dot(ap, ab) / len2, 0.f, 1.f);
        const f32 r = sqrtf(sft_v_length2(sft_v_sub(v[i], sft_v_add(a, sft_v_mul(ab, t)))));
        sum_r += r;
        if (r > max_r) max_r = r;
    }
    const f32 r = SFT_MIN((sum_r / n) * 1.8f, max_r * 0.95f);
    const f32 seal = r * 0.625f;
    f32 lo = t0 - seal, hi = t1 + seal;
    if (lo > hi) lo = hi = (t0 + t1) * 0.5f;
    out.posa = sft_v_add(centroid, sft_v_mul(axis, lo));
    out.posb = sft_v_add(centroid, sft_v_mul(axis, hi));
    out.radius = r;
    return out;
}

/* corners of every triangle under a BLAS node, depth first in child order
   (sofren2.c:3151-3196) */
void gather_node_corners(const sft_model* m, u32 node, std::vector<sftvec>& out) {
    const sft_bvh_node& nd = m->blas_nodes[node];
    for (u32 c = 0; c < nd.child_count; c += 1) {
        const u32 ch = nd.child_indices[c];
        if (kInvalid == ch) continue;
        if (!nd.is_leaf_group) {
            gather_node_corners(m, ch, out);
            continue;
        }
        const u32 payload = ch & ~0x80000000u;
        const u32 gi = payload >> SFT_TRIANGLE_BITS, t = payload & ((1u << SFT_TRIANGLE_BITS) - 1);
        const sft_mesh* mesh = &m->meshes[m->meshlet_mesh_inds[gi]];
        const u32 local = gi - mesh->model_meshlet_offset;
        const sft_meshlet& mv = mesh->meshlets[local];
        const u8* in = mesh->inds[local].inds;
        const u32 base = (t / 8) * 32 + (t % 8);
        for (int k = 0; k < 3; k += 1) {
            const u8 i = in[base + 8 * k];
            sftvec p;
            p.x = mv.x[i];
            p.y = mv.y[i];
            p.z = mv.z[i];
            out.push_back(p);
        }
    }
}

/* static models: split the largest-volume internal node among the chosen
   subtrees (its first child in place, the others appended) until the budget
   is reached or only leaf groups remain, then fit one unbound capsule per
   subtree (sofren2.c:3198-3265) */
void bvh_proxies(sft_model* m, sft_arena* arena, i32 budget) {
    std::vector<u32> chosen((size_t)budget);
    i32 n = 1;
    chosen[0] = m->blas_root_ind;
    while (n < budget) {
        i32 pick = -1;
        f32 pick_vol = -1.f;
        for (i32 i = 0; i < n; i += 1) {
            const sft_bvh_node& nd = m->blas_nodes[chosen[i]];
            if (nd.is_leaf_group) continue;
            f32 lo[3] = { 1e30f, 1e30f, 1e30f }, hi[3] = { -1e30f, -1e30f, -1e30f };
            for (u32 c = 0; c < nd.child_count; c += 1) {
                const f32 cl[3] = { nd.min_x[c], nd.min_y[c], nd.min_z[c] }, chh[3] = { nd.max_x[c], nd.max_y[c], nd.max_z[c] };
                for (int k = 0; k < 3; k += 1) {
                    if (cl[k] < lo[k]) lo[k] = cl[k];
                    if (chh[k] > hi[k]) hi[k] = chh[k];
                }
            }
            const f32 vol = (hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]);
            if (vol > pick_vol) {
                pick_vol = vol;
                pick = i;
            }
        }
        if (pick < 0) break;
        const sft_bvh_node& sp = m->blas_nodes[chosen[pick]];
        chosen[pick] = sp.child_indices[0];
        for (u32 c = 1; c < sp.child_count && n < budget; c += 1) {
            if (kInvalid != sp.child_indices[c]) chosen[n++] = sp.child_indices[c];
        }
    }
    m->shadow_proxies = (sft_shadow_proxy*)sft_arena_alloc(arena, (u64)n * sizeof(sft_shadow_proxy));
    m->shadow_proxy_count = (u32)n;
    std::vector<sftvec> corners;
    for (i32 i = 0; i < n; i += 1) {
        corners.clear();
        gather_node_corners(m, chosen[i], corners);
        m->shadow_proxies[i] = fit_capsule(corners.data(), (i32)corners.size());
        m->shadow_proxies[i].bone_ind = kInvalid;
    }
}

u16 cluster_root(const std::vector<u16>& map, u16 bone) {
    u16 b = bone;
    while (map[b] != b) b = map[b];
    return b;
}

struct Cluster {
    f32 area;
    sftvec centroid_sum;
    i32 tri_count;
};

/* the joint with the largest weight among four; ties keep the earlier one */
template <class W>
void heaviest(const sft_meshlet_skin& s, u32 v, W* best_w, u16* best_j) {
    const u8 w[4] = { s.w0[v], s.w1[v], s.w2[v], s.w3[v] };
    const u16 j[4] = { s.j0[v], s.j1[v], s.j2[v], s.j3[v] };
    for (int k = 0; k < 4; k += 1) {
        if ((W)w[k] > *best_w) {
            *best_w = (W)w[k];
            *best_j = j[k];
        }
    }
}

/* skinned models: triangles go to their dominant joint's cluster (area and
   centre sums); the smallest cluster merges into its parent's (or, for a
   root, the largest) until the budget holds; each cluster's capsule runs
   from its joint's bind position toward the centre of its triangles,
   radius min(2.2 x mean, 0.9 x max) over the vertices whose heaviest joint
   falls in it (sofren2.c:3267-3509) */
void bone_proxies(sft_model* m, sft_arena* arena, i32 budget) {
    const i32 nb = m->skeleton.bone_count;
    std::vector<u16> map((size_t)nb);
    std::vector<Cluster> cl((size_t)nb);
    memset(cl.data(), 0, cl.size() * sizeof(Cluster));
    for (i32 i = 0; i < nb; i += 1) map[i] = (u16)i;

    for (u32 mi = 0; mi < m->mesh_count; mi += 1) {
        const sft_mesh* mesh = &m->meshes[mi];
        if (!mesh->meshlet_skins) continue;
        for (u32 ml = 0; ml < mesh->meshlet_count; ml += 1) {
            const sft_meshlet_skin& sk = mesh->meshlet_skins[ml];
            const sft_meshlet& mv = mesh->meshlets[ml];
            const u8* in = mesh->inds[ml].inds;
            for (u32 t = 0; t < kTris; t += 1) {
                const u32 base = (t / 8) * 32 + (t % 8);
                const u8 c3[3] = { in[base], in[base + 8], in[base + 16] };
                if (c3[0] == c3[1] && c3[1] == c3[2]) continue;
                f32 best_w = -1.f;
                u16 best_j = 0;
                for (int k = 0; k < 3; k += 1) heaviest(sk, c3[k], &best_w, &best_j);
                if ((i32)best_j >= nb) continue;
                sftvec p[3];
                for (int k = 0; k < 3; k += 1) {
                    p[k].x = mv.x[c3[k]];
                    p[k].y = mv.y[c3[k]];
                    p[k].z = mv.z[c3[k]];
                }
                const f32 area = sft_v_length(sft_v_cross(sft_v_sub(p[1], p[0]), sft_v_sub(p[2], p[0]))) * 0.5f;
                const sftvec centre = sft_v_mul(sft_v_add(p[0], sft_v_add(p[1], p[2])), 0.333333f);
                cl[best_j].area += area;
                cl[best_j].centroid_sum = sft_v_add(cl[best_j].centroid_sum, centre);
                cl[best_j].tri_count += 1;
            }
        }
    }

    i32 active = 0;
    for (i32 i = 0; i < nb; i += 1) active += cl[i].tri_count > 0;
    while (active > budget && active > 1) {
        f32 min_area = 1e30f;
        i32 small = -1;
        for (i32 i = 0; i < nb; i += 1) {
            if (cl[i].tri_count > 0 && cl[i].area < min_area) {
                min_area = cl[i].area;
                small = i;
            }
        }
        if (small < 0) break;
        i32 into = -1;
        const i32 parent = m->skeleton.bones[small].parent_ind;
        if (parent != -1) {
            into = cluster_root(map, (u16)parent);
        }
        else {
            /* a root merges into the largest other cluster; emptied clusters
               keep their area and can still be chosen, as in the reference */
            f32 max_area = -1.f;
            for (i32 i = 0; i < nb; i += 1) {
                if (i != small && cl[i].area > max_area) {
                    max_area = cl[i].area;
                    into = i;
                }
            }
        }
        if (into < 0 || into == small) break;
        map[small] = (u16)into;
        cl[into].area += cl[small].area;
        cl[into].centroid_sum = sft_v_add(cl[into].centroid_sum, cl[small].centroid_sum);
        cl[into].tri_count += cl[small].tri_count;
        cl[small].tri_count = 0;
        active -= 1;
    }

    m->shadow_proxies = (sft_shadow_proxy*)sft_arena_alloc(arena, (u64)active * sizeof(sft_shadow_proxy));
    m->shadow_proxy_count = (u32)active;

    std::vector<std::vector<sftvec>> verts((size_t)nb);
    for (u32 mi = 0; mi < m->mesh_count; mi += 1) {
        const sft_mesh* mesh = &m->meshes[mi];
        if (!mesh->meshlet_skins) continue;
        for (u32 ml = 0; ml < mesh->meshlet_count; ml += 1) {
            const sft_meshlet_skin& sk = mesh->meshlet_skins[ml];
            const sft_meshlet& mv = mesh->meshlets[ml];
            for (u32 v = 0; v < kVerts; v += 1) {
                u8 best_w = sk.w0[v];
                u16 best_j = sk.j0[v];
                heaviest(sk, v, &best_w, &best_j);
                if (best_w > 0 && (i32)best_j < nb) {
                    sftvec p;
                    p.x = mv.x[v];
                    p.y = mv.y[v];
                    p.z = mv.z[v];
                    verts[cluster_root(map, best_j)].push_back(p);
                }
            }
        }
    }

    i32 out = 0;
    for (i32 i = 0; i < nb; i += 1) {
        const std::vector<sftvec>& vs = verts[i];
        if (cl[i].tri_count <= 0 || vs.empty()) continue;
        const sftvec centroid = sft_v_mul(cl[i].centroid_sum, 1.f / cl[i].tri_count);
        const sftvec anchor = m->skeleton.bones[i].bind_pos;
        sftvec axis = sft_v_sub(centroid, anchor);
        const f32 axis_len = sft_v_length(axis);
        if (axis_len > SFT_EPSILON) {
            axis = sft_v_mul(axis, 1.f / axis_len);
        }
        else {
            axis.x = 0.f;
            axis.y = 1.f;
            axis.z = 0.f;
        }
        f32 t0 = 1e9f, t1 = -1e9f, sum_r = 0.f, max_r = 0.f;
        const i32 n = (i32)vs.size();
        for (i32 v = 0; v < n; v += 1) {
            const f32 t = sft_v_dot(sft_v_sub(vs[v], anchor), axis);
            if (t < t0) t0 = t;
            if (t > t1) t1 = t;
            const f32 r = sqrtf(sft_v_length2(sft_v_sub(vs[v], sft_v_add(anchor, sft_v_mul(axis, t)))));
            sum_r += r;
            if (r > max_r) max_r = r;
        }
        const f32 r = SFT_MIN((sum_r / n) * 2.2f, max_r * 0.9f);
        f32 lo = SFT_MIN(0.f, t0 - (r * 0.25f)), hi = SFT_MAX(0.f, t1 + (r * 0.25f));
        if (lo > hi) lo = hi = (lo + hi) * 0.5f;
        sft_shadow_proxy& p = m->shadow_proxies[out++];
        p.posa = sft_v_add(anchor, sft_v_mul(axis, lo));
        p.posb = sft_v_add(anchor, sft_v_mul(axis, hi));
        p.radius = r;
        p.bone_ind = (u32)i;
    }
}

}  // namespace

/* Capsule shadow proxies of a model with a BLAS: per bone cluster when it
   has a skeleton, per BVH subtree otherwise; none without a budget or a
   BLAS (sofren2.c:3511-3524). */
extern "C" void sftx_shadow_proxies_build(sft_model* model, sft_arena* arena, i32 budget) {
    if (!model) return;
    if (budget <= 0 || 0 == model->blas_node_count || !arena) {
        model->shadow_proxies = NULL;
        model->shadow_proxy_count = 0;
        return;
    }
    if (model->skeleton.bone_count > 0) bone_proxies(model, arena, budget);
    else bvh_proxies(model, arena, budget);
}
