// LBVH traversal on the device and the host-side builder interface.
//
// Replaces the RTX traversals of nolmoonen/gpuupbp (optixTrace in
// gpuupbp/kernel/intersector.cuh:120,214 and gpuupbp/kernel/pde/functs_*.cuh) by
// stack traversal of a binary LBVH whose inner nodes carry both child boxes.
// Leaf boxes are inflated at build time so that box tests are conservative with
// respect to the exact FP32 acceptance predicates evaluated at the leaves.
#pragma once

#include "upbp_device.h"

namespace upbp {

constexpr int kBvhStack = 64;

struct RayBox {
    f3 o, inv;
    float tmin, tmax;
};

UPBP_HD RayBox make_raybox(f3 o, f3 d, float tmin, float tmax)
{
    return {o, mk(1.f / d.x, 1.f / d.y, 1.f / d.z), tmin, tmax};
}

// Slab test; NaNs from 0*inf are ignored by fminf/fmaxf, which keeps the test conservative.
UPBP_HD bool slab(const RayBox& r, float lx, float ly, float lz, float ux, float uy, float uz, float tmax,
                  float& tentry)
{
    const float tx0 = (lx - r.o.x) * r.inv.x, tx1 = (ux - r.o.x) * r.inv.x;
    const float ty0 = (ly - r.o.y) * r.inv.y, ty1 = (uy - r.o.y) * r.inv.y;
    const float tz0 = (lz - r.o.z) * r.inv.z, tz1 = (uz - r.o.z) * r.inv.z;
    const float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), r.tmin));
    const float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tmax));
    tentry = tn;
    return tn <= tf;
}

UPBP_HD bool inside(f3 p, float lx, float ly, float lz, float ux, float uy, float uz)
{
    return p.x >= lx && p.y >= ly && p.z >= lz && p.x <= ux && p.y <= uy && p.z <= uz;
}

// Visits every leaf whose box contains p. fn(sorted_leaf) -> void.
template <typename F>
__device__ __forceinline__ bool bvh_point_query(const Bvh& b, f3 p, F&& fn)
{
    if (b.n == 0) return true;
    int stack[kBvhStack];
    int sp = 0;
    int node = 0;
    for (;;) {
        const BNode nd = b.nodes[node];
        const bool h0 = inside(p, nd.lx0, nd.ly0, nd.lz0, nd.ux0, nd.uy0, nd.uz0);
        const bool h1 = inside(p, nd.lx1, nd.ly1, nd.lz1, nd.ux1, nd.uy1, nd.uz1);
        int next = -1;
        if (h0) {
            if (nd.c0 < 0) fn(uint32_t(~nd.c0));
            else next = nd.c0;
        }
        if (h1) {
            if (nd.c1 < 0) fn(uint32_t(~nd.c1));
            else if (next < 0) next = nd.c1;
            else {
                if (sp == kBvhStack) return false;
                stack[sp++] = nd.c1;
            }
        }
        if (next < 0) {
            if (sp == 0) return true;
            next = stack[--sp];
        }
        node = next;
    }
}

// Visits every leaf whose box meets the segment [tmin, tmax] of the ray.
// fn(sorted_leaf, float& tmax) may shrink tmax (closest-hit queries).
template <typename F>
__device__ __forceinline__ bool bvh_ray_query(const Bvh& b, f3 o, f3 d, float tmin, float tmax, F&& fn)
{
    if (b.n == 0) return true;
    const RayBox r = make_raybox(o, d, tmin, tmax);
    float tcur = tmax;
    int stack[kBvhStack];
    int sp = 0;
    int node = 0;
    for (;;) {
        const BNode nd = b.nodes[node];
        float e0, e1;
        const bool h0 = slab(r, nd.lx0, nd.ly0, nd.lz0, nd.ux0, nd.uy0, nd.uz0, tcur, e0);
        const bool h1 = slab(r, nd.lx1, nd.ly1, nd.lz1, nd.ux1, nd.uy1, nd.uz1, tcur, e1);
        int n0 = -1, n1 = -1; // inner children to visit, nearest first
        if (h0) {
            if (nd.c0 < 0) fn(uint32_t(~nd.c0), tcur);
            else n0 = nd.c0;
        }
        if (h1) {
            if (nd.c1 < 0) fn(uint32_t(~nd.c1), tcur);
            else n1 = nd.c1;
        }
        int next;
        if (n0 >= 0 && n1 >= 0) {
            if (e1 < e0) {
                const int t = n0;
                n0 = n1;
                n1 = t;
            }
            if (sp == kBvhStack) return false;
            stack[sp++] = n1;
            next = n0;
        } else {
            next = n0 >= 0 ? n0 : n1;
        }
        if (next < 0) {
            if (sp == 0) return true;
            next = stack[--sp];
        }
        node = next;
    }
}

// Visits every leaf of a wide LBVH whose box meets the segment [tmin, tmax] of the ray.
template <typename F>
__device__ __forceinline__ bool wbvh_ray_query(const WBvh& b, f3 o, f3 d, float tmin, float tmax, F&& fn)
{
    if (b.n == 0) return true;
    const RayBox r = make_raybox(o, d, tmin, tmax);
    int stack[kBvhStack];
    int sp = 0;
    int node = 0;
    for (;;) {
        const WNode nd = b.nodes[node];
        const float sx = wnode_step(nd.meta, 0), sy = wnode_step(nd.meta, 1), sz = wnode_step(nd.meta, 2);
        const int nc = wnode_count(nd.meta);
        int next = -1;
        for (int k = 0; k < nc; k++) {
            float e, lx, ly, lz, ux, uy, uz;
            wnode_box(nd, k, sx, sy, sz, lx, ly, lz, ux, uy, uz);
            if (!slab(r, lx, ly, lz, ux, uy, uz, tmax, e)) continue;
            const int ch = nd.child[k];
            if (ch < 0) {
                fn(uint32_t(~ch));
            } else if (next < 0) {
                next = ch;
            } else {
                if (sp == kBvhStack) return false;
                stack[sp++] = ch;
            }
        }
        if (next < 0) {
            if (sp == 0) return true;
            next = stack[--sp];
        }
        node = next;
    }
}

// Moller-Trumbore (Moller & Trumbore 1997); replaces OptiX's built-in triangle test.
UPBP_HD bool tri_test(float4 p0, float4 e1, float4 e2, f3 o, f3 d, float tmin, float tmax, float& th, float& bu,
                      float& bv)
{
    const f3 E1 = mk(e1.x, e1.y, e1.z), E2 = mk(e2.x, e2.y, e2.z);
    const f3 pv = cross(d, E2);
    const float det = dot(E1, pv);
    if (det == 0.f) return false;
    const float inv = 1.f / det;
    const f3 tv = o - mk(p0.x, p0.y, p0.z);
    const float u = dot(tv, pv) * inv;
    if (u < 0.f || u > 1.f) return false;
    const f3 qv = cross(tv, E1);
    const float v = dot(d, qv) * inv;
    if (v < 0.f || u + v > 1.f) return false;
    const float t = dot(E2, qv) * inv;
    if (!(t > tmin && t < tmax)) return false;
    th = t;
    bu = u;
    bv = v;
    return true;
}

// Analytic sphere (centre s.xyz, radius s.w) against a ray: the roots of |o + t d - c| = r with
// the near root refined by one more solve from the first estimate when the ray starts more
// than a radius away, which removes the cancellation of b^2 - c for distant spheres
// (gpuupbp/kernel/intersector.cu:35-95). Writes the roots in (tmin, tmax), near first, and the
// outward normals at them; returns how many (0..2). A closest-hit caller takes the first, an
// all-hits caller every one.
UPBP_HD int sphere_test(float4 s, f3 ro, f3 rd, float tmin, float tmax, float* th, f3* nh)
{
#pragma clang fp contract(off)
    const float radius = s.w;
    const f3 o = ro - mk(s.x, s.y, s.z);
    const float l = 1.f / sqrtf(dot(rd, rd));
    const f3 d = rd * l;
    float b = dot(o, d);
    float c = dot(o, o) - radius * radius;
    float disc = b * b - c;
    if (!(disc > 0.f)) return 0;
    float sdisc = sqrtf(disc);
    const float root1 = -b - sdisc;
    float root11 = 0.f;
    const bool refine = fabsf(root1) > radius;
    if (refine) {
        const f3 o1 = o + d * root1;
        b = dot(o1, d);
        c = dot(o1, o1) - radius * radius;
        disc = b * b - c;
        if (disc > 0.f) {
            sdisc = sqrtf(disc);
            root11 = -b - sdisc;
        }
    }
    int n = 0;
    const float near = root1 + root11;
    float t = near * l;
    if (t > tmin && t < tmax) {
        th[n] = t;
        nh[n++] = (o + d * near) / radius;
    }
    const float root2 = (-b + sdisc) + (refine ? root1 : 0.f);
    t = root2 * l;
    if (t > tmin && t < tmax) {
        th[n] = t;
        nh[n++] = (o + d * root2) / radius;
    }
    return n;
}

} // namespace upbp

// ---- host-side builder ---------------------------------------------------------
namespace upbp_host {

// Device storage of one LBVH: max(n-1,1) inner nodes plus the sorted order.
struct BvhStore {
    upbp::BNode* nodes = nullptr;
    uint32_t* order = nullptr;
    uint32_t n = 0, cap = 0;
};

// Scratch shared by all builds of one renderer (grow-only, one capacity per buffer).
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    int ensure(size_t need)
    {
        if (need <= bytes && p) return 0;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        const size_t b = need < 4096 ? 4096 : need + need / 4;
        if (hipMalloc(&p, b) != hipSuccess) return 1;
        bytes = b;
        return 0;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <typename T>
    T* as() const
    {
        return static_cast<T*>(p);
    }
};

struct BvhScratch {
    DevBuf keys, keys2, vals2, hist, scan, lvl, meta, bounds;
};

// Device storage of a wide LBVH collapsed from a binary one (at most n-1 nodes) and the
// wide-node -> binary-node map used while building it.
struct WBvhStore {
    upbp::WNode* nodes = nullptr;
    uint32_t* bin = nullptr;
    uint32_t n = 0, cap = 0;
};

// Primitive boxes are given as lo/hi float4 arrays (w unused), already inflated.
// Builds into st on stream s; returns 0 or a hipError_t value. With n_dev, the primitive count
// is read from device memory when the kernels run (at most n, which then sizes the buffers and
// launches): no host synchronisation, and zero or one primitive are handled on the device.
int lbvh_build(BvhStore& st, BvhScratch& sc, const float4* lo, const float4* hi, uint32_t n, hipStream_t s,
               const uint32_t* n_dev = nullptr);
void lbvh_free(BvhStore& st);
// Collapses the refitted binary LBVH `bin` (n leaves) into a kWide-wide one; same leaf order.
int wbvh_build(WBvhStore& w, const BvhStore& bin, BvhScratch& sc, uint32_t n, hipStream_t s,
               const uint32_t* n_dev = nullptr);
void wbvh_free(WBvhStore& w);
void lbvh_scratch_free(BvhScratch& sc);

// Stable LSD radix sort of (key, value) pairs on the low `bits` bits; keys/vals updated in place.
int radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t n, int bits, BvhScratch& sc, hipStream_t s,
                     const uint32_t* n_dev = nullptr);
// Exclusive prefix sum of n u32 values; in and out may alias. Returns the total through *total_dev
// (device pointer, may be null).
int scan_exclusive(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* total_dev, BvhScratch& sc,
                   hipStream_t s);

} // namespace upbp_host
