// Test hooks: LBVH range queries and closest-hit queries on caller-provided
// primitives, returning exact index sets for comparison with brute force.
// The acceptance predicates are those of nolmoonen/gpuupbp gpuupbp/kernel/intersection.cuh:33-173,
// applied as in gpuupbp/kernel/pde/program_{surf,pp3d,pb2d,bb1d,bp2d}.cu.
#include "env_tables.h"
#include "kernels.h"

#include "../../../include/upbp_capi.h"

#include <algorithm>
#include <vector>

using namespace upbp;
using namespace upbp_host;

namespace {

__global__ void k_point_boxes(const float* prims, uint32_t n, float r, float4* lo, float4* hi)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* q = prims + 3 * size_t(i);
    const float mag = fmaxf(fabsf(q[0]), fmaxf(fabsf(q[1]), fabsf(q[2])));
    const float rr = r * (1.f + 1e-5f) + 1e-6f * mag;
    lo[i] = make_float4(q[0] - rr, q[1] - rr, q[2] - rr, 0.f);
    hi[i] = make_float4(q[0] + rr, q[1] + rr, q[2] + rr, 0.f);
}

// pass 0 counts accepted primitives per query, pass 1 writes them at offsets[q]
// Beams (kinds 2, 3) enter the LBVH as sub-beams: leaf -> (beam, owned parameter range), and a
// pair counts only for the sub-beam whose range holds the predicate's beam parameter.
__global__ void k_range(int kind, int pass, Bvh bvh, WBvh wide, const float* prims, const uint32_t* sub_prim,
                        const float2* sub_rng, const float* queries, uint32_t nq, float r2, uint32_t* counts,
                        const uint32_t* offsets, uint32_t* out, uint32_t* stack_fail)
{
    const uint32_t qi = blockIdx.x * blockDim.x + threadIdx.x;
    if (qi >= nq) return;
    uint32_t c = 0;
    const uint32_t base = pass ? offsets[qi] : 0u;
    auto accept = [&](uint32_t prim) {
        if (pass) out[base + c] = prim;
        c++;
    };
    bool ok = true;
    if (kind == 0 || kind == 2) {
        const float* q = queries + 3 * size_t(qi);
        const f3 P = mk(q[0], q[1], q[2]);
        ok = bvh_point_query(bvh, P, [&](uint32_t k) {
            if (kind == 0) {
                const uint32_t prim = bvh.order[k];
                const float* x = prims + 3 * size_t(prim);
                if (pred_point_point(P, mk(x[0], x[1], x[2]), r2)) accept(prim);
            } else {
                const uint32_t prim = sub_prim[bvh.order[k]];
                const float2 rg = sub_rng[bvh.order[k]];
                const float* x = prims + 7 * size_t(prim);
                const f3 bo = mk(x[0], x[1], x[2]), bd = mk(x[3], x[4], x[5]);
                float t, h;
                if (pred_point_beam(bo, bd, x[6], P, r2) && pred_bre(t, h, bo, bd, 0.f, x[6], P, r2) && t >= rg.x &&
                    t < rg.y)
                    accept(prim);
            }
        });
    } else {
        const float* q = queries + 8 * size_t(qi);
        const f3 O = mk(q[0], q[1], q[2]), D = mk(q[3], q[4], q[5]);
        const float t0 = q[6], t1 = q[7];
        auto leaf = [&](uint32_t k) {
            if (kind == 1) {
                const uint32_t prim = bvh.order[k];
                const float* x = prims + 3 * size_t(prim);
                const f3 X = mk(x[0], x[1], x[2]);
                float d, h;
                if (pred_point_beam(O, D, t1, X, r2) && pred_bre(d, h, O, D, t0, t1, X, r2)) accept(prim);
            } else {
                const uint32_t prim = sub_prim[bvh.order[k]];
                const float2 rg = sub_rng[bvh.order[k]];
                const float* x = prims + 7 * size_t(prim);
                float dd, st, ta, tb;
                if (pred_beam_beam(O, D, t0, t1, mk(x[0], x[1], x[2]), mk(x[3], x[4], x[5]), 0.f, x[6], r2, dd, st, ta,
                                   tb) &&
                    tb >= rg.x && tb < rg.y)
                    accept(prim);
            }
        };
        if (wide.nodes) ok = wbvh_ray_query(wide, O, D, t0, t1, leaf);
        else ok = bvh_ray_query(bvh, O, D, t0, t1, [&](uint32_t k, float&) { leaf(k); });
    }
    if (!ok) atomicAdd(stack_fail, 1u);
    if (!pass) counts[qi] = c;
}

__global__ void k_closest(TriSet T, const float* rays, uint32_t n, float* ht, int32_t* hp, uint32_t* stack_fail)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* R = rays + 8 * size_t(i);
    const f3 o = mk(R[0], R[1], R[2]), d = mk(R[3], R[4], R[5]);
    const float tmin = R[6];
    int best = -1;
    float bt = INFINITY;
    const bool ok = bvh_ray_query(T.bvh, o, d, tmin, R[7], [&](uint32_t k, float& thi) {
        float t, u, v;
        if (tri_test(T.tri[3 * k], T.tri[3 * k + 1], T.tri[3 * k + 2], o, d, tmin, thi, t, u, v)) {
            best = int(T.bvh.order[k]);
            bt = t;
            thi = t;
        }
    });
    if (!ok) atomicAdd(stack_fail, 1u);
    ht[i] = bt;
    hp[i] = best;
}

__global__ void k_sort_tris(const uint32_t* order, const float4* tri, uint32_t n, float4* out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int k = 0; k < 3; k++) out[3 * i + k] = tri[3 * order[i] + k];
}

struct Tmp {
    std::vector<void*> ptrs;
    ~Tmp()
    {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <typename T>
    T* alloc(size_t n)
    {
        void* p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return nullptr;
        ptrs.push_back(p);
        return static_cast<T*>(p);
    }
};

inline uint32_t nb(size_t n) { return uint32_t((n + 255) / 256); }

// plain streaming copy: the reference kernel for HBM counter calibration and bandwidth
__global__ void __launch_bounds__(256) k_copy_calibration(const float4* __restrict__ in, float4* __restrict__ out,
                                                          size_t n)
{
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = in[i];
}

} // namespace

extern "C" int upbp_test_range_query(int kind, const float* prims, uint32_t n_prims, const float* queries,
                                     uint32_t n_queries, float radius, float subbeam_len, uint32_t* offsets,
                                     uint32_t* indices, uint64_t max_indices, uint64_t* n_indices)
{
    return upbp_test_range_query_wide(kind, prims, n_prims, queries, n_queries, radius, subbeam_len, 0, offsets,
                                      indices, max_indices, n_indices);
}

extern "C" int upbp_test_range_query_wide(int kind, const float* prims, uint32_t n_prims, const float* queries,
                                          uint32_t n_queries, float radius, float subbeam_len, int wide,
                                          uint32_t* offsets, uint32_t* indices, uint64_t max_indices,
                                          uint64_t* n_indices)
{
    if (kind < 0 || kind > 3 || !offsets || (n_prims && !prims) || (n_queries && !queries)) return UPBP_E_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return UPBP_E_NODEVICE;
    const size_t pw = (kind <= 1) ? 3 : 7, qw = (kind == 0 || kind == 2) ? 3 : 8;
    // beams: sub-beam boxes, owning beam and parameter range, built on the host
    std::vector<float4> blo, bhi;
    std::vector<uint32_t> sprim;
    std::vector<float2> srng;
    if (kind >= 2) {
        float mag = 0.f;
        for (size_t i = 0; i < n_prims * pw; i++) mag = fmaxf(mag, fabsf(prims[i]));
        for (size_t i = 0; i < n_queries * qw; i++) mag = fmaxf(mag, fabsf(queries[i]));
        const float slack = 2e-4f * mag;
        for (uint32_t i = 0; i < n_prims; i++) {
            const float* x = prims + 7 * size_t(i);
            const f3 o = mk(x[0], x[1], x[2]), d = mk(x[3], x[4], x[5]);
            const uint32_t n = subbeam_count(x[6], subbeam_len);
            for (uint32_t j = 0; j < n; j++) {
                float lo, hi, ta, tb;
                float4 a, b;
                subbeam_range(x[6], j, n, lo, hi, ta, tb);
                subbeam_box(o, d, x[6], ta, tb, radius, slack, a, b);
                blo.push_back(a);
                bhi.push_back(b);
                sprim.push_back(i);
                srng.push_back(make_float2(lo, hi));
            }
        }
    }
    const uint32_t n_leaves = kind >= 2 ? uint32_t(sprim.size()) : n_prims;
    Tmp t;
    float* dp = t.alloc<float>(n_prims * pw);
    float* dq = t.alloc<float>(n_queries * qw);
    float4* lo = t.alloc<float4>(n_leaves);
    float4* hi = t.alloc<float4>(n_leaves);
    uint32_t* dsp = t.alloc<uint32_t>(n_leaves);
    float2* dsr = t.alloc<float2>(n_leaves);
    uint32_t* cnt = t.alloc<uint32_t>(n_queries + 1);
    uint32_t* off = t.alloc<uint32_t>(n_queries + 1);
    uint32_t* fail_count = t.alloc<uint32_t>(1);
    if (!dp || !dq || !lo || !hi || !dsp || !dsr || !cnt || !off || !fail_count) return UPBP_E_HIP;
    hipStream_t s = nullptr;
    if (n_prims) (void)hipMemcpy(dp, prims, n_prims * pw * 4, hipMemcpyHostToDevice);
    if (n_queries) (void)hipMemcpy(dq, queries, n_queries * qw * 4, hipMemcpyHostToDevice);
    if (kind >= 2 && n_leaves) {
        (void)hipMemcpy(lo, blo.data(), n_leaves * 16, hipMemcpyHostToDevice);
        (void)hipMemcpy(hi, bhi.data(), n_leaves * 16, hipMemcpyHostToDevice);
        (void)hipMemcpy(dsp, sprim.data(), n_leaves * 4, hipMemcpyHostToDevice);
        (void)hipMemcpy(dsr, srng.data(), n_leaves * 8, hipMemcpyHostToDevice);
    }
    (void)hipMemset(fail_count, 0, 4);
    BvhStore st;
    WBvhStore wst;
    BvhScratch sc;
    if (kind <= 1 && n_prims)
        hipLaunchKernelGGL(k_point_boxes, dim3(nb(n_prims)), dim3(256), 0, s, dp, n_prims, radius, lo, hi);
    const bool use_wide = wide && (kind == 1 || kind == 3);
    if (lbvh_build(st, sc, lo, hi, n_leaves, s) || (use_wide && wbvh_build(wst, st, sc, n_leaves, s))) {
        lbvh_free(st);
        wbvh_free(wst);
        lbvh_scratch_free(sc);
        return UPBP_E_HIP;
    }
    const Bvh bvh{st.nodes, st.order, n_leaves};
    const WBvh wbvh{use_wide ? wst.nodes : nullptr, n_leaves};
    const float r2 = radius * radius;
    if (n_queries)
        hipLaunchKernelGGL(k_range, dim3(nb(n_queries)), dim3(256), 0, s, kind, 0, bvh, wbvh, dp, dsp, dsr, dq, n_queries,
                           r2, cnt, off, nullptr, fail_count);
    uint32_t* total = off + n_queries;
    int rc = scan_exclusive(cnt, off, n_queries, total, sc, s);
    uint32_t host_total = 0, fails = 0;
    if (!rc) rc = int(hipMemcpy(&host_total, total, 4, hipMemcpyDeviceToHost));
    uint32_t* dout = rc ? nullptr : t.alloc<uint32_t>(host_total);
    if (!rc && n_queries && dout)
        hipLaunchKernelGGL(k_range, dim3(nb(n_queries)), dim3(256), 0, s, kind, 1, bvh, wbvh, dp, dsp, dsr, dq, n_queries,
                           r2, cnt, off, dout, fail_count);
    if (!rc) rc = int(hipDeviceSynchronize());
    if (!rc) rc = int(hipMemcpy(offsets, off, (size_t(n_queries) + 1) * 4, hipMemcpyDeviceToHost));
    if (!rc && indices && host_total)
        rc = int(hipMemcpy(indices, dout, std::min<uint64_t>(host_total, max_indices) * 4, hipMemcpyDeviceToHost));
    if (!rc) rc = int(hipMemcpy(&fails, fail_count, 4, hipMemcpyDeviceToHost));
    lbvh_free(st);
    wbvh_free(wst);
    lbvh_scratch_free(sc);
    if (n_queries == 0) offsets[0] = 0;
    if (n_indices) *n_indices = host_total;
    if (rc) return UPBP_E_HIP;
    if (fails) return UPBP_E_CAPACITY;
    return host_total <= max_indices ? UPBP_OK : UPBP_E_CAPACITY;
}

extern "C" int upbp_test_closest_hit(const float* verts, uint32_t n_tris, const float* rays, uint32_t n_rays,
                                     float* hit_t, int32_t* hit_prim)
{
    if ((n_tris && !verts) || (n_rays && (!rays || !hit_t || !hit_prim))) return UPBP_E_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return UPBP_E_NODEVICE;
    std::vector<float4> tri(3 * size_t(n_tris)), lo(n_tris), hi(n_tris);
    for (uint32_t i = 0; i < n_tris; i++) {
        const float* v = verts + 9 * size_t(i);
        tri[3 * i] = make_float4(v[0], v[1], v[2], 0.f);
        tri[3 * i + 1] = make_float4(v[3] - v[0], v[4] - v[1], v[5] - v[2], 0.f);
        tri[3 * i + 2] = make_float4(v[6] - v[0], v[7] - v[1], v[8] - v[2], 0.f);
        float mn[3], mx[3], mag = 0.f;
        for (int a = 0; a < 3; a++) {
            mn[a] = fminf(v[a], fminf(v[3 + a], v[6 + a]));
            mx[a] = fmaxf(v[a], fmaxf(v[3 + a], v[6 + a]));
            mag = fmaxf(mag, fmaxf(fabsf(mn[a]), fabsf(mx[a])));
        }
        const float ext = fmaxf(mx[0] - mn[0], fmaxf(mx[1] - mn[1], mx[2] - mn[2]));
        const float pad = 1e-5f * ext + 1e-6f * mag + 1e-7f;
        lo[i] = make_float4(mn[0] - pad, mn[1] - pad, mn[2] - pad, 0.f);
        hi[i] = make_float4(mx[0] + pad, mx[1] + pad, mx[2] + pad, 0.f);
    }
    Tmp t;
    float4* dtri = t.alloc<float4>(tri.size());
    float4* dsorted = t.alloc<float4>(tri.size());
    float4* dlo = t.alloc<float4>(n_tris);
    float4* dhi = t.alloc<float4>(n_tris);
    float* drays = t.alloc<float>(8 * size_t(n_rays));
    float* dt = t.alloc<float>(n_rays);
    int32_t* dp = t.alloc<int32_t>(n_rays);
    uint32_t* fail_count = t.alloc<uint32_t>(1);
    if (!dtri || !dsorted || !dlo || !dhi || !drays || !dt || !dp || !fail_count) return UPBP_E_HIP;
    (void)hipMemcpy(dtri, tri.data(), tri.size() * 16, hipMemcpyHostToDevice);
    (void)hipMemcpy(dlo, lo.data(), lo.size() * 16, hipMemcpyHostToDevice);
    (void)hipMemcpy(dhi, hi.data(), hi.size() * 16, hipMemcpyHostToDevice);
    (void)hipMemcpy(drays, rays, 8 * size_t(n_rays) * 4, hipMemcpyHostToDevice);
    (void)hipMemset(fail_count, 0, 4);
    BvhStore st;
    BvhScratch sc;
    if (lbvh_build(st, sc, dlo, dhi, n_tris, nullptr)) {
        lbvh_free(st);
        lbvh_scratch_free(sc);
        return UPBP_E_HIP;
    }
    if (n_tris) hipLaunchKernelGGL(k_sort_tris, dim3(nb(n_tris)), dim3(256), 0, 0, st.order, dtri, n_tris, dsorted);
    TriSet T;
    T.tri = dsorted;
    T.nrm = nullptr;
    T.rec = nullptr;
    T.bvh = {st.nodes, st.order, n_tris};
    if (n_rays) hipLaunchKernelGGL(k_closest, dim3(nb(n_rays)), dim3(256), 0, 0, T, drays, n_rays, dt, dp, fail_count);
    int rc = int(hipDeviceSynchronize());
    uint32_t fails = 0;
    if (!rc) rc = int(hipMemcpy(hit_t, dt, size_t(n_rays) * 4, hipMemcpyDeviceToHost));
    if (!rc) rc = int(hipMemcpy(hit_prim, dp, size_t(n_rays) * 4, hipMemcpyDeviceToHost));
    if (!rc) rc = int(hipMemcpy(&fails, fail_count, 4, hipMemcpyDeviceToHost));
    lbvh_free(st);
    lbvh_scratch_free(sc);
    if (rc) return UPBP_E_HIP;
    return fails ? UPBP_E_CAPACITY : UPBP_OK;
}

// one thread per (ray, sphere): root count and the roots (INFINITY where absent)
__global__ void k_sphere_pairs(const float4* sph, uint32_t n_sph, const float* rays, uint32_t n_rays, float* out)
{
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= uint64_t(n_sph) * n_rays) return;
    const uint32_t r = uint32_t(i / n_sph), k = uint32_t(i % n_sph);
    const float* R = rays + 8 * size_t(r);
    float th[2] = {INFINITY, INFINITY};
    f3 nh[2];
    const int n = sphere_test(sph[k], mk(R[0], R[1], R[2]), mk(R[3], R[4], R[5]), R[6], R[7], th, nh);
    out[3 * i] = float(n);
    out[3 * i + 1] = n > 0 ? th[0] : INFINITY;
    out[3 * i + 2] = n > 1 ? th[1] : INFINITY;
}

extern "C" int upbp_test_sphere_hits(const float* spheres, uint32_t n_sph, const float* rays, uint32_t n_rays,
                                     float* out)
{
    if ((n_sph && !spheres) || (n_rays && !rays) || (uint64_t(n_sph) * n_rays && !out)) return UPBP_E_ARG;
    const uint64_t pairs = uint64_t(n_sph) * n_rays;
    if (pairs > (1ull << 30)) return UPBP_E_ARG;
    if (pairs == 0) return UPBP_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return UPBP_E_NODEVICE;
    Tmp t;
    float4* ds = t.alloc<float4>(n_sph);
    float* dr = t.alloc<float>(8 * size_t(n_rays));
    float* dout = t.alloc<float>(3 * pairs);
    if (!ds || !dr || !dout) return UPBP_E_HIP;
    (void)hipMemcpy(ds, spheres, size_t(n_sph) * 16, hipMemcpyHostToDevice);
    (void)hipMemcpy(dr, rays, 8 * size_t(n_rays) * 4, hipMemcpyHostToDevice);
    hipLaunchKernelGGL(k_sphere_pairs, dim3(uint32_t((pairs + 255) / 256)), dim3(256), 0, 0, ds, n_sph, dr, n_rays,
                       dout);
    int rc = int(hipDeviceSynchronize());
    if (!rc) rc = int(hipMemcpy(out, dout, 3 * pairs * 4, hipMemcpyDeviceToHost));
    return rc ? UPBP_E_HIP : UPBP_OK;
}

// per sample: sampled direction (3), its pdf, radiance (3), then lookup of that direction:
// pdf and radiance (3) -- 11 floats
__global__ void k_env_samples(DEnvMap m, const float* u, uint32_t n, float* out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float pdf, lpdf;
    f3 rad;
    const f3 d = env_sample(m, u[2 * i], u[2 * i + 1], pdf, rad);
    const f3 lrad = env_lookup(m, d, &lpdf);
    float* o = out + 11 * size_t(i);
    o[0] = d.x;
    o[1] = d.y;
    o[2] = d.z;
    o[3] = pdf;
    o[4] = rad.x;
    o[5] = rad.y;
    o[6] = rad.z;
    o[7] = lpdf;
    o[8] = lrad.x;
    o[9] = lrad.y;
    o[10] = lrad.z;
}

extern "C" int upbp_test_env_samples(const upbp_env_map* env, const float* u, uint32_t n, float* out)
{
    if (!env || !env->rgb || env->height == 0 || env->width < 2 * env->height || (n && (!u || !out)))
        return UPBP_E_ARG;
    if (n == 0) return UPBP_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return UPBP_E_NODEVICE;
    EnvTables t;
    build_env_tables(*env, t);
    Tmp tmp;
    auto up = [&](const std::vector<float>& v) {
        float* d = tmp.alloc<float>(v.size());
        if (d) (void)hipMemcpy(d, v.data(), v.size() * 4, hipMemcpyHostToDevice);
        return d;
    };
    DEnvMap m;
    m.img = up(t.img);
    m.cond_func = up(t.cond_func);
    m.cond_cdf = up(t.cond_cdf);
    m.cond_int = up(t.cond_int);
    m.marg_func = up(t.marg_func);
    m.marg_cdf = up(t.marg_cdf);
    m.marg_int = t.marg_int;
    m.width = int(env->width);
    m.height = int(env->height);
    m.nu = 2 * m.height;
    m.nv = m.height;
    m.norm = .5f * kInvPi * kInvPi;
    float* du = tmp.alloc<float>(2 * size_t(n));
    float* dout = tmp.alloc<float>(11 * size_t(n));
    if (!m.img || !m.cond_func || !m.cond_cdf || !m.cond_int || !m.marg_func || !m.marg_cdf || !du || !dout)
        return UPBP_E_HIP;
    (void)hipMemcpy(du, u, 2 * size_t(n) * 4, hipMemcpyHostToDevice);
    hipLaunchKernelGGL(k_env_samples, dim3((n + 255) / 256), dim3(256), 0, 0, m, du, n, dout);
    int rc = int(hipDeviceSynchronize());
    if (!rc) rc = int(hipMemcpy(out, dout, 11 * size_t(n) * 4, hipMemcpyDeviceToHost));
    return rc ? UPBP_E_HIP : UPBP_OK;
}

__global__ void k_guide_samples(const float4* rec, f3 arrive, float g, int surface, const float* u, uint32_t n,
                                float* out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const PhaseLobes h = surface ? cosine_lobes(arrive) : phase_lobes(arrive, g);
    const f3 d = vmf_prod_sample(rec, h, u[3 * i], u[3 * i + 1], u[3 * i + 2]);
    out[4 * i] = d.x;
    out[4 * i + 1] = d.y;
    out[4 * i + 2] = d.z;
    out[4 * i + 3] = vmf_prod_pdf(rec, h, d);
}

extern "C" int upbp_test_guide_samples(const float* lobes, const float* arrive, float g, int32_t surface,
                                       const float* u, uint32_t n, float* out)
{
    if (!lobes || !arrive || (n && (!u || !out))) return UPBP_E_ARG;
    if (n == 0) return UPBP_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return UPBP_E_NODEVICE;
    float4 rec[kGuideRec];
    pack_guide_record(lobes, rec);
    Tmp t;
    float4* drec = t.alloc<float4>(kGuideRec);
    float* du = t.alloc<float>(3 * size_t(n));
    float* dout = t.alloc<float>(4 * size_t(n));
    if (!drec || !du || !dout) return UPBP_E_HIP;
    (void)hipMemcpy(drec, rec, sizeof(rec), hipMemcpyHostToDevice);
    (void)hipMemcpy(du, u, 3 * size_t(n) * 4, hipMemcpyHostToDevice);
    hipLaunchKernelGGL(k_guide_samples, dim3((n + 255) / 256), dim3(256), 0, 0, drec, mk(arrive[0], arrive[1], arrive[2]),
                       g, int(surface), du, n, dout);
    int rc = int(hipDeviceSynchronize());
    if (!rc) rc = int(hipMemcpy(out, dout, 4 * size_t(n) * 4, hipMemcpyDeviceToHost));
    return rc ? UPBP_E_HIP : UPBP_OK;
}

extern "C" int upbp_test_copy(uint64_t bytes, int reps, double* ms_per_copy)
{
    if (bytes < 16 || reps < 1 || !ms_per_copy) return UPBP_E_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return UPBP_E_NODEVICE;
    const size_t n = size_t(bytes / 16);
    Tmp t;
    float4* a = t.alloc<float4>(n);
    float4* b = t.alloc<float4>(n);
    if (!a || !b) return UPBP_E_HIP;
    if (hipMemset(a, 0, n * 16) != hipSuccess) return UPBP_E_HIP;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return UPBP_E_HIP;
    int dev = 0, cus = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const uint32_t blocks = uint32_t(std::max(1, cus) * 8);
    hipLaunchKernelGGL(k_copy_calibration, dim3(blocks), dim3(256), 0, 0, a, b, n);
    (void)hipEventRecord(e0, 0);
    for (int k = 0; k < reps; k++) hipLaunchKernelGGL(k_copy_calibration, dim3(blocks), dim3(256), 0, 0, a, b, n);
    (void)hipEventRecord(e1, 0);
    int rc = int(hipEventSynchronize(e1));
    float ms = 0.f;
    if (!rc) rc = int(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_per_copy = double(ms) / reps;
    return rc ? UPBP_E_HIP : UPBP_OK;
}
