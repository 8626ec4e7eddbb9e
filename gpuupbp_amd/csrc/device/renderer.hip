// Host orchestration of one UPBP iteration on the GPU and the C ABI around it.
//
// upbp_create replaces Renderer::init (nolmoonen/gpuupbp gpuupbp/renderer/renderer_init.cpp:675-929):
// scene tables go to HBM, the real and imaginary triangle LBVHs are built once, and the
// wavefront buffers are sized from N_L (max path length - 1) vertex slots per light path.
// upbp_run_iteration replaces Renderer::run_iteration (gpuupbp/renderer/renderer.cpp:37-418):
// radius schedule and MIS factors (:56-119), light wavefront, deterministic photon lists +
// three photon LBVH builds, camera wavefront, FP64 accumulation on the device.
#include "env_tables.h"
#include "guide_mlp.h"
#include "kernels.h"

#include "../../../include/upbp_capi.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

using namespace upbp;
using namespace upbp_host;

namespace {

constexpr int kMaxBounces = 64;

__global__ void k_fill(uint32_t* p, uint32_t v, size_t n)
{
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// per light path: how many surface points, medium points and beams it contributes
__global__ void k_count_photons(Bufs b, uint32_t* cs, uint32_t* cm, uint32_t* cb, int want_pts, int want_beams)
{
    const IterParams& p = *b.params;
    const uint32_t pid = blockIdx.x * blockDim.x + threadIdx.x;
    if (pid >= p.n_light) return;
    uint32_t ns = 0, nm = 0;
    if (want_pts) {
        const uint32_t nv = b.nvert[pid];
        for (uint32_t k = 0; k < nv; k++) {
            const int fl = __float_as_int(b.verts[size_t(pid) * p.vstride + k].g.z);
            if (fl & VF_IN_MEDIUM) nm++;
            else ns++;
        }
    }
    uint32_t nb = 0;
    if (want_beams) {
        const uint32_t n = b.nbeam[pid];
        for (uint32_t k = 0; k < n; k++) nb += subbeam_count(b.beams[size_t(pid) * p.bstride + k].a.w, p.sub_len);
    }
    cs[pid] = ns;
    cm[pid] = nm;
    cb[pid] = nb;
}

// writes the photon lists in light-path order together with inflated primitive boxes; entries at
// or beyond a list's capacity (cap: surface, medium, sub-beams) are not written (the iteration is
// then discarded, see k_map_counts)
__global__ void k_emit_photons(Bufs b, const uint32_t* os, const uint32_t* om, const uint32_t* ob, int want_surf,
                               int want_med, int want_beams, uint32_t* ls, float4* slo, float4* shi, uint32_t* lm,
                               float4* mlo, float4* mhi, float4* geo, uint2* ref, float4* blo, float4* bhi, uint3 cap)
{
    const IterParams& p = *b.params;
    const uint32_t pid = blockIdx.x * blockDim.x + threadIdx.x;
    if (pid >= p.n_light) return;
    uint32_t js = os[pid], jm = om[pid];
    if (want_surf || want_med) {
        const uint32_t nv = b.nvert[pid];
        for (uint32_t k = 0; k < nv; k++) {
            const uint32_t slot = pid * p.vstride + k;
            const VertexRec& v = b.verts[slot];
            const int fl = __float_as_int(v.g.z);
            const float4 a = v.a;
            const float mag = fmaxf(fabsf(a.x), fmaxf(fabsf(a.y), fabsf(a.z)));
            if (fl & VF_IN_MEDIUM) {
                const float r = p.rad_point_med * (1.f + 1e-5f) + 1e-6f * mag;
                if (want_med && jm < cap.y) {
                    lm[jm] = slot;
                    mlo[jm] = make_float4(a.x - r, a.y - r, a.z - r, 0.f);
                    mhi[jm] = make_float4(a.x + r, a.y + r, a.z + r, 0.f);
                }
                jm++;
            } else {
                const float r = p.r_surf * (1.f + 1e-5f) + 1e-6f * mag;
                if (want_surf && js < cap.x) {
                    ls[js] = slot;
                    slo[js] = make_float4(a.x - r, a.y - r, a.z - r, 0.f);
                    shi[js] = make_float4(a.x + r, a.y + r, a.z + r, 0.f);
                }
                js++;
            }
        }
    }
    if (want_beams) { // sub-beams: box and (beam, owned parameter range) payload, unsorted
        const uint32_t nb = b.nbeam[pid];
        uint32_t jb = ob[pid];
        for (uint32_t k = 0; k < nb; k++) {
            const uint32_t slot = pid * p.bstride + k;
            const BeamRec& r = b.beams[slot];
            const float4 a = r.a, d = r.b, dd = r.d;
            const int flags = uint32_t(__float_as_int(dd.z)) < p.bb1d_used ? SB_BB1D : 0;
            const f3 o = mk(a.x, a.y, a.z), dir = mk(d.x, d.y, d.z);
            const uint32_t n = subbeam_count(a.w, p.sub_len);
            for (uint32_t j = 0; j < n && jb < cap.z; j++, jb++) {
                float lo, hi, ta, tb;
                subbeam_range(a.w, j, n, lo, hi, ta, tb);
                subbeam_box(o, dir, a.w, ta, tb, p.rad_beam, p.box_slack, blo[jb], bhi[jb]);
                geo[3 * size_t(jb)] = a;
                geo[3 * size_t(jb) + 1] = make_float4(d.x, d.y, d.z, __int_as_float(int(slot)));
                geo[3 * size_t(jb) + 2] = make_float4(lo, hi, d.w, __int_as_float(flags));
                ref[jb] = make_uint2(slot, j | (n << 16));
            }
        }
    }
}

__global__ void k_gather_points(const uint32_t* order, const uint32_t* list, const VertexRec* verts, uint32_t cap,
                                const uint32_t* n_dev, float4* out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= min(cap, *n_dev)) return;
    const uint32_t slot = list[order[i]];
    const float4 a = verts[slot].a;
    out[i] = make_float4(a.x, a.y, a.z, __int_as_float(int(slot)));
}

__global__ void k_gather_beams(const uint32_t* order, const float4* geo_u, const uint2* ref_u, uint32_t cap,
                               const uint32_t* n_dev, float4* out, uint2* ref_out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= min(cap, *n_dev)) return;
    const size_t j = order[i];
    out[3 * size_t(i)] = geo_u[3 * j];
    out[3 * size_t(i) + 1] = geo_u[3 * j + 1];
    out[3 * size_t(i) + 2] = geo_u[3 * j + 2];
    ref_out[i] = ref_u[j];
}
__global__ void k_gather_tris(const uint32_t* order, const float4* tri, const float4* nrm, const int2* rec, uint32_t n,
                              float4* tri_s, float4* nrm_s, int2* rec_s)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = order[i];
    for (int k = 0; k < 3; k++) tri_s[3 * i + k] = tri[3 * j + k];
    for (int k = 0; k < 4; k++) nrm_s[4 * i + k] = nrm[4 * j + k];
    rec_s[i] = rec[j];
}

// the photon counts of this iteration (totals of the photon-list scans, tot[0..2]), each clamped
// to its list's capacity, into n[0..2] for the builds; a total beyond a capacity raises ovf[0],
// which makes k_accumulate discard the iteration (the host then grows the lists and renders it
// again)
__global__ void k_map_counts(const uint32_t* tot, uint3 cap, uint3 want, uint32_t* n, uint32_t* ovf)
{
    if (threadIdx.x != 0) return;
    n[0] = want.x ? min(tot[0], cap.x) : 0u;
    n[1] = want.y ? min(tot[1], cap.y) : 0u;
    n[2] = want.z ? min(tot[2], cap.z) : 0u;
    if ((want.x && tot[0] > cap.x) || (want.y && tot[1] > cap.y) || (want.z && tot[2] > cap.z)) ovf[0] = 1u;
}

// the counts into the map descriptor (whose pointers the host has just written)
__global__ void k_map_n(const uint32_t* n, PhotonMaps* m)
{
    if (threadIdx.x != 0) return;
    m->surf.n = n[0];
    m->med.n = m->wmed.n = n[1];
    m->beam.n = m->wbeam.n = n[2];
}

inline uint32_t nblk(size_t n, uint32_t t = 256) { return uint32_t((n + t - 1) / t); }

struct TriStore {
    float4 *tri = nullptr, *nrm = nullptr;
    int2* rec = nullptr;
    BvhStore bvh;
};

} // namespace

// per-iteration values the device reads, staged in pinned host memory: two slots used in turn,
// each reused only after the iteration that last used it has consumed it (its event)
struct Staging {
    IterParams p;
    PhotonMaps m;
    uint32_t ovf; // the iteration's photon-list overflow flag, copied back
};

struct upbp_renderer {
    std::string err;
    hipStream_t stream = nullptr;
    upbp_config cfg;
    DScene sc;
    std::vector<DevBuf> owned; // scene tables
    TriStore real, imag;
    BvhScratch scratch;
    float r0_bb1d, r0_bp2d, r0_pb2d, r0_pp3d, r0_surf;
    uint32_t n_light, n_pixels, bb1d_used, vstride, bstride;
    bool want_beams;
    Bufs bufs;
    DevBuf lstate, lstack, cstate, cstack, q0, q1, qcount, verts, nvert, beams, nbeam, fb32, accum, stats;
    DevBuf cnt, off, tot, list_s, list_m, lo_s, hi_s, lo_m, hi_m, lo_b, hi_b, pts_s, pts_m, geo_u, geo_b, ref_u, ref_b;
    DevBuf mcolor, segq, pieces, pieces_x, piece_cell, xwork, pkeys, pvals, unitn, units;
    DevBuf guide; // learned proposal records
    DevBuf mislog, miscount; // MIS weight log (upbp_set_mis_log)
    // device-resident iteration: parameters and photon-map descriptor in device memory, photon
    // counts (clamped) and the overflow flag; grow-only photon-list capacities
    DevBuf dparams, dmaps, mapn, ovf;
    Staging* stage = nullptr;
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    int pending[2] = {-1, -1}; // iteration whose overflow flag a staging slot holds (-1: none)
    int slot = 0;
    bool need_sizing = true;
    uint32_t cap_s = 0, cap_m = 0, cap_b = 0;
    PhotonMaps hmaps{};
    uint64_t gen = 0; // bumped when buffers or settings baked into kernel arguments change
    // captured iterations (upbp_tuning.graph): one per staging slot, with the generation it was
    // captured at; eager_gen: the generation of the last iteration run without capture
    bool use_graph = false;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    uint64_t gexec_gen[2] = {~0ull, ~0ull};
    uint64_t eager_gen = ~0ull;
    uint64_t graph_launches = 0;
    uint64_t resizes = 0; // photon-list sizing passes
    // online training of the proposal (upbp_set_guide)
    upbp_guide_config gcfg{};
    bool gtrain = false;
    GuideNet* gnet = nullptr;
    DevBuf gpend, graw, gsamp, gcount;
    uint32_t gcap = 0, gring = 0;
    double gloss = 0.0, gms = 0.0;
    uint64_t gsamples = 0;
    BvhStore bvh_s, bvh_m, bvh_b;
    WBvhStore wbvh_m, wbvh_b;
    float scene_mag = 1.f;       // largest coordinate magnitude of the scene and camera
    upbp_tuning tune;
    hipEvent_t ev[4 * kMaxBounces + 8];
};

namespace {

thread_local std::string g_err;

int fail(upbp_renderer* r, const std::string& m, int code)
{
    if (r) r->err = m;
    g_err = m;
    return code;
}

#define HIPCHK(expr)                                                                                                    \
    do {                                                                                                                \
        hipError_t e_ = (expr);                                                                                         \
        if (e_ != hipSuccess) return fail(r, std::string(#expr) + ": " + hipGetErrorString(e_), UPBP_E_HIP);            \
    } while (0)

template <typename T>
T* upload(upbp_renderer* r, const std::vector<T>& v, DevBuf& buf)
{
    if (buf.ensure(std::max<size_t>(v.size(), 1) * sizeof(T))) return nullptr;
    if (!v.empty() && hipMemcpy(buf.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
        return nullptr;
    return buf.as<T>();
}

f3 F3(const float* p) { return mk(p[0], p[1], p[2]); }

// host/medium.hpp:39-85
DMedium make_medium(const upbp_medium& in)
{
    DMedium m;
    m.cont_prob = in.continuation_prob;
    m.mean_cosine = in.mean_cosine;
    m.absorption = mk(fabsf(in.absorption[0]), fabsf(in.absorption[1]), fabsf(in.absorption[2]));
    m.emission = mk(fabsf(in.emission[0]), fabsf(in.emission[1]), fabsf(in.emission[2]));
    m.scattering = mk(fabsf(in.scattering[0]), fabsf(in.scattering[1]), fabsf(in.scattering[2]));
    m.attenuation = m.absorption + m.scattering;
    m.min_pos_att = m.attenuation.x;
    m.min_pos_idx = 0;
    if (m.attenuation.y > 0.f && (m.min_pos_att == 0 || m.attenuation.y < m.min_pos_att)) {
        m.min_pos_att = m.attenuation.y;
        m.min_pos_idx = 1;
    }
    if (m.attenuation.z > 0.f && (m.min_pos_att == 0 || m.attenuation.z < m.min_pos_att)) {
        m.min_pos_att = m.attenuation.z;
        m.min_pos_idx = 2;
    }
    m.has_scattering = m.scattering.x > 0 || m.scattering.y > 0 || m.scattering.z > 0;
    m.has_attenuation = m.attenuation.x > 0 || m.attenuation.y > 0 || m.attenuation.z > 0;
    m.max_beam_length = m.has_attenuation ? -logf(1e-9f) / maxc(m.attenuation) : kInf;
    return m;
}

int build_tris(upbp_renderer* r, const upbp_triangles& t, TriStore& out, TriSet& set)
{
    const uint32_t n = t.count;
    set.bvh.n = n;
    if (n == 0) {
        set.tri = nullptr;
        set.nrm = nullptr;
        set.rec = nullptr;
        set.bvh.nodes = nullptr;
        set.bvh.order = nullptr;
        return 0;
    }
    std::vector<float4> tri(3 * size_t(n)), nrm(4 * size_t(n)), lo(n), hi(n);
    std::vector<int2> rec(n);
    for (uint32_t i = 0; i < n; i++) {
        const float* v = t.verts + 9 * size_t(i);
        const f3 p0 = F3(v), p1 = F3(v + 3), p2 = F3(v + 6);
        const f3 e1 = p1 - p0, e2 = p2 - p0;
        tri[3 * i] = make_float4(p0.x, p0.y, p0.z, 0.f);
        tri[3 * i + 1] = make_float4(e1.x, e1.y, e1.z, 0.f);
        tri[3 * i + 2] = make_float4(e2.x, e2.y, e2.z, 0.f);
        for (int k = 0; k < 4; k++) {
            const float* q = t.normals + 12 * size_t(i) + 3 * k;
            nrm[4 * i + k] = make_float4(q[0], q[1], q[2], 0.f);
        }
        rec[i] = make_int2(t.mat_id[i], t.light_id[i]);
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
    DevBuf dtri, dnrm, drec, dlo, dhi;
    if (!upload(r, tri, dtri) || !upload(r, nrm, dnrm) || !upload(r, rec, drec) || !upload(r, lo, dlo) ||
        !upload(r, hi, dhi))
        return fail(r, "triangle upload failed", UPBP_E_HIP);
    if (lbvh_build(out.bvh, r->scratch, dlo.as<float4>(), dhi.as<float4>(), n, r->stream))
        return fail(r, "triangle LBVH build failed", UPBP_E_HIP);
    HIPCHK(hipMalloc(&out.tri, tri.size() * sizeof(float4)));
    HIPCHK(hipMalloc(&out.nrm, nrm.size() * sizeof(float4)));
    HIPCHK(hipMalloc(&out.rec, rec.size() * sizeof(int2)));
    hipLaunchKernelGGL(k_gather_tris, dim3(nblk(n)), dim3(256), 0, r->stream, out.bvh.order, dtri.as<float4>(),
                       dnrm.as<float4>(), drec.as<int2>(), n, out.tri, out.nrm, out.rec);
    HIPCHK(hipStreamSynchronize(r->stream));
    for (DevBuf* b : {&dtri, &dnrm, &drec, &dlo, &dhi}) b->release();
    set.tri = out.tri;
    set.nrm = out.nrm;
    set.rec = out.rec;
    set.bvh.nodes = out.bvh.nodes;
    set.bvh.order = out.bvh.order;
    return 0;
}

// renderer.cpp:56-119 + renderer_init.cpp:713-736, 891-916
void setup_params(const upbp_renderer* r, int iteration, IterParams& p)
{
    const upbp_config& c = r->cfg;
    const uint32_t f = c.algorithm_flags;
    memset(&p, 0, sizeof(p));
    p.trace_light = p.trace_camera = f != 0;
    p.do_bpt = (f & T_BPT) != 0;
    p.do_bb1d = (f & T_BB1D) != 0;
    p.do_bp2d = (f & T_BP2D) != 0;
    p.do_pb2d = (f & T_PB2D) != 0;
    p.do_pp3d = (f & T_PP3D) != 0;
    p.do_surf = (f & T_SURF) != 0;
    if (f & T_SPECULAR_ONLY) {
        p.trace_light = 0;
        p.trace_camera = 1;
        p.do_bpt = p.do_bb1d = p.do_bp2d = p.do_pb2d = p.do_pp3d = p.do_surf = 0;
    }
    p.connect_from_surf = (f & (T_PREVIOUS | T_COMPATIBLE)) == 0;
    p.n_pixels = r->n_pixels;
    p.n_light = r->n_light;
    const bool merging = c.max_path_length > 1;
    p.build_bb1d = p.trace_light && merging && p.do_bb1d;
    p.build_bp2d = p.trace_light && merging && p.do_bp2d;
    p.build_pb2d = p.trace_light && merging && p.do_pb2d;
    p.build_pp3d = p.trace_light && merging && p.do_pp3d;
    p.build_surf = p.trace_light && merging && p.do_surf;
    p.min_len = c.min_path_length;
    p.max_len = c.max_path_length;
    p.pbt = c.photon_beam_type;
    p.qbt = c.query_beam_type;
    p.techniques = f;
    p.bb1d_used = r->bb1d_used;
    p.vstride = r->vstride;
    p.bstride = r->bstride;
    p.ignore_spec = c.ignore_fully_spec_paths != 0;
    p.iter_key = uint64_t(iteration) + c.iteration_offset;
    const float n_light = float(r->n_light);
    const float bb1d_used = float(r->bb1d_used);
    p.n_bpt = float(p.do_bpt);
    const float n_bb1d = float(p.do_bb1d) * bb1d_used, n_bp2d = float(p.do_bp2d) * n_light,
                n_pb2d = float(p.do_pb2d) * n_light, n_pp3d = float(p.do_pp3d) * n_light,
                n_surf = float(p.do_surf) * n_light;
    const float eff = 1.f + float(iteration);
    const float r_bb1d = fmaxf(
        r->r0_bb1d * powf(1.f + float(iteration) * bb1d_used / n_light, c.bb1d_radius_alpha - 1.f), 1e-7f);
    const float r_bp2d = fmaxf(r->r0_bp2d * powf(eff, (c.bp2d_radius_alpha - 1.f) * .5f), 1e-7f);
    const float r_pb2d = fmaxf(r->r0_pb2d * powf(eff, (c.pb2d_radius_alpha - 1.f) * .5f), 1e-7f);
    const float r_pp3d = fmaxf(r->r0_pp3d * powf(eff, (c.pp3d_radius_alpha - 1.f) * (1.f / 3.f)), 1e-7f);
    const float r_surf = fmaxf(r->r0_surf * powf(eff, (c.surf_radius_alpha - 1.f) * .5f), 1e-7f);
    p.mf_bb1d = .5f * r_bb1d * n_bb1d;
    p.mf_bp2d = (kPi * (r_bp2d * r_bp2d)) * n_bp2d;
    p.mf_pb2d = (kPi * (r_pb2d * r_pb2d)) * n_pb2d;
    p.mf_pp3d = (4.f / 3.f) * (kPi * (r_pp3d * r_pp3d * r_pp3d)) * n_pp3d;
    p.mf_surf = (kPi * (r_surf * r_surf)) * n_surf;
    p.norm_bb1d = 1.f / bb1d_used;
    p.norm_bp2d = 1.f / n_light;
    p.norm_pb2d = 1.f / n_light;
    p.norm_pp3d = 1.f / p.mf_pp3d;
    p.norm_surf = 1.f / p.mf_surf;
    p.r_bb1d = r_bb1d;
    p.r_bp2d = r_bp2d;
    p.r_pb2d = r_pb2d;
    p.r_pp3d = r_pp3d;
    p.r_surf = r_surf;
    p.rad_point_med = 0.f;
    if (p.do_pp3d) p.rad_point_med = fmaxf(p.rad_point_med, r_pp3d);
    if (p.do_pb2d) p.rad_point_med = fmaxf(p.rad_point_med, r_pb2d);
    p.rad_beam = 0.f;
    if (p.do_bb1d) p.rad_beam = fmaxf(p.rad_beam, r_bb1d);
    if (p.do_bp2d) p.rad_beam = fmaxf(p.rad_beam, r_bp2d);
    // beam map: sub-beam length and the box slack for the FP32 error of the beam predicates'
    // line parameters near the parallel cut-off, which grows with the coordinate magnitude
    p.sub_len = r->tune.subbeam_factor > 0.f ? r->tune.subbeam_factor * p.rad_beam : 0.f;
    p.box_slack = 2e-4f * r->scene_mag;
    p.merge_bpc = r->tune.merge_blocks_per_cu;
    p.piece_order = r->tune.piece_order;
    p.packet_bounces = r->tune.packet_bounces;
    p.beam_leaf = r->tune.beam_leaf;
    // bundles list into pieces_x and read the sub-beam records: not with a sorted piece order,
    // beam references or packets
    p.bundle_bounces =
        r->tune.piece_order == 0 && r->tune.beam_leaf == 0 && r->tune.packet_bounces == 0 ? r->tune.bundle_bounces : 0;
    p.bundle_radius = r->tune.bundle_factor * p.rad_beam;
    p.bundle_box = p.sub_len > 0.f ? 0.5f * p.sub_len + p.rad_beam : 8.f * p.rad_beam;
    p.block = r->tune.block_size > 0 ? uint32_t(r->tune.block_size) : 64u;
    p.cap_med = r->cap_m;
    p.cap_beam = r->cap_b;
    {
        float rq = 0.f;
        if (p.do_pb2d) rq = fmaxf(rq, r_pb2d);
        if (p.do_bb1d) rq = fmaxf(rq, r_bb1d);
        p.piece_len = r->tune.query_piece_factor > 0.f ? r->tune.query_piece_factor * rq : 0.f;
        p.bundle_piece_len = p.piece_len > 0.f && r->tune.bundle_piece_factor > 0.f
                                 ? r->tune.bundle_piece_factor * rq
                                 : p.piece_len;
    }
    p.r2_bb1d = r_bb1d * r_bb1d;
    p.r2_bp2d = r_bp2d * r_bp2d;
    p.r2_pb2d = r_pb2d * r_pb2d;
    p.r2_pp3d = r_pp3d * r_pp3d;
    p.r2_surf = r_surf * r_surf;
}

} // namespace

extern "C" const char* upbp_last_error(const upbp_renderer* r) { return r ? r->err.c_str() : g_err.c_str(); }

// Every index the kernels dereference must be in range: a bad one would read outside a
// device table.
static const char* invalid_scene(const upbp_scene& s)
{
    const int64_t nm = s.mat_count, nd = s.med_count, nl = s.light_count;
    if (nm == 0 || nd == 0) return "scene needs at least one material and one medium";
    if (int64_t(s.global_medium_idx) >= nd) return "global medium index out of range";
    if (s.background_light_idx >= 0 && s.background_light_idx >= nl) return "background light index out of range";
    for (uint32_t i = 0; i < s.mat_count; i++)
        if (s.materials[i].med_idx < -1 || s.materials[i].med_idx >= nd) return "material medium index out of range";
    for (uint32_t i = 0; i < s.light_count; i++)
        if (s.lights[i].mat_id < -1 || s.lights[i].mat_id >= nm || s.lights[i].med_id >= nd)
            return "light material or medium index out of range";
    if (s.camera.mat_id < -1 || s.camera.mat_id >= nm || s.camera.med_id >= nd)
        return "camera material or medium index out of range";
    if (s.env_map.rgb && (s.env_map.height == 0 || s.env_map.width < 2 * s.env_map.height ||
                          s.env_map.height > (1u << 15)))
        return "environment map must be at least twice as wide as high";
    for (const upbp_triangles* t : {&s.real, &s.imag})
        for (uint32_t i = 0; i < t->count; i++)
            if (t->mat_id[i] < 0 || t->mat_id[i] >= nm || t->light_id[i] < -1 || t->light_id[i] >= nl)
                return "triangle material or light index out of range";
    for (const upbp_spheres* t : {&s.sph_real, &s.sph_imag})
        for (uint32_t i = 0; i < t->count; i++)
            if (t->mat_id[i] < 0 || t->mat_id[i] >= nm || (t->light_id && (t->light_id[i] < -1 || t->light_id[i] >= nl)) ||
                !(t->spheres[4 * i + 3] > 0.f))
                return "sphere material, light index or radius out of range";
    return nullptr;
}

extern "C" int upbp_create(const upbp_scene* scene, const upbp_config* cfg, upbp_renderer** out)
{
    upbp_renderer* r = nullptr;
    if (!scene || !cfg || !out) return fail(r, "null argument", UPBP_E_ARG);
    *out = nullptr;
    if (scene->light_count == 0) return fail(r, "scene has no lights", UPBP_E_ARG);
    if (const char* bad = invalid_scene(*scene)) return fail(r, bad, UPBP_E_ARG);
    if (scene->mat_count > 32767u || scene->med_count > 32767u)
        return fail(r, "at most 32767 materials and 32767 media", UPBP_E_ARG);
    if (cfg->path_count_per_iter < 1.f) return fail(r, "light path count must be positive", UPBP_E_ARG);
    if (cfg->max_path_length >= uint32_t(kMaxBounces)) return fail(r, "max path length must be below 64", UPBP_E_ARG);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(r, "no GPU visible", UPBP_E_NODEVICE);
    r = new upbp_renderer();
    r->cfg = *cfg;
    const upbp_config& c = *cfg;
    HIPCHK(hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking));
    for (auto& e : r->ev) HIPCHK(hipEventCreate(&e));

    // radii relative to the scene radius when negative (renderer_init.cpp:681-711)
    const float rad = scene->scene_radius;
    auto rel = [&](float v) { return v < 0.f ? -v * rad : v; };
    r->r0_bb1d = rel(c.bb1d_radius_initial);
    r->r0_bp2d = rel(c.bp2d_radius_initial);
    r->r0_pb2d = rel(c.pb2d_radius_initial);
    r->r0_pp3d = rel(c.pp3d_radius_initial);
    r->r0_surf = rel(c.surf_radius_initial);
    r->n_pixels = c.res_x * c.res_y;
    r->n_light = uint32_t(c.path_count_per_iter);
    float used = c.bb1d_used_light_subpath_count;
    if (used < 0.f) used = floorf(-used * float(r->n_light));
    r->bb1d_used = uint32_t(used);
    r->vstride = std::max<uint32_t>(c.max_path_length - 1, 1);
    r->want_beams = (c.algorithm_flags & (T_BB1D | T_BP2D)) != 0 && !(c.algorithm_flags & T_SPECULAR_ONLY);
    r->bstride = r->want_beams ? 4 * r->vstride : 0;

    // scene tables
    DScene& sc = r->sc;
    memset(&sc, 0, sizeof(sc));
    std::vector<DMaterial> mats(scene->mat_count);
    for (uint32_t i = 0; i < scene->mat_count; i++) {
        const upbp_material& m = scene->materials[i];
        mats[i] = {F3(m.diffuse), F3(m.phong), F3(m.mirror), m.phong_exponent, m.ior, m.priority, m.real, m.med_idx};
    }
    std::vector<DMedium> media(scene->med_count);
    for (uint32_t i = 0; i < scene->med_count; i++) media[i] = make_medium(scene->media[i]);
    std::vector<DLight> lights(scene->light_count);
    for (uint32_t i = 0; i < scene->light_count; i++) {
        const upbp_light& l = scene->lights[i];
        DLight& o = lights[i];
        memset(&o, 0, sizeof(o));
        o.type = l.type;
        o.mat_id = l.mat_id;
        o.med_id = l.med_id;
        o.intensity = F3(l.intensity);
        o.scale = l.scale;
        o.uses_env = l.type == UPBP_LIGHT_BACKGROUND && l.uses_env_map && scene->env_map.rgb ? 1 : 0;
        if (l.type == UPBP_LIGHT_AREA) { // host/light.hpp:30-47
            o.p0 = F3(l.p0);
            o.e1 = F3(l.p1) - o.p0;
            o.e2 = F3(l.p2) - o.p0;
            const f3 nrm = cross(o.e1, o.e2);
            o.inv_area = 2.f / sqrtf(dot(nrm, nrm));
            o.frame.set_from_z(nrm);
        } else if (l.type == UPBP_LIGHT_DIRECTIONAL) {
            o.frame.set_from_z(F3(l.position));
        } else if (l.type == UPBP_LIGHT_POINT) {
            o.position = F3(l.position);
        }
    }
    // analytic spheres (predefined scenes)
    std::vector<float4> sph[2];
    std::vector<int2> sph_rec[2];
    for (int k = 0; k < 2; k++) {
        const upbp_spheres& S = k == 0 ? scene->sph_real : scene->sph_imag;
        for (uint32_t i = 0; i < S.count; i++) {
            sph[k].push_back(make_float4(S.spheres[4 * i], S.spheres[4 * i + 1], S.spheres[4 * i + 2], S.spheres[4 * i + 3]));
            sph_rec[k].push_back(make_int2(S.mat_id[i], S.light_id ? S.light_id[i] : -1));
        }
    }
    // environment map: image and luminance distribution
    EnvTables env;
    if (scene->env_map.rgb && scene->env_map.width && scene->env_map.height)
        build_env_tables(scene->env_map, env);
    r->owned.resize(14);
    sc.mats = upload(r, mats, r->owned[0]);
    sc.media = upload(r, media, r->owned[1]);
    sc.lights = upload(r, lights, r->owned[2]);
    sc.sph_real = {upload(r, sph[0], r->owned[3]), upload(r, sph_rec[0], r->owned[4]), uint32_t(sph[0].size())};
    sc.sph_imag = {upload(r, sph[1], r->owned[5]), upload(r, sph_rec[1], r->owned[6]), uint32_t(sph[1].size())};
    if (!env.img.empty()) {
        DEnvMap& m = sc.env;
        m.img = upload(r, env.img, r->owned[7]);
        m.cond_func = upload(r, env.cond_func, r->owned[8]);
        m.cond_cdf = upload(r, env.cond_cdf, r->owned[9]);
        m.cond_int = upload(r, env.cond_int, r->owned[10]);
        m.marg_func = upload(r, env.marg_func, r->owned[11]);
        m.marg_cdf = upload(r, env.marg_cdf, r->owned[12]);
        m.marg_int = env.marg_int;
        m.width = int(scene->env_map.width);
        m.height = int(scene->env_map.height);
        m.nu = 2 * m.height;
        m.nv = m.height;
        m.norm = .5f * kInvPi * kInvPi;
        if (!m.img || !m.cond_func || !m.cond_cdf || !m.cond_int || !m.marg_func || !m.marg_cdf) {
            upbp_destroy(r);
            return fail(nullptr, "environment map upload failed", UPBP_E_HIP);
        }
    }
    if (!sc.mats || !sc.media || !sc.lights || !sc.sph_real.sph || !sc.sph_real.rec || !sc.sph_imag.sph ||
        !sc.sph_imag.rec) {
        upbp_destroy(r);
        return fail(nullptr, "scene upload failed", UPBP_E_HIP);
    }
    sc.n_mats = scene->mat_count;
    sc.n_media = scene->med_count;
    sc.n_lights = scene->light_count;
    sc.global_medium = scene->global_medium_idx;
    sc.background_light = scene->background_light_idx;
    sc.center = F3(scene->scene_center);
    sc.radius = scene->scene_radius;
    sc.inv_radius_sqr = 1.f / (sc.radius * sc.radius);
    sc.use_shading_normal = c.use_shading_normal;
    { // camera basis (host/camera.hpp:33-76)
        const upbp_camera& cam = scene->camera;
        DCamera& dc = sc.cam;
        dc.origin = F3(cam.origin);
        dc.forward = normalize(F3(cam.target) - dc.origin);
        dc.right = normalize(cross(F3(cam.roll), -dc.forward));
        dc.up = cross(-dc.forward, dc.right);
        dc.res_x = float(c.res_x);
        dc.res_y = float(c.res_y);
        dc.ipd = float(c.res_x) / (2.f * tanf(cam.horizontal_fov_deg * kPi / 360.f));
        dc.mat_id = cam.mat_id;
        dc.med_id = cam.med_id;
    }
    { // largest coordinate magnitude of geometry and camera (scales the beam-box slack)
        float mag = 0.f;
        for (const upbp_triangles* t : {&scene->real, &scene->imag})
            for (size_t i = 0; i < size_t(t->count) * 9; i++) mag = fmaxf(mag, fabsf(t->verts[i]));
        for (int k = 0; k < 2; k++)
            for (const float4& q : sph[k])
                mag = fmaxf(mag, fmaxf(fabsf(q.x), fmaxf(fabsf(q.y), fabsf(q.z))) + q.w);
        for (int a = 0; a < 3; a++) mag = fmaxf(mag, fabsf(scene->camera.origin[a]));
        r->scene_mag = fmaxf(mag, 1e-3f);
    }
    upbp_tuning_defaults(&r->tune);
    int e = build_tris(r, scene->real, r->real, sc.real);
    if (!e) e = build_tris(r, scene->imag, r->imag, sc.imag);
    if (e) {
        const std::string m = r->err;
        upbp_destroy(r);
        return fail(nullptr, m, e);
    }

    // wavefront buffers
    const size_t NL = r->n_light, NP = r->n_pixels, NQ = std::max(NL, NP);
    bool ok = !r->lstate.ensure(NL * sizeof(PathState)) && !r->lstack.ensure(NL * sizeof(StackRec)) &&
              !r->cstate.ensure(NP * sizeof(PathState)) && !r->cstack.ensure(NP * sizeof(StackRec)) &&
              !r->q0.ensure(NQ * 4) && !r->q1.ensure(NQ * 4) && !r->qcount.ensure(kQcountWords * 4) &&
              !r->mcolor.ensure(NP * 12) &&
              !r->verts.ensure(NL * r->vstride * sizeof(VertexRec)) && !r->nvert.ensure(NL * 4) &&
              !r->nbeam.ensure(NL * 4) && !r->fb32.ensure(NP * 12) && !r->accum.ensure(NP * 24) &&
              !r->stats.ensure(STAT_COUNT * 8) && !r->cnt.ensure(3 * NL * 4) && !r->off.ensure(3 * NL * 4) &&
              !r->tot.ensure(16);
    if (ok && r->want_beams) ok = !r->beams.ensure(NL * r->bstride * sizeof(BeamRec));
    // segment queries of one camera bounce: at most kMaxSegments per live camera path
    const bool seg_merges = (c.algorithm_flags & (T_PB2D | T_BB1D)) != 0;
    const size_t seg_cap = seg_merges ? NP * kMaxSegments : 1;
    const size_t piece_cap = seg_merges ? NP * (kPiecesPerRay + kMaxSegments + 1) : 1;
    if (ok)
        ok = !r->segq.ensure(seg_cap * sizeof(SegRec)) && !r->pieces.ensure(piece_cap * sizeof(uint2)) &&
             !r->pieces_x.ensure((piece_cap + 63) / 64 * 64 * sizeof(uint2)) && !r->piece_cell.ensure(piece_cap) &&
             !r->xwork.ensure(size_t(kMaxBounces) * kXWork * 4) &&
             !r->pkeys.ensure((piece_cap + 63) / 64 * 4) && !r->pvals.ensure((piece_cap + 63) / 64 * 4) &&
             !r->unitn.ensure(16);
    // piece bundles of the B-B1D merge: at most one per piece
    if (ok && (c.algorithm_flags & T_BB1D)) ok = !r->units.ensure(piece_cap * sizeof(uint2));
    if (!ok) {
        upbp_destroy(r);
        return fail(nullptr, "device allocation failed", UPBP_E_HIP);
    }
    if (r->dparams.ensure(sizeof(IterParams)) || r->dmaps.ensure(sizeof(PhotonMaps)) || r->mapn.ensure(16) ||
        r->ovf.ensure(16) || hipHostMalloc(reinterpret_cast<void**>(&r->stage), 2 * sizeof(Staging)) != hipSuccess) {
        upbp_destroy(r);
        return fail(nullptr, "device allocation failed", UPBP_E_HIP);
    }
    memset(r->stage, 0, 2 * sizeof(Staging));
    for (auto& e : r->stage_ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIPCHK(hipMemsetAsync(r->accum.p, 0, NP * 24, r->stream));
    HIPCHK(hipMemsetAsync(r->dmaps.p, 0, sizeof(PhotonMaps), r->stream));
    HIPCHK(hipStreamSynchronize(r->stream));
    Bufs& b = r->bufs;
    memset(&b, 0, sizeof(b));
    b.params = r->dparams.as<IterParams>();
    b.maps = r->dmaps.as<PhotonMaps>();
    b.lstate = r->lstate.as<PathState>();
    b.lstack = r->lstack.as<StackRec>();
    b.cstate = r->cstate.as<PathState>();
    b.cstack = r->cstack.as<StackRec>();
    b.queue[0] = r->q0.as<uint32_t>();
    b.queue[1] = r->q1.as<uint32_t>();
    b.qcount = r->qcount.as<uint32_t>();
    b.verts = r->verts.as<VertexRec>();
    b.nvert = r->nvert.as<uint32_t>();
    b.beams = r->beams.as<BeamRec>();
    b.nbeam = r->nbeam.as<uint32_t>();
    b.fb32 = r->fb32.as<float>();
    b.mcolor = r->mcolor.as<float>();
    b.segq = r->segq.as<SegRec>();
    b.seg_cap = uint32_t(seg_cap);
    b.pieces = r->pieces.as<uint2>();
    b.pieces_x = r->pieces_x.as<uint2>();
    b.units = (c.algorithm_flags & T_BB1D) ? r->units.as<uint2>() : nullptr;
    b.piece_cell = r->piece_cell.as<uint8_t>();
    b.xwork = r->xwork.as<uint32_t>();
    b.piece_cap = uint32_t(piece_cap);
    b.accum = r->accum.as<double>();
    b.stats = r->stats.as<unsigned long long>();
    *out = r;
    return UPBP_OK;
}

namespace {

// Photon lists (path order) -> boxes -> three LBVHs -> sorted payloads. The lists live at
// grow-only capacities; their counts stay on the device (k_map_counts), so the build, like the
// rest of the iteration, needs no host synchronisation. A sizing pass (the first iteration, and
// the one after a capacity was exceeded) reads the totals back once and sets the capacities with
// a quarter to spare.
int build_photon_maps(upbp_renderer* r, const IterParams& p, Staging& st)
{
    Bufs& b = r->bufs;
    hipStream_t s = r->stream;
    const uint32_t NL = r->n_light;
    const int want_surf = p.build_surf, want_med = p.build_pp3d || p.build_pb2d;
    const int want_pts = want_surf || want_med;
    const int want_beams = (p.build_bb1d || p.build_bp2d) && r->want_beams;
    uint32_t* cnt = r->cnt.as<uint32_t>();
    uint32_t* off = r->off.as<uint32_t>();
    uint32_t* tot = r->tot.as<uint32_t>();
    hipLaunchKernelGGL(k_count_photons, dim3(nblk(NL)), dim3(256), 0, s, b, cnt, cnt + NL, cnt + 2 * NL, want_pts,
                       want_beams);
    for (int k = 0; k < 3; k++)
        if (scan_exclusive(cnt + size_t(k) * NL, off + size_t(k) * NL, NL, tot + k, r->scratch, s))
            return fail(r, "photon list scan failed", UPBP_E_HIP);
    if (r->need_sizing) {
        uint32_t host_tot[3];
        HIPCHK(hipMemcpyAsync(host_tot, tot, 12, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        auto grow = [](uint32_t cap, uint32_t need) {
            const uint64_t c = std::max<uint64_t>(uint64_t(need) + need / 4, 1024);
            return uint32_t(std::min<uint64_t>(std::max<uint64_t>(cap, c), 0xffffffffu));
        };
        r->cap_s = want_surf ? grow(r->cap_s, host_tot[0]) : r->cap_s;
        r->cap_m = want_med ? grow(r->cap_m, host_tot[1]) : r->cap_m;
        r->cap_b = want_beams ? grow(r->cap_b, host_tot[2]) : r->cap_b;
        r->need_sizing = false;
        r->resizes++;
        r->gen++;
    }
    const uint32_t cs = want_surf ? r->cap_s : 0, cm = want_med ? r->cap_m : 0, cb = want_beams ? r->cap_b : 0;
    bool ok = !r->list_s.ensure(std::max<size_t>(cs, 1) * 4) && !r->lo_s.ensure(std::max<size_t>(cs, 1) * 16) &&
              !r->hi_s.ensure(std::max<size_t>(cs, 1) * 16) && !r->pts_s.ensure(std::max<size_t>(cs, 1) * 16) &&
              !r->list_m.ensure(std::max<size_t>(cm, 1) * 4) && !r->lo_m.ensure(std::max<size_t>(cm, 1) * 16) &&
              !r->hi_m.ensure(std::max<size_t>(cm, 1) * 16) && !r->pts_m.ensure(std::max<size_t>(cm, 1) * 16) &&
              !r->geo_u.ensure(std::max<size_t>(cb, 1) * 48) && !r->lo_b.ensure(std::max<size_t>(cb, 1) * 16) &&
              !r->hi_b.ensure(std::max<size_t>(cb, 1) * 16) && !r->geo_b.ensure(std::max<size_t>(cb, 1) * 48) &&
              !r->ref_u.ensure(std::max<size_t>(cb, 1) * 8) && !r->ref_b.ensure(std::max<size_t>(cb, 1) * 8);
    if (!ok) return fail(r, "photon list allocation failed", UPBP_E_HIP);
    const uint3 cap = make_uint3(cs, cm, cb);
    hipLaunchKernelGGL(k_emit_photons, dim3(nblk(NL)), dim3(256), 0, s, b, off, off + NL, off + 2 * NL, want_surf,
                       want_med, want_beams, r->list_s.as<uint32_t>(), r->lo_s.as<float4>(), r->hi_s.as<float4>(),
                       r->list_m.as<uint32_t>(), r->lo_m.as<float4>(), r->hi_m.as<float4>(), r->geo_u.as<float4>(),
                       r->ref_u.as<uint2>(),
                       r->lo_b.as<float4>(), r->hi_b.as<float4>(), cap);
    uint32_t* n = r->mapn.as<uint32_t>();
    hipLaunchKernelGGL(k_map_counts, dim3(1), dim3(64), 0, s, tot, cap, make_uint3(want_surf, want_med, want_beams), n,
                       r->ovf.as<uint32_t>());
    if (lbvh_build(r->bvh_s, r->scratch, r->lo_s.as<float4>(), r->hi_s.as<float4>(), cs, s, n) ||
        lbvh_build(r->bvh_m, r->scratch, r->lo_m.as<float4>(), r->hi_m.as<float4>(), cm, s, n + 1) ||
        lbvh_build(r->bvh_b, r->scratch, r->lo_b.as<float4>(), r->hi_b.as<float4>(), cb, s, n + 2) ||
        wbvh_build(r->wbvh_m, r->bvh_m, r->scratch, cm, s, n + 1) ||
        wbvh_build(r->wbvh_b, r->bvh_b, r->scratch, cb, s, n + 2))
        return fail(r, "photon LBVH build failed", UPBP_E_HIP);
    // the descriptor: pointers and capacities from the host, counts from the device
    PhotonMaps& M = r->hmaps;
    M.surf = {r->bvh_s.nodes, r->bvh_s.order, cs};
    M.med = {r->bvh_m.nodes, r->bvh_m.order, cm};
    M.beam = {r->bvh_b.nodes, r->bvh_b.order, cb};
    M.wmed = {r->wbvh_m.nodes, cm};
    M.wbeam = {r->wbvh_b.nodes, cb};
    M.surf_pts = r->pts_s.as<float4>();
    M.med_pts = r->pts_m.as<float4>();
    M.beam_geo = r->geo_b.as<float4>();
    M.beam_ref = r->ref_b.as<uint2>();
    st.m = M;
    HIPCHK(hipMemcpyAsync(r->dmaps.p, &st.m, sizeof(PhotonMaps), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_map_n, dim3(1), dim3(64), 0, s, n, r->dmaps.as<PhotonMaps>());
    if (cs)
        hipLaunchKernelGGL(k_gather_points, dim3(nblk(cs)), dim3(256), 0, s, r->bvh_s.order, r->list_s.as<uint32_t>(),
                           b.verts, cs, n, r->pts_s.as<float4>());
    if (cm)
        hipLaunchKernelGGL(k_gather_points, dim3(nblk(cm)), dim3(256), 0, s, r->bvh_m.order, r->list_m.as<uint32_t>(),
                           b.verts, cm, n + 1, r->pts_m.as<float4>());
    if (cb)
        hipLaunchKernelGGL(k_gather_beams, dim3(nblk(cb)), dim3(256), 0, s, r->bvh_b.order, r->geo_u.as<float4>(),
                           r->ref_u.as<uint2>(), cb, n + 2, r->geo_b.as<float4>(), r->ref_b.as<uint2>());
    return 0;
}

float elapsed(hipEvent_t a, hipEvent_t b)
{
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, a, b);
    return ms;
}

} // namespace

// Adam steps of the proposal on the samples the camera pass collected, then new records; the
// next iteration samples from them. One host synchronisation (the sample count).
static int train_guide(upbp_renderer* r)
{
    hipStream_t s = r->stream;
    uint32_t n = 0;
    HIPCHK(hipEventRecord(r->ev[4], s));
    HIPCHK(hipMemcpyAsync(&n, r->gcount.as<uint32_t>() + 3, 4, hipMemcpyDeviceToHost, s)); // samples in the ring
    HIPCHK(hipStreamSynchronize(s));
    n = std::min(n, r->gring);
    r->gsamples = n;
    r->gloss = 0.0;
    r->gms = 0.0;
    if (n < 256) return UPBP_OK;
    const upbp_guide_config& g = r->gcfg;
    const uint32_t batch = std::min<uint32_t>(uint32_t(g.batch), n);
    double loss = 0.0;
    for (int k = 0; k < g.steps_per_iteration; k++) {
        if (r->gnet->gradient_drawn(r->gsamp.as<float>(), n, batch, uint32_t(r->gnet->step), s) ||
            r->gnet->adam_step(g.learning_rate, s))
            return fail(r, "guide training failed", UPBP_E_HIP);
        float l = 0.f;
        HIPCHK(hipMemcpyAsync(&l, r->gnet->loss.p, 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        loss += double(l) / double(batch);
    }
    if (g.steps_per_iteration > 0) {
        if (r->gnet->records(r->sc.guide.res, r->guide.as<float4>(), s)) return fail(r, "guide refresh failed", UPBP_E_HIP);
        if (r->sc.guide.alpha != g.alpha || r->sc.guide.alpha_s != g.surface_alpha) r->gen++;
        r->sc.guide.alpha = g.alpha;
        r->sc.guide.alpha_s = g.surface_alpha;
        r->gloss = loss / double(g.steps_per_iteration);
    }
    HIPCHK(hipEventRecord(r->ev[5], s));
    HIPCHK(hipEventSynchronize(r->ev[5]));
    r->gms = elapsed(r->ev[4], r->ev[5]);
    return UPBP_OK;
}

namespace {

int settle(upbp_renderer* r, bool wait);

// Enqueues one iteration on r's stream. Nothing in it waits for the device, except a sizing
// pass of the photon lists and the statistics (stats != null), which synchronise.
// The work of one iteration on r's stream, reading its parameters from the staging slot st.
// p.cap_* are set to the photon-list capacities the launches below depend on.
// The iteration's start on r's stream: its parameters from the staging slot to the device and
// the per-iteration counters and buffers cleared.
int enqueue_head(upbp_renderer* r, Staging& st)
{
    hipStream_t s = r->stream;
    Bufs& b = r->bufs;
    HIPCHK(hipMemcpyAsync(r->dparams.p, &st.p, sizeof(IterParams), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_fill, dim3(nblk(kQcountWords)), dim3(256), 0, s, b.qcount, 0u, size_t(kQcountWords));
    hipLaunchKernelGGL(k_fill, dim3(nblk(size_t(kMaxBounces) * kXWork)), dim3(256), 0, s, b.xwork, 0u,
                       size_t(kMaxBounces) * kXWork);
    hipLaunchKernelGGL(k_fill, dim3(nblk(3 * size_t(r->n_pixels))), dim3(256), 0, s, reinterpret_cast<uint32_t*>(b.fb32),
                       0u, 3 * size_t(r->n_pixels));
    hipLaunchKernelGGL(k_fill, dim3(nblk(3 * size_t(r->n_pixels))), dim3(256), 0, s,
                       reinterpret_cast<uint32_t*>(b.mcolor), 0u, 3 * size_t(r->n_pixels));
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, s, reinterpret_cast<uint32_t*>(b.stats), 0u,
                       size_t(2 * STAT_COUNT));
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, s, r->ovf.as<uint32_t>(), 0u, size_t(1));
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, s, r->mapn.as<uint32_t>(), 0u, size_t(3));
    if (b.mislog) hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, s, b.miscount, 0u, size_t(1));
    return UPBP_OK;
}

// The photon maps of the iteration's light paths on r's stream; p.cap_* are then the photon-list
// capacities the launches after it depend on.
int enqueue_maps(upbp_renderer* r, IterParams& p, Staging& st)
{
    const int e = build_photon_maps(r, p, st);
    if (e) return e;
    p.cap_med = (p.build_pp3d || p.build_pb2d) ? r->cap_m : 0u;
    p.cap_beam = (p.build_bb1d || p.build_bp2d) && r->want_beams ? r->cap_b : 0u;
    return UPBP_OK;
}

int enqueue_body(upbp_renderer* r, IterParams& p, Staging& st, bool timed)
{
    hipStream_t s = r->stream;
    Bufs& b = r->bufs;
    const int L = int(r->cfg.max_path_length);
    hipEvent_t* ev = r->ev;
    if (timed) HIPCHK(hipEventRecord(ev[0], s));
    {
        const int e = enqueue_head(r, st);
        if (e) return e;
    }
    if (p.trace_light) {
        launch_light_generate(r->sc, p, b, s);
        for (int k = 0; k < L; k++) launch_light_bounce(r->sc, p, b, k, s);
    } else {
        hipLaunchKernelGGL(k_fill, dim3(nblk(r->n_light)), dim3(256), 0, s, b.nvert, 0u, size_t(r->n_light));
        hipLaunchKernelGGL(k_fill, dim3(nblk(r->n_light)), dim3(256), 0, s, b.nbeam, 0u, size_t(r->n_light));
    }
    if (timed) HIPCHK(hipEventRecord(ev[1], s));
    if (p.trace_light) {
        // launch decisions below use the capacities a sizing pass may just have set
        const int e = enqueue_maps(r, p, st);
        if (e) return e;
    } else {
        p.cap_med = p.cap_beam = 0u;
        hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, s, r->dmaps.as<uint32_t>(), 0u,
                           sizeof(PhotonMaps) / sizeof(uint32_t));
    }
    if (timed) HIPCHK(hipEventRecord(ev[2], s));
    const bool collect = r->gtrain && p.trace_camera;
    b.gpend = collect ? r->gpend.as<float>() : nullptr;
    b.graw = collect ? r->graw.as<float>() : nullptr;
    b.gsamp = collect ? r->gsamp.as<float>() : nullptr;
    b.gcount = collect ? r->gcount.as<uint32_t>() : nullptr;
    b.gcap = collect ? r->gcap : 0u;
    b.gring = collect ? r->gring : 1u;
    if (collect) {
        hipLaunchKernelGGL(k_fill, dim3(nblk(size_t(r->n_pixels) * kGuidePend)), dim3(256), 0, s,
                           r->gpend.as<uint32_t>(), 0u, size_t(r->n_pixels) * kGuidePend);
        // the raw and kept counters restart every iteration; the ring's position and fill only
        // without history
        hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, s, r->gcount.as<uint32_t>(), 0u,
                           size_t(r->gcfg.history ? 2 : 4));
    }
    if (p.trace_camera) {
        launch_camera_generate(r->sc, p, b, s);
        for (int k = 0; k < L; k++) {
            hipEvent_t* e = ev + 8 + 4 * k;
            if (timed) HIPCHK(hipEventRecord(e[0], s));
            launch_camera_bounce(r->sc, p, b, k, s);
            if (timed) HIPCHK(hipEventRecord(e[1], s));
            launch_seg_merge(r->sc, p, b, 0, k, s);
            if (timed) HIPCHK(hipEventRecord(e[2], s));
            if (p.piece_order == 2 && p.build_bb1d && p.cap_beam > 0) {
                // Morton order of the pieces: count to the host, sort (key, index), gather
                uint32_t cnt = 0;
                HIPCHK(hipMemcpyAsync(&cnt, b.qcount + 320 + k, 4, hipMemcpyDeviceToHost, s));
                HIPCHK(hipStreamSynchronize(s));
                const uint32_t n = std::min(cnt, b.piece_cap);
                if (r->pkeys.ensure(size_t(b.piece_cap) * 4) || r->pvals.ensure(size_t(b.piece_cap) * 4))
                    return fail(r, "device allocation failed", UPBP_E_HIP);
                launch_piece_keys(b, k, n, r->pkeys.as<uint32_t>(), r->pvals.as<uint32_t>(), s);
                if (radix_sort_pairs(r->pkeys.as<uint32_t>(), r->pvals.as<uint32_t>(), n, 30, r->scratch, s))
                    return fail(r, "piece sort failed", UPBP_E_HIP);
                launch_piece_gather(b, n, r->pvals.as<uint32_t>(), s);
            }
            if (p.piece_order >= 3 && p.build_bb1d && p.cap_beam > 0) {
                // whole units of 64 pieces in spatial order, all on the device
                const uint32_t cu = (b.piece_cap + 63u) / 64u;
                uint32_t* keys = r->pkeys.as<uint32_t>();
                uint32_t* vals = r->pvals.as<uint32_t>();
                uint32_t* nu = r->unitn.as<uint32_t>();
                launch_unit_keys(b, k, cu, keys, vals, nu, s);
                if (radix_sort_pairs(keys, vals, cu, 30, r->scratch, s, nu))
                    return fail(r, "piece unit sort failed", UPBP_E_HIP);
                launch_unit_gather(b, k, cu, vals, p.piece_order == 3, s);
            }
            launch_seg_merge(r->sc, p, b, 1, k, s);
            if (timed) HIPCHK(hipEventRecord(e[3], s));
        }
    }
    launch_camera_finish(p, b, float(r->cfg.res_x), float(r->cfg.res_y), r->ovf.as<uint32_t>(), s);
    if (timed) HIPCHK(hipEventRecord(ev[3], s));
    HIPCHK(hipGetLastError());
    if (collect) { // train the proposal on this iteration's samples, then refresh its records
        const int e = train_guide(r);
        if (e) return e;
    }
    // the overflow flag comes back with the slot; settle() checks it once the slot's event passed
    HIPCHK(hipMemcpyAsync(&st.ovf, r->ovf.p, 4, hipMemcpyDeviceToHost, s));
    return UPBP_OK;
}

// Enqueues one iteration on r's stream. Nothing in it waits for the device, except a sizing
// pass of the photon lists, online training of the proposal and the statistics (stats != null).
int run_one(upbp_renderer* r, int iteration, upbp_iter_stats* stats)
{
    IterParams p;
    setup_params(r, iteration, p);
    hipStream_t s = r->stream;
    const int L = int(r->cfg.max_path_length);
    const bool timed = stats != nullptr; // per-kernel events only when statistics are asked for
    hipEvent_t* ev = r->ev;
    // a staging slot is reused only once the iteration that used it last has read it
    const int sl = r->slot;
    r->slot ^= 1;
    if (r->pending[sl] >= 0) {
        const int e = settle(r, true);
        if (e) return e;
    }
    Staging& st = r->stage[sl];
    st.p = p;
    st.ovf = 0;
    // Nothing in the body reads back to the host when no statistics are asked for, the photon
    // lists are sized, the proposal is not training and pieces are not Morton-sorted. Then the
    // body is captured once -- after one eager iteration with the same settings, which allocated
    // every buffer -- and replayed; the replay reads this iteration's parameters from the slot.
    const bool capturable = r->use_graph && !timed && !r->need_sizing && !r->gtrain && p.piece_order != 2 &&
                            p.trace_light && p.trace_camera;
    if (capturable && r->eager_gen == r->gen) {
        if (!r->gexec[sl] || r->gexec_gen[sl] != r->gen) {
            if (r->gexec[sl]) (void)hipGraphExecDestroy(r->gexec[sl]);
            r->gexec[sl] = nullptr;
            HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            const int e = enqueue_body(r, p, st, false);
            hipGraph_t g = nullptr;
            const hipError_t ce = hipStreamEndCapture(s, &g);
            if (e || ce != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return e ? e : fail(r, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce), UPBP_E_HIP);
            }
            const hipError_t ie = hipGraphInstantiate(&r->gexec[sl], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ie != hipSuccess) return fail(r, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie), UPBP_E_HIP);
            r->gexec_gen[sl] = r->gen;
        }
        HIPCHK(hipGraphLaunch(r->gexec[sl], s));
        r->graph_launches++;
    } else {
        const int e = enqueue_body(r, p, st, timed);
        if (e) return e;
        r->eager_gen = r->gen;
    }
    HIPCHK(hipEventRecord(r->stage_ev[sl], s));
    r->pending[sl] = iteration;
    if (stats) {
        const Bufs& b = r->bufs;
        HIPCHK(hipEventSynchronize(ev[3]));
        uint32_t overflowed = 0;
        HIPCHK(hipMemcpy(&overflowed, r->ovf.p, 4, hipMemcpyDeviceToHost));
        if (overflowed) { // not accumulated: resize the photon lists and render it again
            r->pending[sl] = -1;
            r->need_sizing = true;
            return run_one(r, iteration, stats);
        }
        memset(stats, 0, sizeof(*stats));
        uint32_t q[kQcountWords];
        unsigned long long st[STAT_COUNT];
        HIPCHK(hipMemcpy(q, b.qcount, sizeof(q), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(st, b.stats, sizeof(st), hipMemcpyDeviceToHost));
        for (int k = 0; k < L; k++) {
            stats->light_extend_steps += q[k];
            stats->camera_extend_steps += q[64 + k];
        }
        stats->ms_light = elapsed(ev[0], ev[1]);
        stats->ms_build = elapsed(ev[1], ev[2]);
        stats->ms_camera = elapsed(ev[2], ev[3]);
        stats->ms_total = elapsed(ev[0], ev[3]);
        if (p.trace_camera)
            for (int k = 0; k < L; k++) {
                const hipEvent_t* e = ev + 8 + 4 * k;
                stats->ms_bounce += elapsed(e[0], e[1]);
                stats->ms_merge_pb2d += elapsed(e[1], e[2]);
                stats->ms_merge_bb1d += elapsed(e[2], e[3]);
            }
        uint32_t counts[3] = {0, 0, 0};
        if (p.trace_light) HIPCHK(hipMemcpy(counts, r->mapn.p, 12, hipMemcpyDeviceToHost));
        stats->surface_vertices = counts[0];
        stats->medium_vertices = counts[1];
        stats->light_vertices = uint64_t(counts[0]) + counts[1];
        stats->subbeams = counts[2];
        stats->beams = 0;
        {
            std::vector<uint32_t> nbeam(r->want_beams ? r->n_light : 0);
            if (!nbeam.empty()) {
                HIPCHK(hipMemcpy(nbeam.data(), b.nbeam, nbeam.size() * 4, hipMemcpyDeviceToHost));
                for (uint32_t v : nbeam) stats->beams += v;
            }
        }
        stats->bpt_connections = st[STAT_BPT_CONN];
        stats->point_queries = st[STAT_POINT_Q];
        stats->segment_queries = st[STAT_SEG_Q];
        stats->merge_accepts = st[STAT_ACCEPT];
        stats->pb2d_accepts = st[STAT_ACC_PB2D];
        stats->bb1d_accepts = st[STAT_ACC_BB1D];
        stats->seg_nodes = st[STAT_SEG_NODES];
        stats->seg_candidates = st[STAT_SEG_CANDS];
        stats->seg_pieces = st[STAT_SEG_PIECES];
        stats->seg_wave_trips = st[STAT_SEG_TRIPS];
        stats->seg_cand_batches = st[STAT_CAND_BATCHES];
        stats->seg_pair_batches = st[STAT_PAIR_BATCHES];
        stats->vertex_drops = st[STAT_VDROP];
        stats->beam_drops = st[STAT_BDROP];
        stats->path_errors = st[STAT_ERRORS];
        stats->merge_stack_overflows = st[STAT_MERGE_STACK];
        stats->bundles = st[STAT_BUNDLES];
        stats->bundle_pieces_left = st[STAT_BUNDLE_LEFT];
        stats->bundle_nodes = st[STAT_BUNDLE_NODES];
        stats->bundle_leaves = st[STAT_BUNDLE_LEAVES];
        stats->bundle_tests = st[STAT_BUNDLE_TESTS];
        stats->photon_list_resizes = r->resizes;
        stats->graph_launches = r->graph_launches;
        // Compulsory HBM bytes. Camera bounce kernels: 128-byte state read and write plus 64
        // bytes of stack per live path-step, 112 bytes per light vertex read by a connection or
        // an accepted point merge, 112 bytes per segment query written. Segment merge kernels:
        // the 112-byte query read once per kernel, 112 bytes per accepted medium point (P-B2D),
        // 128 bytes per accepted beam (B-B1D).
        const double point_acc = double(st[STAT_ACCEPT] - st[STAT_ACC_PB2D] - st[STAT_ACC_BB1D]);
        const double nq = double(st[STAT_SEG_Q]);
        const double b_bounce = 320.0 * double(stats->camera_extend_steps) + 112.0 * double(st[STAT_BPT_CONN]) +
                                112.0 * point_acc + 112.0 * nq;
        const double b_pb2d = (p.build_pb2d ? 112.0 * nq : 0.0) + 112.0 * double(st[STAT_ACC_PB2D]);
        const double b_bb1d = (p.build_bb1d ? 112.0 * nq : 0.0) + 128.0 * double(st[STAT_ACC_BB1D]);
        struct Group {
            const char* name;
            double ms, bytes;
        } g[3] = {{"k_camera_bounce", stats->ms_bounce, b_bounce},
                  {"k_seg_merge<0> (P-B2D)", stats->ms_merge_pb2d, b_pb2d},
                  {"B-B1D merge (bundle+piece)", stats->ms_merge_bb1d, b_bb1d}};
        int dom = 0;
        for (int k = 1; k < 3; k++)
            if (g[k].ms > g[dom].ms) dom = k;
        stats->ms_dominant = g[dom].ms;
        stats->dominant_bytes = g[dom].bytes;
        stats->dominant_launches = p.trace_camera ? uint64_t(L) : 0;
        snprintf(stats->dominant_kernel, sizeof(stats->dominant_kernel), "%s", g[dom].name);
        stats->guide_samples = r->gsamples;
        stats->guide_loss = r->gloss;
        stats->ms_guide = r->gms;
        stats->algorithmic_bytes = b_bounce + b_pb2d + b_bb1d + 320.0 * double(stats->light_extend_steps) +
                                   112.0 * double(counts[0] + counts[1]) + 128.0 * double(stats->beams) +
                                   88.0 * double(counts[0] + counts[1] + counts[2]) + 60.0 * double(r->n_pixels);
    }
    return UPBP_OK;
}

// Checks the iterations whose overflow flags are back (all of them, waiting, when `wait`): one
// that overflowed a photon list was not accumulated, so the lists are resized from its totals and
// it is rendered again.
int settle(upbp_renderer* r, bool wait)
{
    for (int k = 0; k < 2; k++) {
        const int it = r->pending[k];
        if (it < 0) continue;
        if (wait) HIPCHK(hipEventSynchronize(r->stage_ev[k]));
        else if (hipEventQuery(r->stage_ev[k]) != hipSuccess) continue;
        r->pending[k] = -1;
        if (r->stage[k].ovf) {
            r->need_sizing = true;
            const int e = run_one(r, it, nullptr);
            if (e) return e;
        }
    }
    return UPBP_OK;
}

} // namespace

extern "C" int upbp_run_iteration(upbp_renderer* r, int iteration, upbp_iter_stats* stats)
{
    if (!r) return fail(r, "null renderer", UPBP_E_ARG);
    const int e = settle(r, false);
    return e ? e : run_one(r, iteration, stats);
}

// waits for every iteration enqueued so far and renders again any that overflowed a photon list
static int settle_all(upbp_renderer* r)
{
    while (r->pending[0] >= 0 || r->pending[1] >= 0) {
        const int e = settle(r, true);
        if (e) return e;
    }
    return UPBP_OK;
}

extern "C" int upbp_read_accum(upbp_renderer* r, double* rgb, int iters_done)
{
    if (!r || !rgb || iters_done < 1) return fail(r, "invalid argument", UPBP_E_ARG);
    const int e = settle_all(r);
    if (e) return e;
    HIPCHK(hipStreamSynchronize(r->stream));
    const size_t n = 3 * size_t(r->n_pixels);
    HIPCHK(hipMemcpy(rgb, r->bufs.accum, n * sizeof(double), hipMemcpyDeviceToHost));
    const double inv = 1.0 / double(iters_done);
    for (size_t i = 0; i < n; i++) rgb[i] *= inv;
    return UPBP_OK;
}

extern "C" void upbp_tuning_defaults(upbp_tuning* t)
{
    if (!t) return;
    memset(t, 0, sizeof(*t));
    t->subbeam_factor = 12.f;
    t->query_piece_factor = 64.f;
    t->merge_blocks_per_cu = 0;
    t->packet_bounces = 0;
    t->graph = 0;
    t->bundle_bounces = 1;
    t->bundle_factor = 32.f;
    t->bundle_piece_factor = 0.f;
}

extern "C" int upbp_set_guide_lobes(upbp_renderer* r, const float* lobes, int32_t res, const float* lo, const float* hi,
                                    float alpha)
{
    if (!r) return UPBP_E_ARG;
    if (!(alpha >= 0.f && alpha <= 1.f)) return fail(r, "guide alpha must lie in [0, 1]", UPBP_E_ARG);
    r->gen++;
    if (alpha == 0.f) {
        r->sc.guide.alpha = r->sc.guide.alpha_s = 0.f;
        return UPBP_OK;
    }
    if (!lobes || !lo || !hi || res < 1 || res > 256) return fail(r, "invalid guide grid", UPBP_E_ARG);
    for (int a = 0; a < 3; a++)
        if (!(hi[a] > lo[a])) return fail(r, "empty guide bounds", UPBP_E_ARG);
    const size_t n = size_t(res) * res * res;
    std::vector<float4> recs(n * kGuideRec);
    for (size_t i = 0; i < n; i++) pack_guide_record(lobes + i * kGuideLobes * 5, &recs[i * kGuideRec]);
    if (r->guide.ensure(recs.size() * sizeof(float4)) ||
        hipMemcpy(r->guide.p, recs.data(), recs.size() * sizeof(float4), hipMemcpyHostToDevice) != hipSuccess)
        return fail(r, "guide upload failed", UPBP_E_HIP);
    DGuide& g = r->sc.guide;
    g.recs = r->guide.as<float4>();
    g.res = res;
    g.lo = F3(lo);
    g.inv_cell = mk(float(res) / (hi[0] - lo[0]), float(res) / (hi[1] - lo[1]), float(res) / (hi[2] - lo[2]));
    g.alpha = g.alpha_s = alpha;
    r->gen++;
    return UPBP_OK;
}

extern "C" void upbp_guide_defaults(upbp_guide_config* g)
{
    if (!g) return;
    memset(g, 0, sizeof(*g));
    g->alpha = 0.f;
    g->grid_res = 64;
    g->train = 1;
    g->steps_per_iteration = 10;
    g->batch = 65536;
    g->learning_rate = 8e-3f;
    g->seed = 3;
    g->max_samples = 1u << 24;
    g->surface_alpha = 0.65f;
    g->history = 1u << 22;
}

extern "C" int upbp_set_guide(upbp_renderer* r, const upbp_guide_config* g)
{
    if (!r || !g) return UPBP_E_ARG;
    if (!(g->alpha >= 0.f && g->alpha <= 1.f) || !(g->surface_alpha >= 0.f && g->surface_alpha <= 1.f) ||
        g->grid_res < 1 || g->grid_res > 64 || g->steps_per_iteration < 0 ||
        g->batch < 16 || g->batch > (1 << 22) || !(g->learning_rate > 0.f) || g->max_samples < 16 ||
        (g->history != 0 && g->history < 16))
        return fail(r, "invalid guide configuration", UPBP_E_ARG);
    r->gcfg = *g;
    r->gen++;
    delete r->gnet;
    r->gnet = nullptr;
    r->gtrain = false;
    r->sc.guide.alpha = r->sc.guide.alpha_s = 0.f;
    if (g->alpha == 0.f && g->surface_alpha == 0.f) return UPBP_OK;
    // grid over the bounding cube of the scene sphere
    const f3 c = r->sc.center;
    const float rad = r->sc.radius;
    const int res = g->grid_res;
    const size_t nrec = size_t(res) * res * res;
    if (r->guide.ensure(nrec * kGuideRec * sizeof(float4))) return fail(r, "guide allocation failed", UPBP_E_HIP);
    DGuide& dg = r->sc.guide;
    dg.recs = r->guide.as<float4>();
    dg.res = res;
    dg.lo = c - mk(rad);
    dg.inv_cell = mk(float(res) / (2.f * rad));
    r->gnet = new GuideNet();
    if (r->gnet->init(g->seed) || r->gnet->records(res, r->guide.as<float4>(), r->stream))
        return fail(r, "guide network setup failed", UPBP_E_HIP);
    r->gcap = g->max_samples;
    r->gring = g->history ? g->history : g->max_samples;
    if (r->gpend.ensure(size_t(r->n_pixels) * kGuidePend * 4) || r->graw.ensure(size_t(r->gcap) * kGuideRaw * 4) ||
        r->gsamp.ensure(size_t(r->gring) * 10 * 4) || r->gcount.ensure(16))
        return fail(r, "guide allocation failed", UPBP_E_HIP);
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, r->stream, r->gcount.as<uint32_t>(), 0u, size_t(4));
    r->gtrain = g->train != 0;
    // samples only from the scatterings whose proposal is on
    dg.train_media = g->alpha > 0.f;
    dg.train_surfaces = g->surface_alpha > 0.f;
    // an untrained network proposes nearly uniform lobes: the first iteration samples the phase
    // function only and collects training data; the mixture switches on after the first steps
    if (!r->gtrain) {
        dg.alpha = g->alpha;
        dg.alpha_s = g->surface_alpha;
    }
    HIPCHK(hipStreamSynchronize(r->stream));
    return UPBP_OK;
}

extern "C" int upbp_set_tuning(upbp_renderer* r, const upbp_tuning* t)
{
    if (!r || !t) return fail(r, "null argument", UPBP_E_ARG);
    if (!(t->subbeam_factor >= 0.f) || !(t->query_piece_factor >= 0.f) || t->merge_blocks_per_cu < 0 ||
        t->piece_order < 0 || t->piece_order > 4 || t->packet_bounces < 0 || t->beam_leaf < 0 || t->beam_leaf > 1 ||
        (t->block_size != 0 && t->block_size != 64 && t->block_size != 128 && t->block_size != 256) ||
        t->bundle_bounces < 0 || !(t->bundle_factor >= 0.f) || !(t->bundle_piece_factor >= 0.f))
        return fail(r, "invalid tuning", UPBP_E_ARG);
    r->tune = *t;
    r->use_graph = t->graph != 0;
    r->gen++;
    return UPBP_OK;
}

// Coordinate search over the tuning grid: each knob in turn takes the value that renders
// `iteration` fastest with the others held, starting from the current tuning. A candidate is
// timed by host wall clock over `repeats` iterations after one untimed run (which also grows
// buffers a new setting needs). The knobs never change which pairs a query accepts, so the
// trial iterations render the same image; the accumulator is saved before and restored after
// them, and online training of the learned proposal is paused meanwhile.
extern "C" int upbp_autotune(upbp_renderer* r, int32_t iteration, int32_t repeats, upbp_tuning* chosen,
                             float* ms_per_iteration)
{
    if (!r) return fail(r, "null renderer", UPBP_E_ARG);
    if (repeats < 1 || repeats > 64 || iteration < 0) return fail(r, "invalid autotune arguments", UPBP_E_ARG);
    const size_t acc_bytes = size_t(r->n_pixels) * 3 * sizeof(double);
    DevBuf saved;
    struct Release {
        DevBuf& b;
        ~Release() { b.release(); }
    } release_saved{saved};
    if (saved.ensure(acc_bytes)) return fail(r, "autotune allocation failed", UPBP_E_HIP);
    if (const int es = settle_all(r)) return es;
    HIPCHK(hipMemcpyAsync(saved.p, r->accum.p, acc_bytes, hipMemcpyDeviceToDevice, r->stream));
    const bool train = r->gtrain;
    r->gtrain = false;
    const upbp_tuning start = r->tune;
    auto timed = [&](const upbp_tuning& t, double& ms) -> int {
        r->tune = t;
        for (int k = 0; k <= repeats; k++) {
            HIPCHK(hipStreamSynchronize(r->stream));
            const auto t0 = std::chrono::steady_clock::now();
            const int e = upbp_run_iteration(r, iteration, nullptr);
            if (e) return e;
            HIPCHK(hipStreamSynchronize(r->stream));
            const double dt = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            if (k == 0) ms = 0.0;
            else ms += dt / double(repeats);
        }
        return UPBP_OK;
    };
    upbp_tuning best = start;
    double best_ms = 0.0;
    int e = timed(best, best_ms);
    const float subbeam[] = {8.f, 12.f, 16.f, 24.f}, piece[] = {32.f, 64.f, 128.f};
    const int blocks[] = {0, 1, 2}, order[] = {0, 2}, packets[] = {0, 1, 2, 64}, threads[] = {64, 128, 256},
              bundles[] = {0, 1, 2};
    const int counts[] = {4, 3, 3, 2, 4, 3, 3};
    for (int knob = 0; knob < 7 && !e; knob++) {
        const int n = counts[knob];
        const upbp_tuning base = best;
        for (int v = 0; v < n && !e; v++) {
            upbp_tuning t = base;
            if (knob == 0) t.subbeam_factor = subbeam[v];
            else if (knob == 1) t.query_piece_factor = piece[v];
            else if (knob == 2) t.merge_blocks_per_cu = blocks[v];
            else if (knob == 3) t.piece_order = order[v];
            else if (knob == 4) t.packet_bounces = packets[v];
            else if (knob == 5) t.block_size = threads[v];
            else t.bundle_bounces = bundles[v];
            if (memcmp(&t, &base, sizeof(t)) == 0) continue;
            double ms = 0.0;
            e = timed(t, ms);
            if (!e && ms < best_ms) {
                best = t;
                best_ms = ms;
            }
        }
    }
    if (!e) e = settle_all(r);
    r->tune = e ? start : best;
    r->gtrain = train;
    if (hipMemcpyAsync(r->accum.p, saved.p, acc_bytes, hipMemcpyDeviceToDevice, r->stream) != hipSuccess ||
        hipStreamSynchronize(r->stream) != hipSuccess)
        return fail(r, "autotune restore failed", UPBP_E_HIP);
    if (e) return e;
    if (chosen) *chosen = best;
    if (ms_per_iteration) *ms_per_iteration = float(best_ms);
    return UPBP_OK;
}

// One iteration of each of n independent renderers at once. Each renderer owns its stream, so
// host threads that drive different renderers put their kernels on the GPU side by side:
// small jobs, which fill only part of the chip on their own, overlap. Threads take the next
// job from a shared counter; the first error (lowest job index) is returned.
extern "C" int upbp_run_batch(upbp_renderer* const* rs, const int32_t* iterations, int32_t n, int32_t threads)
{
    if (n < 0 || (n > 0 && (!rs || !iterations)) || threads < 1) return fail(nullptr, "invalid batch", UPBP_E_ARG);
    for (int32_t i = 0; i < n; i++)
        if (!rs[i]) return fail(nullptr, "null renderer in batch", UPBP_E_ARG);
    { // two threads on one renderer would share its buffers and stream: reject duplicates
        std::vector<upbp_renderer*> sorted(rs, rs + n);
        std::sort(sorted.begin(), sorted.end());
        if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
            return fail(nullptr, "a renderer appears twice in the batch", UPBP_E_ARG);
    }
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) return fail(nullptr, "no HIP device", UPBP_E_HIP);
    std::atomic<int32_t> next{0};
    std::vector<int> rc(size_t(n), UPBP_OK);
    auto work = [&]() {
        if (hipSetDevice(device) != hipSuccess) return;
        for (int32_t i = next++; i < n; i = next++) {
            rc[size_t(i)] = upbp_run_iteration(rs[i], iterations[i], nullptr);
            if (!rc[size_t(i)]) rc[size_t(i)] = upbp_sync(rs[i]);
        }
    };
    const int32_t t = std::min(threads, std::max<int32_t>(n, 1));
    std::vector<std::thread> pool;
    for (int32_t k = 1; k < t; k++) pool.emplace_back(work);
    work();
    for (std::thread& th : pool) th.join();
    for (int32_t i = 0; i < n; i++)
        if (rc[size_t(i)]) return rc[size_t(i)];
    return UPBP_OK;
}

// ---- batched jobs in one launch per stage -----------------------------------------------------
// A round renders one iteration of every job. The wavefront stages -- light generate, each light
// bounce, camera generate, each camera bounce with its two segment merges, splat and accumulate --
// run as one launch each over all jobs: the block maps give each block its job and its block in
// that job's own grid, and the persistent merge kernels walk every job's pieces. The per-job work
// between them (parameters and counters, photon lists and their LBVH builds) is enqueued on each
// job's own stream, forked from the batch stream and joined back, so the jobs' builds overlap. The
// round is captured in a HIP graph once (after one eager round with the same settings) and replayed.
struct upbp_batch {
    std::string err;
    std::vector<upbp_renderer*> rs;
    hipStream_t stream = nullptr;
    hipEvent_t fork[2] = {nullptr, nullptr};
    std::vector<hipEvent_t> join; // 2 per job
    std::vector<BatchJob> hjobs;
    DevBuf jobs, map_light, map_camgen, map_cam, map_pix, map_acc;
    // light and camera bounce maps list the jobs by decreasing max path length: bounce k runs the
    // first light_upto[k] / cam_upto[k] blocks, those of the jobs whose paths reach bounce k
    std::vector<uint32_t> light_upto, cam_upto;
    uint32_t nb_light = 0, nb_camgen = 0, nb_pix = 0, nb_acc = 0;
    int lmax = 0;
    hipGraphExec_t gexec = nullptr;
    std::vector<uint64_t> key_graph, key_eager; // the jobs' settings generations at capture / last eager round
    uint64_t rounds = 0, graph_rounds = 0, eager_rounds = 0, solo_rounds = 0;
};

namespace {

int bfail(upbp_batch* bt, const std::string& m, int code)
{
    if (bt) bt->err = m;
    g_err = m;
    return code;
}

#define BHIPCHK(expr)                                                                                                   \
    do {                                                                                                                \
        hipError_t e_ = (expr);                                                                                         \
        if (e_ != hipSuccess) return bfail(bt, std::string(#expr) + ": " + hipGetErrorString(e_), UPBP_E_HIP);          \
    } while (0)

// the work of one round on the batch stream (eager or under capture)
int batch_enqueue(upbp_batch* bt, std::vector<IterParams>& ps)
{
    hipStream_t S = bt->stream;
    const uint32_t n = uint32_t(bt->rs.size());
    const BatchJob* jobs = bt->jobs.as<BatchJob>();
    BHIPCHK(hipEventRecord(bt->fork[0], S));
    for (uint32_t j = 0; j < n; j++) {
        upbp_renderer* r = bt->rs[j];
        BHIPCHK(hipStreamWaitEvent(r->stream, bt->fork[0], 0));
        if (enqueue_head(r, r->stage[0])) return bfail(bt, "job " + std::to_string(j) + ": " + r->err, UPBP_E_HIP);
        BHIPCHK(hipEventRecord(bt->join[2 * j], r->stream));
        BHIPCHK(hipStreamWaitEvent(S, bt->join[2 * j], 0));
    }
    launch_light_generate_batch(jobs, bt->map_light.as<uint2>(), bt->nb_light, S);
    for (int k = 0; k < bt->lmax; k++) launch_light_bounce_batch(jobs, bt->map_light.as<uint2>(), bt->light_upto[k], k, S);
    BHIPCHK(hipEventRecord(bt->fork[1], S));
    for (uint32_t j = 0; j < n; j++) {
        upbp_renderer* r = bt->rs[j];
        BHIPCHK(hipStreamWaitEvent(r->stream, bt->fork[1], 0));
        if (enqueue_maps(r, ps[j], r->stage[0])) return bfail(bt, "job " + std::to_string(j) + ": " + r->err, UPBP_E_HIP);
        BHIPCHK(hipEventRecord(bt->join[2 * j + 1], r->stream));
        BHIPCHK(hipStreamWaitEvent(S, bt->join[2 * j + 1], 0));
    }
    launch_camera_generate_batch(jobs, bt->map_camgen.as<uint2>(), bt->nb_camgen, S);
    for (int k = 0; k < bt->lmax; k++) {
        launch_camera_bounce_batch(jobs, bt->map_cam.as<uint2>(), bt->cam_upto[k], k, S);
        launch_seg_merge_batch(jobs, n, 0, k, S);
        launch_seg_merge_batch(jobs, n, 1, k, S);
    }
    launch_camera_finish_batch(jobs, bt->map_pix.as<uint2>(), bt->nb_pix, bt->map_acc.as<uint2>(), bt->nb_acc, S);
    for (uint32_t j = 0; j < n; j++) {
        upbp_renderer* r = bt->rs[j];
        BHIPCHK(hipMemcpyAsync(&r->stage[0].ovf, r->ovf.p, 4, hipMemcpyDeviceToHost, S));
    }
    BHIPCHK(hipGetLastError());
    return UPBP_OK;
}

} // namespace

extern "C" int upbp_batch_create(upbp_renderer* const* rs, int32_t n, upbp_batch** out)
{
    if (!out || n < 1 || !rs) return bfail(nullptr, "invalid batch", UPBP_E_ARG);
    *out = nullptr;
    for (int32_t i = 0; i < n; i++)
        if (!rs[i]) return bfail(nullptr, "null renderer in batch", UPBP_E_ARG);
    {
        std::vector<upbp_renderer*> sorted(rs, rs + n);
        std::sort(sorted.begin(), sorted.end());
        if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
            return bfail(nullptr, "a renderer appears twice in the batch", UPBP_E_ARG);
    }
    auto* bt = new upbp_batch();
    bt->rs.assign(rs, rs + n);
    auto cleanup = [&](int code) {
        upbp_batch_destroy(bt);
        return code;
    };
    if (hipStreamCreateWithFlags(&bt->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&bt->fork[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&bt->fork[1], hipEventDisableTiming) != hipSuccess)
        return cleanup(bfail(nullptr, "batch stream or event creation failed", UPBP_E_HIP));
    bt->join.assign(size_t(2 * n), nullptr);
    for (hipEvent_t& e : bt->join)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
            return cleanup(bfail(nullptr, "batch event creation failed", UPBP_E_HIP));
    // block maps; the bounce maps by decreasing max path length
    std::vector<uint32_t> order(static_cast<size_t>(n));
    for (int32_t i = 0; i < n; i++) order[size_t(i)] = uint32_t(i);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        return rs[a]->cfg.max_path_length > rs[b]->cfg.max_path_length;
    });
    std::vector<uint2> ml, mg, mc, mp, ma;
    for (uint32_t j : order) {
        const upbp_renderer* r = rs[j];
        for (uint32_t k = 0; k < (r->n_light + 255) / 256; k++) ml.push_back(make_uint2(j, k));
        for (uint32_t k = 0; k < (r->n_pixels + 255) / 256; k++) mc.push_back(make_uint2(j, k));
        bt->lmax = std::max(bt->lmax, int(r->cfg.max_path_length));
    }
    bt->light_upto.assign(size_t(bt->lmax), 0u);
    bt->cam_upto.assign(size_t(bt->lmax), 0u);
    {
        uint32_t al = 0, ac = 0;
        for (uint32_t j : order) {
            al += (rs[j]->n_light + 255) / 256;
            ac += (rs[j]->n_pixels + 255) / 256;
            for (int k = 0; k < int(rs[j]->cfg.max_path_length); k++) {
                bt->light_upto[size_t(k)] = al;
                bt->cam_upto[size_t(k)] = ac;
            }
        }
    }
    bt->hjobs.resize(size_t(n));
    for (int32_t j = 0; j < n; j++) {
        const upbp_renderer* r = rs[j];
        const uint32_t res_x = uint32_t(r->sc.cam.res_x);
        const uint32_t res_y = (r->n_pixels + res_x - 1) / res_x;
        const uint32_t tiles_x = (res_x + 15) / 16, tiles_y = (res_y + 15) / 16;
        for (uint32_t k = 0; k < tiles_x * tiles_y; k++) mg.push_back(make_uint2(uint32_t(j), k));
        for (uint32_t k = 0; k < (r->n_pixels + 255) / 256; k++) mp.push_back(make_uint2(uint32_t(j), k));
        for (uint32_t k = 0; k < (3 * r->n_pixels + 255) / 256; k++) ma.push_back(make_uint2(uint32_t(j), k));
        BatchJob& J = bt->hjobs[size_t(j)];
        J.res_x = res_x;
        J.tiles_x = tiles_x;
        J.fres_x = float(r->cfg.res_x);
        J.fres_y = float(r->cfg.res_y);
    }
    bt->nb_light = uint32_t(ml.size());
    bt->nb_camgen = uint32_t(mg.size());
    bt->nb_pix = uint32_t(mp.size());
    bt->nb_acc = uint32_t(ma.size());
    struct M {
        DevBuf* buf;
        const std::vector<uint2>* v;
    };
    for (const M& m : {M{&bt->map_light, &ml}, M{&bt->map_camgen, &mg}, M{&bt->map_cam, &mc}, M{&bt->map_pix, &mp},
                       M{&bt->map_acc, &ma}})
        if (m.buf->ensure(std::max<size_t>(m.v->size(), 1) * sizeof(uint2)) ||
            (!m.v->empty() &&
             hipMemcpy(m.buf->p, m.v->data(), m.v->size() * sizeof(uint2), hipMemcpyHostToDevice) != hipSuccess))
            return cleanup(bfail(nullptr, "batch block maps: device allocation failed", UPBP_E_HIP));
    if (bt->jobs.ensure(size_t(n) * sizeof(BatchJob)))
        return cleanup(bfail(nullptr, "batch jobs: device allocation failed", UPBP_E_HIP));
    *out = bt;
    return UPBP_OK;
}

extern "C" int upbp_batch_run(upbp_batch* bt, const int32_t* iterations, int32_t use_graph)
{
    if (!bt || !iterations) return bfail(bt, "invalid argument", UPBP_E_ARG);
    const uint32_t n = uint32_t(bt->rs.size());
    // iterations still pending on a job's own stream come first (they may render again)
    for (upbp_renderer* r : bt->rs) {
        const int e = settle_all(r);
        if (e) return bfail(bt, r->err, e);
        BHIPCHK(hipStreamSynchronize(r->stream));
    }
    std::vector<IterParams> ps(n);
    bool batched = true;
    for (uint32_t j = 0; j < n; j++) {
        upbp_renderer* r = bt->rs[j];
        setup_params(r, iterations[j], ps[j]);
        // a round in one launch per stage needs every job sized, plain (no training, no MIS log)
        // and tracing both subpaths
        if (r->need_sizing || r->gtrain || r->bufs.mislog || !ps[j].trace_light || !ps[j].trace_camera)
            batched = false;
    }
    bt->rounds++;
    if (!batched) { // each job alone on its own stream (a sizing pass, training, a special mode)
        for (uint32_t j = 0; j < n; j++) {
            upbp_renderer* r = bt->rs[j];
            int e = upbp_run_iteration(r, iterations[j], nullptr);
            if (!e) e = settle_all(r);
            if (e) return bfail(bt, "job " + std::to_string(j) + ": " + r->err, e);
        }
        bt->solo_rounds++;
        return UPBP_OK;
    }
    std::vector<uint64_t> key(n);
    for (uint32_t j = 0; j < n; j++) {
        upbp_renderer* r = bt->rs[j];
        Staging& st = r->stage[0];
        st.p = ps[j];
        st.ovf = 0;
        BatchJob& J = bt->hjobs[j];
        J.sc = r->sc;
        J.b = r->bufs;
        J.b.gpend = J.b.graw = J.b.gsamp = nullptr; // no training samples are collected in a batch
        J.b.gcount = nullptr;
        J.b.gcap = 0u;
        J.b.gring = 1u;
        J.ovf = r->ovf.as<uint32_t>();
        key[j] = r->gen;
    }
    BHIPCHK(hipMemcpy(bt->jobs.p, bt->hjobs.data(), n * sizeof(BatchJob), hipMemcpyHostToDevice));
    const bool capture = use_graph && key == bt->key_eager;
    if (capture && (!bt->gexec || key != bt->key_graph)) {
        if (bt->gexec) (void)hipGraphExecDestroy(bt->gexec);
        bt->gexec = nullptr;
        BHIPCHK(hipStreamBeginCapture(bt->stream, hipStreamCaptureModeThreadLocal));
        const int e = batch_enqueue(bt, ps);
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(bt->stream, &g);
        if (e || ce != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return e ? e : bfail(bt, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce), UPBP_E_HIP);
        }
        const hipError_t ie = hipGraphInstantiate(&bt->gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess) return bfail(bt, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie), UPBP_E_HIP);
        bt->key_graph = key;
    }
    if (capture) {
        BHIPCHK(hipGraphLaunch(bt->gexec, bt->stream));
        bt->graph_rounds++;
    } else {
        const int e = batch_enqueue(bt, ps);
        if (e) return e;
        bt->key_eager = key;
        bt->eager_rounds++;
    }
    BHIPCHK(hipStreamSynchronize(bt->stream));
    // a job whose photon lists overflowed was not accumulated: it sizes them and renders again alone
    for (uint32_t j = 0; j < n; j++) {
        upbp_renderer* r = bt->rs[j];
        if (!r->stage[0].ovf) continue;
        r->need_sizing = true;
        int e = run_one(r, iterations[j], nullptr);
        if (!e) e = settle_all(r);
        if (e) return bfail(bt, "job " + std::to_string(j) + ": " + r->err, e);
    }
    return UPBP_OK;
}

extern "C" int upbp_batch_counts(const upbp_batch* bt, uint64_t* rounds, uint64_t* graph_rounds,
                                 uint64_t* eager_rounds, uint64_t* solo_rounds)
{
    if (!bt) return bfail(nullptr, "null batch", UPBP_E_ARG);
    if (rounds) *rounds = bt->rounds;
    if (graph_rounds) *graph_rounds = bt->graph_rounds;
    if (eager_rounds) *eager_rounds = bt->eager_rounds;
    if (solo_rounds) *solo_rounds = bt->solo_rounds;
    return UPBP_OK;
}

extern "C" const char* upbp_batch_last_error(const upbp_batch* bt) { return bt ? bt->err.c_str() : g_err.c_str(); }

extern "C" void upbp_batch_destroy(upbp_batch* bt)
{
    if (!bt) return;
    if (bt->stream) (void)hipStreamSynchronize(bt->stream);
    if (bt->gexec) (void)hipGraphExecDestroy(bt->gexec);
    for (hipEvent_t e : bt->fork)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : bt->join)
        if (e) (void)hipEventDestroy(e);
    for (DevBuf* d : {&bt->jobs, &bt->map_light, &bt->map_camgen, &bt->map_cam, &bt->map_pix, &bt->map_acc}) d->release();
    if (bt->stream) (void)hipStreamDestroy(bt->stream);
    delete bt;
}

extern "C" int upbp_set_mis_log(upbp_renderer* r, uint32_t capacity, uint32_t pixel_stride, uint32_t light_paths)
{
    if (!r) return fail(r, "null renderer", UPBP_E_ARG);
    Bufs& b = r->bufs;
    if (capacity == 0) {
        HIPCHK(hipStreamSynchronize(r->stream));
        r->mislog.release();
        r->miscount.release();
        b.mislog = nullptr;
        b.miscount = nullptr;
        b.miscap = 0;
        r->gen++;
        return UPBP_OK;
    }
    if (pixel_stride == 0 || capacity > (1u << 28)) return fail(r, "invalid MIS log arguments", UPBP_E_ARG);
    if (r->mislog.ensure(size_t(capacity) * sizeof(uint4)) || r->miscount.ensure(4))
        return fail(r, "MIS log allocation failed", UPBP_E_HIP);
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, r->stream, r->miscount.as<uint32_t>(), 0u, size_t(1));
    HIPCHK(hipStreamSynchronize(r->stream));
    b.mislog = r->mislog.as<uint4>();
    b.miscount = r->miscount.as<uint32_t>();
    b.miscap = capacity;
    b.mis_stride = pixel_stride;
    b.mis_lt = light_paths;
    r->gen++;
    return UPBP_OK;
}

extern "C" int upbp_read_mis_log(upbp_renderer* r, uint32_t* records, uint64_t max_records, uint64_t* n_records)
{
    if (!r || !n_records || (max_records && !records)) return fail(r, "null argument", UPBP_E_ARG);
    const Bufs& b = r->bufs;
    if (!b.mislog) return fail(r, "the MIS log is off", UPBP_E_ARG);
    uint32_t n = 0;
    HIPCHK(hipStreamSynchronize(r->stream));
    HIPCHK(hipMemcpy(&n, b.miscount, 4, hipMemcpyDeviceToHost));
    *n_records = n;
    const uint64_t m = std::min<uint64_t>(std::min<uint64_t>(n, b.miscap), max_records);
    if (m) HIPCHK(hipMemcpy(records, b.mislog, m * sizeof(uint4), hipMemcpyDeviceToHost));
    if (n > b.miscap) return fail(r, "MIS log overflow: " + std::to_string(n) + " records", UPBP_E_CAPACITY);
    return UPBP_OK;
}

extern "C" int upbp_read_guide_samples(upbp_renderer* r, float* samples, uint64_t max_samples, uint64_t* n_samples,
                                       float* grid)
{
    if (!r || !n_samples || (max_samples && !samples)) return fail(r, "null argument", UPBP_E_ARG);
    if (!r->gtrain || !r->gsamp.p) return fail(r, "online training is off", UPBP_E_ARG);
    uint32_t n = 0;
    HIPCHK(hipStreamSynchronize(r->stream));
    HIPCHK(hipMemcpy(&n, r->gcount.as<uint32_t>() + 3, 4, hipMemcpyDeviceToHost)); // the ring's fill
    *n_samples = n;
    const uint64_t m = std::min<uint64_t>(std::min<uint64_t>(n, r->gring), max_samples);
    if (m) HIPCHK(hipMemcpy(samples, r->gsamp.p, m * 10 * sizeof(float), hipMemcpyDeviceToHost));
    if (grid) {
        const DGuide& g = r->sc.guide;
        grid[0] = g.lo.x;
        grid[1] = g.lo.y;
        grid[2] = g.lo.z;
        grid[3] = float(g.res) / g.inv_cell.x;
    }
    return UPBP_OK;
}

extern "C" int upbp_memory_info(uint64_t* free_bytes, uint64_t* total_bytes)
{
    if (!free_bytes || !total_bytes) return UPBP_E_ARG;
    size_t f = 0, t = 0;
    if (hipMemGetInfo(&f, &t) != hipSuccess) return UPBP_E_HIP;
    *free_bytes = f;
    *total_bytes = t;
    return UPBP_OK;
}

extern "C" int upbp_sync(upbp_renderer* r)
{
    if (!r) return fail(r, "null renderer", UPBP_E_ARG);
    const int e = settle_all(r);
    if (e) return e;
    HIPCHK(hipStreamSynchronize(r->stream));
    return UPBP_OK;
}

extern "C" void upbp_destroy(upbp_renderer* r)
{
    if (!r) return;
    if (r->stream) (void)hipStreamSynchronize(r->stream);
    for (DevBuf& d : r->owned) d.release();
    for (auto& g : r->gexec)
        if (g) (void)hipGraphExecDestroy(g);
    for (TriStore* t : {&r->real, &r->imag}) {
        if (t->tri) (void)hipFree(t->tri);
        if (t->nrm) (void)hipFree(t->nrm);
        if (t->rec) (void)hipFree(t->rec);
        lbvh_free(t->bvh);
    }
    for (DevBuf* d : {&r->lstate, &r->lstack, &r->cstate, &r->cstack, &r->q0, &r->q1, &r->qcount, &r->verts, &r->nvert,
                      &r->beams, &r->nbeam, &r->fb32, &r->accum, &r->stats, &r->cnt, &r->off, &r->tot, &r->list_s,
                      &r->list_m, &r->lo_s, &r->hi_s, &r->lo_m, &r->hi_m, &r->lo_b, &r->hi_b, &r->pts_s,
                      &r->pts_m, &r->geo_u, &r->geo_b, &r->ref_u, &r->ref_b, &r->mcolor, &r->segq, &r->pieces, &r->pieces_x, &r->piece_cell,
                      &r->xwork, &r->pkeys, &r->pvals, &r->unitn, &r->units, &r->guide, &r->gpend, &r->graw, &r->gsamp, &r->gcount,
                      &r->mislog, &r->miscount, &r->dparams, &r->dmaps, &r->mapn, &r->ovf})
        d->release();
    if (r->stage) (void)hipHostFree(r->stage);
    for (auto& e : r->stage_ev)
        if (e) (void)hipEventDestroy(e);
    delete r->gnet;
    r->gnet = nullptr;
    lbvh_free(r->bvh_s);
    lbvh_free(r->bvh_m);
    lbvh_free(r->bvh_b);
    wbvh_free(r->wbvh_m);
    wbvh_free(r->wbvh_b);
    lbvh_scratch_free(r->scratch);
    for (auto& e : r->ev)
        if (e) (void)hipEventDestroy(e);
    if (r->stream) (void)hipStreamDestroy(r->stream);
    delete r;
}
