// Wavefront path state, light-vertex / beam records, the medium-aware intersector and
// the fused connect / merge estimators with their constant-time MIS weights.
//
// The per-segment logic restates nolmoonen/gpuupbp gpuupbp/kernel/trace_light.cu:84-321,
// trace_camera.cu:79-382, intersector.cuh:45-418, functs_shared.cuh:37-187,
// functs_light.cuh:86-388, functs_camera.cuh:41-527, path_weight.cuh:35-51 and
// pde/functs_{pp,pb2d,bb1d,bp2d}.cuh. The data layout is this build's own: one
// 128-byte state record and one priority stack per path between bounce kernels,
// fixed-stride vertex and beam slots per light path (the per-path range replaces the
// reference's linked list), and sorted LBVH payloads for the photon maps.
#pragma once

#include "bsdf_light.h"
#include "lbvh.h"

namespace upbp {

// ---- per-iteration parameters (shared/launch_params.h:44-198 subset) -----------------
struct IterParams {
    uint32_t min_len, max_len, pbt, qbt, techniques, n_light, n_pixels, bb1d_used;
    uint32_t vstride, bstride; // vertex / beam slots per light path
    int do_bpt, do_bb1d, do_bp2d, do_pb2d, do_pp3d, do_surf;
    int build_bb1d, build_bp2d, build_pb2d, build_pp3d, build_surf;
    int trace_light, trace_camera, connect_from_surf, ignore_spec;
    float n_bpt;
    float mf_bb1d, mf_bp2d, mf_pb2d, mf_pp3d, mf_surf;
    float norm_bb1d, norm_bp2d, norm_pb2d, norm_pp3d, norm_surf;
    float r_bb1d, r_bp2d, r_pb2d, r_pp3d, r_surf, rad_point_med, rad_beam;
    float r2_bb1d, r2_bp2d, r2_pb2d, r2_pp3d, r2_surf;
    float sub_len;   // target sub-beam length of the beam map
    float piece_len; // target length of one camera-segment query piece
    float bundle_piece_len; // the same on the bounces walked in bundles
    float box_slack; // extra beam-box inflation covering FP32 error of the beam predicates
    int merge_bpc;   // resident blocks per CU of the merge kernels (0: occupancy limit)
    int piece_order; // B-B1D work order: 0 as written, 1 XCD regions, 2 Morton-sorted (pieces_x)
    int packet_bounces; // B-B1D: camera bounces below this walk the beam tree as 64-piece packets
    uint32_t block;     // threads per block of the light and camera bounce kernels (64, 128, 256)
    int beam_leaf;      // B-B1D leaf test reads 0: the sorted sub-beam records (beam_geo); 1: the
                        // sorted (beam slot, sub-beam j | n << 16) references and the beam record
    int bundle_bounces; // B-B1D: camera bounces below this walk the beam tree per bundle of pieces
    float bundle_radius; // largest bundle radius walked as one bundle (world units)
    float bundle_box;    // half the size of a sub-beam's box (world units), for the bundle cost test
    uint32_t cap_med, cap_beam; // capacities of the medium-point and sub-beam lists (host launch
                                // decisions; the counts themselves are only on the device)
    uint64_t iter_key;
};

// ---- recursive MIS quantities (shared/light_vertex.h:57-93) ------------------------
struct Weights {
    float d_shared, d_bpt, d_pde, for_inv, rev_inv, for_ratio, rev_ratio;
    int prev_in_medium, prev_specular;
};
struct StateWeights {
    Weights w;
    float last_sin, d_bpt_a, d_bpt_b, d_pde_a, d_pde_b;
};

// path flags
constexpr uint32_t PF_INFINITE = 1u, PF_SPECULAR_PATH = 2u, PF_ORIGIN_IN_MEDIUM = 4u, PF_ONLY_SPEC_SURF = 8u,
                   PF_PREV_MEDIUM = 16u, PF_PREV_SPECULAR = 32u;

// 128-byte path state between bounce kernels
struct __align__(16) PathState {
    f3 origin;
    uint32_t path_length;
    f3 dir;
    uint32_t flags;
    f3 throughput;
    float d_shared;
    float d_bpt, d_pde, for_inv, rev_inv;
    float for_ratio, rev_ratio, last_sin, d_bpt_a;
    float d_bpt_b, d_pde_a, d_pde_b, sx;
    f3 color; // camera paths: accumulated contribution
    float sy;
};

UPBP_HD void unpack_weights(const PathState& s, StateWeights& sw)
{
    sw.w.d_shared = s.d_shared;
    sw.w.d_bpt = s.d_bpt;
    sw.w.d_pde = s.d_pde;
    sw.w.for_inv = s.for_inv;
    sw.w.rev_inv = s.rev_inv;
    sw.w.for_ratio = s.for_ratio;
    sw.w.rev_ratio = s.rev_ratio;
    sw.w.prev_in_medium = (s.flags & PF_PREV_MEDIUM) != 0;
    sw.w.prev_specular = (s.flags & PF_PREV_SPECULAR) != 0;
    sw.last_sin = s.last_sin;
    sw.d_bpt_a = s.d_bpt_a;
    sw.d_bpt_b = s.d_bpt_b;
    sw.d_pde_a = s.d_pde_a;
    sw.d_pde_b = s.d_pde_b;
}

UPBP_HD void pack_weights(PathState& s, const StateWeights& sw)
{
    s.d_shared = sw.w.d_shared;
    s.d_bpt = sw.w.d_bpt;
    s.d_pde = sw.w.d_pde;
    s.for_inv = sw.w.for_inv;
    s.rev_inv = sw.w.rev_inv;
    s.for_ratio = sw.w.for_ratio;
    s.rev_ratio = sw.w.rev_ratio;
    s.flags = (s.flags & ~(PF_PREV_MEDIUM | PF_PREV_SPECULAR)) | (sw.w.prev_in_medium ? PF_PREV_MEDIUM : 0u) |
              (sw.w.prev_specular ? PF_PREV_SPECULAR : 0u);
    s.last_sin = sw.last_sin;
    s.d_bpt_a = sw.d_bpt_a;
    s.d_bpt_b = sw.d_bpt_b;
    s.d_pde_a = sw.d_pde_a;
    s.d_pde_b = sw.d_pde_b;
}

// ---- medium priority stack (kernel/pstack.cuh:33-137) ------------------------------
struct BStack {
    int n;
    // material and medium ids fit 16 bits (the renderer rejects larger scenes): the arrays
    // live in scratch memory, so narrower entries mean less scratch traffic
    int16_t mat[kMaxStack], med[kMaxStack];
    int pr[kMaxStack];
    // the live entries only: the arrays sit in scratch memory and paths hold one or two entries
    UPBP_HD void copy_from(const BStack& o)
    {
        n = o.n;
        for (int i = 0; i < o.n; i++) {
            mat[i] = o.mat[i];
            med[i] = o.med[i];
            pr[i] = o.pr[i];
        }
    }
    UPBP_HD bool push(int m, int d, int p)
    {
        if (n == kMaxStack) return false;
        int pos = 0;
        for (int i = n - 1; i >= 0; i--)
            if (pr[i] <= p) {
                pos = i + 1;
                break;
            }
        for (int i = n; i > pos; i--) {
            mat[i] = mat[i - 1];
            med[i] = med[i - 1];
            pr[i] = pr[i - 1];
        }
        mat[pos] = int16_t(m);
        med[pos] = int16_t(d);
        pr[pos] = p;
        n++;
        return true;
    }
    UPBP_HD bool pop(int m, int d, int p)
    {
        for (int i = n - 1; i >= 0 && pr[i] >= p; i--) {
            if (pr[i] == p && mat[i] == m && med[i] == d) {
                for (int j = i; j < n - 1; j++) {
                    mat[j] = mat[j + 1];
                    med[j] = med[j + 1];
                    pr[j] = pr[j + 1];
                }
                n--;
                return true;
            }
        }
        return false;
    }
};

// stack storage between kernels: entries of path p at stride 1 (AoS, 256 bytes per path)
struct __align__(16) StackRec {
    int n, pad[3];
    int4 e[kMaxStack]; // (mat, med, priority, 0)
};

__device__ __forceinline__ void load_stack(const StackRec* recs, uint32_t p, BStack& s)
{
    const StackRec* r = recs + p;
    s.n = r->n;
    for (int i = 0; i < s.n; i++) {
        const int4 v = r->e[i];
        s.mat[i] = int16_t(v.x);
        s.med[i] = int16_t(v.y);
        s.pr[i] = v.z;
    }
}
__device__ __forceinline__ void store_stack(StackRec* recs, uint32_t p, const BStack& s)
{
    StackRec* r = recs + p;
    r->n = s.n;
    for (int i = 0; i < s.n; i++) r->e[i] = make_int4(s.mat[i], s.med[i], s.pr[i], 0);
}

// ---- light vertex record: 7 x float4 = 112 bytes -----------------------------------
// vertex flag bits: Bsdf flags in bits 0-3
constexpr int VF_IN_MEDIUM = 16, VF_BEHIND_SURF = 32, VF_PREV_MEDIUM = 64, VF_PREV_SPECULAR = 128;

struct __align__(16) VertexRec {
    float4 a; // hit.xyz, throughput.x
    float4 b; // throughput.yz, frame source .xy
    float4 c; // frame source .z, local_dir_fix.xyz
    float4 d; // geom_local.xyz, cont_prob
    float4 e; // diff_prob, phong_prob, d_shared, d_bpt
    float4 f; // d_pde, for_inv, rev_inv, for_ratio
    float4 g; // rev_ratio, id, flags, path_length
};

struct LVertex {
    f3 hit, throughput;
    Bsdf bsdf;
    Weights w;
    uint32_t path_length;
    bool in_medium, behind_surf;
};

__device__ __forceinline__ void store_vertex(VertexRec* v, const LVertex& x, f3 frame_src)
{
    const Bsdf& b = x.bsdf;
    const int fl = b.flags | (x.in_medium ? VF_IN_MEDIUM : 0) | (x.behind_surf ? VF_BEHIND_SURF : 0) |
                   (x.w.prev_in_medium ? VF_PREV_MEDIUM : 0) | (x.w.prev_specular ? VF_PREV_SPECULAR : 0);
    v->a = make_float4(x.hit.x, x.hit.y, x.hit.z, x.throughput.x);
    v->b = make_float4(x.throughput.y, x.throughput.z, frame_src.x, frame_src.y);
    v->c = make_float4(frame_src.z, b.local_dir_fix.x, b.local_dir_fix.y, b.local_dir_fix.z);
    v->d = make_float4(b.geom_normal.x, b.geom_normal.y, b.geom_normal.z, b.cont_prob);
    v->e = make_float4(b.diff_prob, b.phong_prob, x.w.d_shared, x.w.d_bpt);
    v->f = make_float4(x.w.d_pde, x.w.for_inv, x.w.rev_inv, x.w.for_ratio);
    v->g = make_float4(x.w.rev_ratio, __int_as_float(b.id), __int_as_float(fl), __int_as_float(int(x.path_length)));
}

__device__ __forceinline__ void load_vertex(const VertexRec* v, LVertex& x)
{
    const float4 a = v->a, b = v->b, c = v->c, d = v->d, e = v->e, f = v->f, g = v->g;
    x.hit = mk(a.x, a.y, a.z);
    x.throughput = mk(a.w, b.x, b.y);
    const f3 src = mk(b.z, b.w, c.x);
    const int fl = __float_as_int(g.z);
    const int id = __float_as_int(g.y);
    if (fl & Bsdf::IN_MEDIUM) bsdf_restore_medium(x.bsdf, src, id, d.w, fl & 15);
    else bsdf_restore_surface(x.bsdf, src, mk(c.y, c.z, c.w), mk(d.x, d.y, d.z), id, e.x, e.y, d.w, fl & 15);
    x.bsdf.pos = x.hit;
    x.w.d_shared = e.z;
    x.w.d_bpt = e.w;
    x.w.d_pde = f.x;
    x.w.for_inv = f.y;
    x.w.rev_inv = f.z;
    x.w.for_ratio = f.w;
    x.w.rev_ratio = g.x;
    x.w.prev_in_medium = (fl & VF_PREV_MEDIUM) != 0;
    x.w.prev_specular = (fl & VF_PREV_SPECULAR) != 0;
    x.path_length = uint32_t(__float_as_int(g.w));
    x.in_medium = (fl & VF_IN_MEDIUM) != 0;
    x.behind_surf = (fl & VF_BEHIND_SURF) != 0;
}

// ---- light beam record: 8 x float4 = 128 bytes -------------------------------------
struct __align__(16) BeamRec {
    float4 a; // origin.xyz, length
    float4 b; // dir.xyz, medium
    float4 c; // throughput.xyz, for_pdf
    float4 d; // rev_pdf, ray sampling flags, path_idx, path_length | infinite << 31
    float4 e; // d_shared, d_bpt, d_pde, for_inv
    float4 f; // rev_inv, for_ratio, rev_ratio, prev bits
    float4 g; // last_sin, d_bpt_a, d_bpt_b, d_pde_a
    float4 h; // d_pde_b
};

struct LBeam {
    f3 origin, dir, throughput;
    float length, for_pdf, rev_pdf;
    int med;
    unsigned flags;
    uint32_t path_idx, path_length;
    bool infinite;
    StateWeights sw;
};

__device__ __forceinline__ void store_beam(BeamRec* r, const LBeam& x)
{
    const int prev = (x.sw.w.prev_in_medium ? 1 : 0) | (x.sw.w.prev_specular ? 2 : 0);
    r->a = make_float4(x.origin.x, x.origin.y, x.origin.z, x.length);
    r->b = make_float4(x.dir.x, x.dir.y, x.dir.z, __int_as_float(x.med));
    r->c = make_float4(x.throughput.x, x.throughput.y, x.throughput.z, x.for_pdf);
    r->d = make_float4(x.rev_pdf, __int_as_float(int(x.flags)), __int_as_float(int(x.path_idx)),
                       __int_as_float(int(x.path_length | (x.infinite ? 0x80000000u : 0u))));
    r->e = make_float4(x.sw.w.d_shared, x.sw.w.d_bpt, x.sw.w.d_pde, x.sw.w.for_inv);
    r->f = make_float4(x.sw.w.rev_inv, x.sw.w.for_ratio, x.sw.w.rev_ratio, __int_as_float(prev));
    r->g = make_float4(x.sw.last_sin, x.sw.d_bpt_a, x.sw.d_bpt_b, x.sw.d_pde_a);
    r->h = make_float4(x.sw.d_pde_b, 0.f, 0.f, 0.f);
}

__device__ __forceinline__ void load_beam(const BeamRec* r, LBeam& x)
{
    const float4 a = r->a, b = r->b, c = r->c, d = r->d, e = r->e, f = r->f, g = r->g, h = r->h;
    x.origin = mk(a.x, a.y, a.z);
    x.length = a.w;
    x.dir = mk(b.x, b.y, b.z);
    x.med = __float_as_int(b.w);
    x.throughput = mk(c.x, c.y, c.z);
    x.for_pdf = c.w;
    x.rev_pdf = d.x;
    x.flags = unsigned(__float_as_int(d.y));
    x.path_idx = uint32_t(__float_as_int(d.z));
    const uint32_t pl = uint32_t(__float_as_int(d.w));
    x.path_length = pl & 0x7fffffffu;
    x.infinite = (pl >> 31) != 0;
    x.sw.w.d_shared = e.x;
    x.sw.w.d_bpt = e.y;
    x.sw.w.d_pde = e.z;
    x.sw.w.for_inv = e.w;
    x.sw.w.rev_inv = f.x;
    x.sw.w.for_ratio = f.y;
    x.sw.w.rev_ratio = f.z;
    const int prev = __float_as_int(f.w);
    x.sw.w.prev_in_medium = (prev & 1) != 0;
    x.sw.w.prev_specular = (prev & 2) != 0;
    x.sw.last_sin = g.x;
    x.sw.d_bpt_a = g.y;
    x.sw.d_bpt_b = g.z;
    x.sw.d_pde_a = g.w;
    x.sw.d_pde_b = h.x;
}

// ---- device buffers of one renderer --------------------------------------------------
// Long beams are cut into sub-beams so that their boxes stay tight. Every sub-beam leaf
// carries the whole beam (the predicates run on it unchanged) plus the parameter range
// [lo, hi) it owns; a pair is accepted only by the sub-beam whose range holds the
// predicate's own beam parameter, so each accepted beam is reported exactly once.
constexpr uint32_t kMaxSubBeams = 1024;
constexpr int SB_BB1D = 1; // sub-beam flag: beam belongs to the B-B1D subset (path < bb1d_used)

UPBP_HD uint32_t subbeam_count(float len, float sub_len)
{
    if (!(sub_len > 0.f) || !(len > sub_len)) return 1u;
    const float n = ceilf(len / sub_len);
    return n >= float(kMaxSubBeams) ? kMaxSubBeams : uint32_t(n);
}

// sub-beam j of n over [0, len): owned parameter range and the box of its beam piece
UPBP_HD void subbeam_range(float len, uint32_t j, uint32_t n, float& lo, float& hi, float& ta, float& tb)
{
    const float step = len / float(n);
    lo = j == 0 ? -kInf : float(j) * step;
    hi = j + 1 == n ? kInf : float(j + 1) * step;
    ta = j == 0 ? 0.f : lo;
    tb = j + 1 == n ? len : hi;
}

UPBP_HD void subbeam_box(f3 o, f3 d, float len, float ta, float tb, float r, float slack, float4& blo, float4& bhi)
{
    const f3 a = o + d * ta, e = o + d * tb;
    const float mag = fmaxf(fmaxf(fmaxf(fabsf(a.x), fabsf(a.y)), fmaxf(fabsf(a.z), fabsf(e.x))),
                            fmaxf(fmaxf(fabsf(e.y), fabsf(e.z)), fmaxf(fabsf(o.x), fmaxf(fabsf(o.y), fabsf(o.z)))));
    const float rr = r * (1.f + 1e-4f) + 1e-5f * mag + 1e-6f * len + slack;
    blo = make_float4(fminf(a.x, e.x) - rr, fminf(a.y, e.y) - rr, fminf(a.z, e.z) - rr, 0.f);
    bhi = make_float4(fmaxf(a.x, e.x) + rr, fmaxf(a.y, e.y) + rr, fmaxf(a.z, e.z) + rr, 0.f);
}

struct PhotonMaps {
    Bvh surf, med, beam;
    WBvh wmed, wbeam; // wide collapses of med and beam (segment merge kernels)
    const float4* surf_pts; // sorted: (hit.xyz, vertex slot)
    const float4* med_pts;
    const float4* beam_geo; // sorted: 3 per sub-beam: (origin.xyz, length), (dir.xyz, beam slot),
                            // (lo, hi, medium, SB_* flags)
    const uint2* beam_ref;  // sorted: per sub-beam (beam slot, j | n << 16), sub-beam j of n
};

// A camera-segment query is walked in pieces, each an independent work item of the merge
// kernels: piece k of n owns the camera parameters t in [lo, hi) and traverses only
// [ta, tb], while the predicates still see the whole segment, so an accepted pair is found
// by exactly one piece. A ray's pieces number at most kPiecesPerRay plus one per segment.
constexpr uint32_t kPiecesPerRay = 32;

UPBP_HD uint32_t piece_count(float len, float piece_len)
{
    if (!(piece_len > 0.f) || !(len > piece_len)) return 1u;
    const float n = ceilf(len / piece_len);
    return n >= float(kPiecesPerRay) ? kPiecesPerRay : uint32_t(n);
}

UPBP_HD void piece_range(float dmin, float dmax, uint32_t k, uint32_t n, float& lo, float& hi, float& ta, float& tb)
{
    const float step = (dmax - dmin) / float(n);
    lo = k == 0 ? -kInf : dmin + float(k) * step;
    hi = k + 1 == n ? kInf : dmin + float(k + 1) * step;
    ta = k == 0 ? dmin : lo;
    tb = k + 1 == n ? dmax : hi;
}

// One camera-segment query (a scattering medium segment of one camera ray), handed from
// the camera bounce kernel to the P-B2D / B-B1D merge kernels: 7 x float4 = 112 bytes.
struct __align__(16) SegRec {
    float4 a; // ro.xyz, dmin
    float4 b; // rd.xyz, dmax
    float4 c; // throughput * attenuation before the segment (xyz), pixel
    float4 d; // rsf, rsr, medium, bits: ray-sampling flags 0-1, prev medium 2, prev specular 3, cam_len << 8
    float4 e; // d_shared, d_bpt, d_pde, for_inv
    float4 f; // rev_inv, for_ratio, rev_ratio, last_sin
    float4 g; // d_bpt_a, d_bpt_b, d_pde_a, d_pde_b
};

struct Bufs {
    PathState* lstate;
    StackRec* lstack;
    PathState* cstate;
    StackRec* cstack;
    uint32_t* queue[2];
    uint32_t* qcount; // paths alive at bounce k: light [k], camera [64 + k]; segment queries
                      // [128 + k]; merge work heads [192 + 2k + kind]; query pieces [320 + k];
                      // piece bundles [384 + k]; pieces left to the per-piece B-B1D walk [448 + k]
    VertexRec* verts;
    uint32_t* nvert; // per light path
    BeamRec* beams;
    uint32_t* nbeam;
    float* fb32;
    float* mcolor;  // per camera path: P-B2D / B-B1D contributions of the merge kernels
    SegRec* segq;   // segment queries of the current camera bounce
    uint2* pieces;  // their pieces: (query, k | n << 16)
    uint2* pieces_x; // the same, grouped by XCD region for the B-B1D merge kernel (or the pieces
                     // the bundle kernel leaves to the per-piece walk)
    uint2* units;    // piece bundles of a camera wave: (first piece, count <= 64), the wave's
                     // j-th pieces of its rays (null when no bounce walks bundles)
    uint8_t* piece_cell; // spatial cell of each piece
    uint32_t* xwork; // per bounce: kXWork words of cell counts, offsets and per-XCD work ranges
    uint32_t seg_cap, piece_cap;
    // learned-proposal training (null when off): per camera path the pending record of its
    // last medium scattering (kGuidePend floats: position in [0,1]^3, zero arrival inputs,
    // scattered direction, training-weight denominator (0 = none), luminance of the path colour
    // after it), the raw pool (kGuideRaw floats: the pending record with the colour snapshot
    // completed by the merge colour, then the pixel), the training samples finalised from it
    // (10 floats: position, arrival, scattered, weight) and the two counters
    float* gpend;
    float* graw;
    float* gsamp;
    uint32_t* gcount;
    uint32_t gcap;  // raw pool capacity (per iteration)
    uint32_t gring; // training-sample ring capacity (samples written at gcount[1] modulo gring)
    double* accum;
    unsigned long long* stats; // see STAT_*
    // this iteration's parameters and photon maps, in device memory: kernel arguments stay the
    // same from one iteration to the next (a captured iteration can be replayed), and the photon
    // counts are written by the device (no host synchronisation before the camera pass)
    const IterParams* params;
    const PhotonMaps* maps;
    // MIS weight log of the MIS oracle test (null when off): one record per weighted contribution
    // of the camera paths whose index is a multiple of mis_stride and of the light paths below
    // mis_lt (light tracing), as (technique | k << 8 | s << 16, camera path or light path, light
    // path or light, weight bits); x_0 .. x_k is the full path, s its light-side vertex count
    uint4* mislog;
    uint32_t* miscount;
    uint32_t miscap, mis_stride, mis_lt;
};

// techniques of the MIS weight log
enum MisTech : uint32_t { MT_LIGHT_HIT = 0, MT_NEE, MT_BPT, MT_LT, MT_SURF, MT_PP3D, MT_PB2D, MT_BP2D, MT_BB1D };

__device__ __forceinline__ void mis_log(const Bufs& b, uint32_t tech, uint32_t k, uint32_t s, uint32_t id,
                                        uint32_t lp, float w)
{
    const uint32_t at = atomicAdd(b.miscount, 1u);
    if (at < b.miscap) b.mislog[at] = make_uint4(tech | (k << 8) | (s << 16), id, lp, __float_as_uint(w));
}

enum Stat : int { STAT_LIGHT_STEPS = 0, STAT_CAMERA_STEPS, STAT_BPT_CONN, STAT_POINT_Q, STAT_SEG_Q, STAT_ACCEPT,
                  STAT_VDROP, STAT_BDROP, STAT_ERRORS, STAT_SEG_NODES, STAT_SEG_CANDS, STAT_ACC_PB2D, STAT_ACC_BB1D,
                  STAT_SEG_PIECES, STAT_SEG_TRIPS, STAT_CAND_BATCHES, STAT_PAIR_BATCHES, STAT_MERGE_STACK,
                  STAT_BUNDLES, STAT_BUNDLE_LEFT, STAT_BUNDLE_NODES, STAT_BUNDLE_LEAVES, STAT_BUNDLE_TESTS,
                  STAT_COUNT };
constexpr int kQcountWords = 512;
// One render job of a batched launch (device memory): its scene, buffers and the few launch
// values the single-job launches pass as arguments. A batched stage kernel runs the blocks of
// every job in one grid; a block map (job, block of that job) per stage says which block is whose.
struct BatchJob {
    DScene sc;
    Bufs b;
    const uint32_t* ovf; // the job's photon-list overflow flag (its iteration is then not accumulated)
    uint32_t res_x, tiles_x;
    float fres_x, fres_y;
};

// XCD-aware work split of the B-B1D merge: 4x4x4 spatial cells, 8 XCDs owning one octant
// (2x2x2 cells) each. Per bounce: cell counts [0,64), cell offsets [64,128), fill cursors
// [128,192), XCD range begin [192,200), end [200,208), heads [208,216).
constexpr int kXcds = 8, kCells = 64, kXWork = 224;

// wave-aggregated counter increment
__device__ __forceinline__ void stat_add(unsigned long long* stats, int k, uint32_t v)
{
    if (v) atomicAdd(&stats[k], (unsigned long long)v);
}

// wave-aggregated queue append (ballot + prefix popcount, one atomic per wave)
__device__ __forceinline__ void queue_append(bool keep, uint32_t item, uint32_t* queue, uint32_t* count)
{
    const uint64_t mask = __ballot(keep);
    if (!keep) return;
    const uint32_t lane = __lane_id();
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if (int(lane) == leader) base = atomicAdd(count, uint32_t(__popcll(mask)));
    base = __shfl(base, leader);
    queue[base + __popcll(mask & lt)] = item;
}

// ---- intersector --------------------------------------------------------------------
struct VolSeg {
    f3 att;
    float dmin, dmax, pf, pr;
    int med;
};
struct LiteSeg {
    float dmin, dmax;
    int med;
};
struct LiteHit {
    float dist;
    int mat, med, enter;
};

UPBP_HD int priority_of(const DScene& sc, int mat) { return mat >= 0 ? sc.mats[mat].priority : kGlobalPriority; }
UPBP_HD bool real_geom(const DScene& sc, int mat) { return mat >= 0 && sc.mats[mat].real; }
UPBP_HD const DMedium& medium_at(const DScene& sc, int id)
{
    return sc.media[int(fminf(float(id), float(sc.n_media - 1)))];
}

// closest real hit in (tmin, tmax) (intersector.cu:103-150 fill semantics)
__device__ __forceinline__ bool trace_real(const DScene& sc, f3 o, f3 d, float tmin, float tmax, Hit& out,
                                           bool& stack_ok)
{
    int best = -1;
    float bt = 0.f, bu = 0.f, bv = 0.f;
    const TriSet& T = sc.real;
    stack_ok = bvh_ray_query(T.bvh, o, d, tmin, tmax, [&](uint32_t i, float& thi) {
        float t, u, v;
        if (tri_test(T.tri[3 * i], T.tri[3 * i + 1], T.tri[3 * i + 2], o, d, tmin, thi, t, u, v)) {
            best = int(i);
            bt = t;
            bu = u;
            bv = v;
            thi = t;
        }
    });
    // analytic spheres (predefined scenes: a handful, tested exhaustively after the triangles)
    int best_s = -1;
    f3 sn{0.f, 0.f, 0.f};
    for (uint32_t i = 0; i < sc.sph_real.n; i++) {
        float th[2];
        f3 nh[2];
        if (sphere_test(sc.sph_real.sph[i], o, d, tmin, best >= 0 || best_s >= 0 ? bt : tmax, th, nh)) {
            best_s = int(i);
            bt = th[0];
            sn = nh[0];
        }
    }
    if (best_s >= 0) {
        const int2 rec = sc.sph_real.rec[best_s];
        out.dist = bt;
        out.mat = rec.x;
        out.med = sc.mats[rec.x].med_idx;
        out.light = rec.y;
        out.normal = sn;
        out.shading = sn;
        out.enter = dot(sn, d) < 0.f;
        return true;
    }
    if (best < 0) return false;
    const int2 rec = T.rec[best];
    const float4 fn = T.nrm[4 * best + 3];
    out.dist = bt;
    out.mat = rec.x;
    out.med = sc.mats[rec.x].med_idx;
    out.light = rec.y;
    out.normal = mk(fn.x, fn.y, fn.z);
    if (sc.use_shading_normal) {
        const float4 n0 = T.nrm[4 * best], n1 = T.nrm[4 * best + 1], n2 = T.nrm[4 * best + 2];
        out.shading = normalize(mk(n0.x, n0.y, n0.z) * (1.f - bu - bv) + mk(n1.x, n1.y, n1.z) * bu +
                                mk(n2.x, n2.y, n2.z) * bv);
    } else {
        out.shading = out.normal;
    }
    out.enter = dot(out.normal, d) < 0.f;
    return true;
}

// all imaginary hits in (tmin, tmax), distance-ordered, at most kMaxImaginary (intersector.cu:152-167)
__device__ __forceinline__ bool trace_imag(const DScene& sc, f3 o, f3 d, float tmin, float tmax, LiteHit* hits,
                                           int& nh, bool& stack_ok)
{
    const TriSet& T = sc.imag;
    bool overflow = false;
    nh = 0;
    auto insert = [&](float t, int mat, f3 n) {
        if (nh == kMaxImaginary) {
            overflow = true;
            return;
        }
        LiteHit h{t, mat, sc.mats[mat].med_idx, dot(n, d) < 0.f ? 1 : 0};
        int pos = 0; // insertion after the last entry with distance <= t (pstack.cuh:70-95)
        for (int k = nh - 1; k >= 0; k--)
            if (hits[k].dist <= t) {
                pos = k + 1;
                break;
            }
        for (int k = nh; k > pos; k--) hits[k] = hits[k - 1];
        hits[pos] = h;
        nh++;
    };
    stack_ok = bvh_ray_query(T.bvh, o, d, tmin, tmax, [&](uint32_t i, float&) {
        float t, u, v;
        if (overflow) return;
        if (!tri_test(T.tri[3 * i], T.tri[3 * i + 1], T.tri[3 * i + 2], o, d, tmin, tmax, t, u, v)) return;
        const float4 fn = T.nrm[4 * i + 3];
        insert(t, T.rec[i].x, mk(fn.x, fn.y, fn.z));
    });
    // imaginary spheres report both crossings (intersector.cu:78-93 under an any-hit program)
    for (uint32_t i = 0; i < sc.sph_imag.n && !overflow; i++) {
        float th[2];
        f3 nrm[2];
        const int k = sphere_test(sc.sph_imag.sph[i], o, d, tmin, tmax, th, nrm);
        for (int j = 0; j < k && !overflow; j++) insert(th[j], sc.sph_imag.rec[i].x, nrm[j]);
    }
    return !overflow;
}

// Products over the medium segments of an occlusion ray, in segment order: the forward and
// reverse ray-sampling pdfs and the attenuation (what acc_for / acc_rev / acc_att give over the
// segment list of a traced ray).
struct SegAcc {
    float pf, pr;
    f3 att;
};

// kernel/intersector.cuh:45-393. OCC selects the occlusion variant (intersector.cuh:396-418):
// no distance sampling, the segments only accumulated into *acc (no segment lists are stored),
// and the caller's stack left unchanged. The traced variant stores the segments in segs and the
// light segments in lite, samples a scattering distance with rng and moves bstack along the ray.
template <bool OCC>
__device__ __noinline__ int intersect_impl(const DScene& sc, bool& hit, f3 ro, f3 rd, Hit& result, BStack& bstack,
                                           unsigned flags, VolSeg* segs, int& nseg, LiteSeg* lite, int& nlite,
                                           Rng* rng, SegAcc* acc)
{
    hit = false;
    nseg = 0;
    nlite = 0;
    Hit tmp = result;
    float dist = 0.f;
    LiteHit hwm[kMaxRealHits];
    int nhwm = 0;
    const int top_priority = bstack.pr[bstack.n - 1];
    const float first_eps = (flags & kOriginInMedium) ? 0.f : kEpsRay;
    float t_max, miss_dist;
    if (OCC) {
        t_max = tmp.dist - kEpsRay;
        miss_dist = tmp.dist;
    } else {
        t_max = kInf;
        miss_dist = kInf;
    }
    BStack cp;
    cp.copy_from(bstack);
    float t_min = first_eps;
    bool not_yet_real;
    bool ok;
    do {
        not_yet_real = false;
        Hit h;
        hit = trace_real(sc, ro, rd, t_min, t_max, h, ok);
        if (!ok) return S_ERR_STACK;
        if (hit) tmp = h;
        else tmp.dist = kInf;
        if (hit) {
            dist = tmp.dist;
            const int priority = priority_of(sc, tmp.mat);
            if (tmp.med < 0) tmp.med = cp.med[cp.n - 1];
            if (priority < top_priority && priority != kThinWall) {
                if (tmp.enter) {
                    if (nhwm == kMaxRealHits) return S_ERR_REAL_HITS;
                    hwm[nhwm++] = {tmp.dist, tmp.mat, tmp.med, 1};
                    if (!cp.push(tmp.mat, tmp.med, priority)) return S_ERR_STACK;
                } else if (cp.pop(tmp.mat, tmp.med, priority)) {
                    if (nhwm == kMaxRealHits) return S_ERR_REAL_HITS;
                    hwm[nhwm++] = {tmp.dist, tmp.mat, tmp.med, 0};
                }
                not_yet_real = true;
                t_min = dist + kEpsRay;
                // the remaining occlusion range is shortened by the hit distance again, as in
                // intersector.cuh:175-177
                t_max = OCC ? miss_dist - dist - kEpsRay : kInf;
            }
        }
    } while (not_yet_real && t_max > t_min);
    if (!hit || not_yet_real) {
        hit = false;
        dist = miss_dist;
    }
    result = tmp;
    result.dist = dist;

    const bool origin_in_medium = flags & kOriginInMedium;
    const bool end_in_medium = flags & kEndInMedium;
    if (OCC) {
        acc->pf = acc->pr = 1.f;
        acc->att = mk(1.f);
    }
    // The occlusion variant consumes each light segment as it is found: the pdfs and attenuation
    // of its medium, multiplied in segment order. A segment without a medium adds nothing.
    auto occ_segment = [&](float dmin, float dmax, int med) {
        if (med < 0) return;
        const DMedium& m = medium_at(sc, med);
        unsigned f = 0;
        if (origin_in_medium && dmin == 0) f |= kOriginInMedium;
        if (end_in_medium && dmax == dist) f |= kEndInMedium;
        float fp = 1.f, rp = 1.f;
        if (m.has_scattering) fp = ray_sample_pdf(m, dmin, dmax, f, &rp);
        acc->att *= eval_attenuation(m, dmax - dmin);
        acc->pf *= fp;
        acc->pr *= rp;
        nseg++;
    };

    bool scattered = false;
    LiteHit imag[kMaxImaginary];
    int nimag = 0;
    bool have_imag = false;
    if (!real_geom(sc, bstack.mat[bstack.n - 1]) && dist > first_eps + kEpsRay) {
        have_imag = true;
        if (!trace_imag(sc, ro, rd, first_eps, dist - kEpsRay, imag, nimag, ok)) return S_ERR_IMAGINARY;
        if (!ok) return S_ERR_STACK;
        cp.copy_from(bstack);
        float dmin = 0.f;
        for (int i = 0; i < nimag; i++) {
            const LiteHit& is = imag[i];
            const int priority = priority_of(sc, is.mat);
            if (is.enter) {
                if (priority >= cp.pr[cp.n - 1]) {
                    if (nlite == kMaxSegments) return S_ERR_SEGMENTS;
                    if (OCC) occ_segment(dmin, is.dist, cp.med[cp.n - 1]);
                    else lite[nlite] = {dmin, is.dist, cp.med[cp.n - 1]};
                    nlite++;
                    dmin = is.dist;
                }
                if (!cp.push(is.mat, is.med, priority)) return S_ERR_STACK;
            } else {
                if (priority == cp.pr[cp.n - 1] && is.mat == cp.mat[cp.n - 1] && is.med == cp.med[cp.n - 1]) {
                    if (nlite == kMaxSegments) return S_ERR_SEGMENTS;
                    if (OCC) occ_segment(dmin, is.dist, cp.med[cp.n - 1]);
                    else lite[nlite] = {dmin, is.dist, cp.med[cp.n - 1]};
                    nlite++;
                    dmin = is.dist;
                }
                cp.pop(is.mat, is.med, priority);
            }
        }
        if (dmin < dist) {
            if (nlite == kMaxSegments) return S_ERR_SEGMENTS;
            if (OCC) occ_segment(dmin, dist, cp.med[cp.n - 1]);
            else lite[nlite] = {dmin, dist, cp.med[cp.n - 1]};
            nlite++;
        }
    } else {
        if (OCC) occ_segment(0.f, dist, bstack.med[bstack.n - 1]);
        else lite[nlite] = {0.f, dist, bstack.med[bstack.n - 1]};
        nlite++;
    }
    if (OCC) return S_OK;

    int popped = 0;
    for (int i = 0; i < nlite; i++) {
        const LiteSeg sg = lite[i];
        if (sg.med < 0) continue;
        const DMedium& m = medium_at(sc, sg.med);
        unsigned f = 0;
        if (origin_in_medium && sg.dmin == 0) f |= kOriginInMedium;
        if (end_in_medium && sg.dmax == dist) f |= kEndInMedium;
        float fp = 1.f, rp = 1.f;
        float dmax = sg.dmax;
        if (m.has_scattering) {
            float r;
            do {
                r = rng->next();
            } while (r == 0.f);
            const float to_next = sg.dmax - sg.dmin;
            const float dm = sample_ray(m, to_next, r, &fp, f, &rp);
            if (dm < to_next) {
                result.dist = sg.dmin + dm;
                result.mat = -1;
                result.med = sg.med;
                result.light = -1;
                result.normal = mk(0.f);
                result.enter = false;
                hit = true;
                scattered = true;
                dmax = result.dist;
            }
        }
        if (nseg == kMaxSegments) return S_ERR_SEGMENTS;
        segs[nseg++] = {eval_attenuation(m, dmax - sg.dmin), sg.dmin, dmax, fp, rp, sg.med};
        if (have_imag) {
            for (int j = popped; j < nimag; j++) {
                const LiteHit& is = imag[j];
                if (is.dist > dmax) break;
                if (is.enter) {
                    if (!bstack.push(is.mat, is.med, priority_of(sc, is.mat))) return S_ERR_STACK;
                } else if (bstack.n > 1) {
                    bstack.pop(is.mat, is.med, priority_of(sc, is.mat));
                }
                popped++;
            }
        }
        if (scattered) break;
    }
    for (int j = 0; j < nhwm; j++) {
        const LiteHit& h = hwm[j];
        if (h.dist > result.dist) break;
        if (h.enter) {
            if (!bstack.push(h.mat, h.med, priority_of(sc, h.mat))) return S_ERR_STACK;
        } else if (bstack.n > 1) {
            bstack.pop(h.mat, h.med, priority_of(sc, h.mat));
        }
    }
    return S_OK;
}

// a traced ray (kernel/intersector.cuh:45-393): distance sampling with rng
__device__ __forceinline__ int intersect(const DScene& sc, bool& hit, f3 ro, f3 rd, Hit& result, BStack& bstack,
                                         unsigned flags, VolSeg* segs, int& nseg, LiteSeg* lite, int& nlite, Rng* rng)
{
    return intersect_impl<false>(sc, hit, ro, rd, result, bstack, flags, segs, nseg, lite, nlite, rng, nullptr);
}

// kernel/intersector.cuh:396-418: whether the segment [0, tmax) of the ray is blocked, and the
// segment products along it (all 1 when the ray crosses no medium)
__device__ __forceinline__ int occluded(const DScene& sc, bool& hit, f3 p, f3 d, float tmax, const BStack& bs,
                                        unsigned flags, SegAcc& acc)
{
    Hit tmp;
    hit_reset(tmp, tmax);
    int nseg, nlite;
    return intersect_impl<true>(sc, hit, p, d, tmp, const_cast<BStack&>(bs), flags, nullptr, nseg, nullptr, nlite,
                                nullptr, &acc);
}

UPBP_HD float acc_for(const VolSeg* s, int n)
{
    float r = 1.f;
    for (int i = 0; i < n; i++) r *= s[i].pf;
    return r;
}
UPBP_HD float acc_rev(const VolSeg* s, int n)
{
    float r = 1.f;
    for (int i = 0; i < n; i++) r *= s[i].pr;
    return r;
}
UPBP_HD f3 acc_att(const VolSeg* s, int n)
{
    f3 r = mk(1.f);
    for (int i = 0; i < n; i++) r *= s[i].att;
    return r;
}

// kernel/scene.cuh:113-152
UPBP_HD float relative_ior(const DScene& sc, const Hit& is, const BStack& st)
{
    if (!is.on_surface()) return -1.f;
    const float ior1 = sc.mats[is.mat].ior;
    if (ior1 < 0.f) return -1.f;
    float ior2 = 1.f;
    const int m2 = is.enter ? st.mat[st.n - 1] : (st.n > 1 ? st.mat[st.n - 2] : st.mat[st.n - 1]);
    if (m2 >= 0 && sc.mats[m2].real) ior2 = sc.mats[m2].ior;
    if (ior2 < 0.f) return -1.f;
    return is.enter ? ior2 / ior1 : ior1 / ior2;
}

// kernel/path_weight.cuh:35-51
UPBP_HD float pde_factor(const IterParams& p, bool in_med, bool spec, float for_inv, float for_ratio, float rev_inv,
                         float rev_ratio, float sin_theta)
{
    const float fp = p.pbt == T_BEAM_LONG ? for_inv : for_ratio;
    const float rp = p.qbt == T_BEAM_LONG ? rev_inv : rev_ratio;
    return float(!in_med && !spec) * p.mf_surf +
           float(in_med) * (p.mf_pp3d + p.mf_pb2d * rp + p.mf_bp2d * fp + p.mf_bb1d * sin_theta * fp * rp);
}

UPBP_HD int init_stack(const DScene& sc, BStack& st, int mat, int med)
{
    st.n = 0;
    st.push(-1, int(sc.global_medium), kGlobalPriority);
    if (mat != -1 && med != -1)
        if (!st.push(mat, med, priority_of(sc, mat))) return S_ERR_STACK;
    return S_OK;
}

struct Sub { // subpath state held in registers within a bounce kernel
    BStack stack;
    f3 origin, dir, throughput;
    uint32_t path_length;
    bool infinite, specular_path;
    StateWeights sw;
};

// kernel/functs_shared.cuh:37-154. guide_den (optional, camera paths): lum(throughput times
// the scattering function times the cosine) / continuation probability of a continuing
// scattering in a medium or at a guidable surface whose proposal is on (0 elsewhere), the
// denominator of the learned proposal's training weight.
__device__ __forceinline__ int sample_scattering(const DScene& sc, const IterParams& p, bool& cont, const Bsdf& b,
                                                 f3 hitp, const Hit& is, Sub& s, Rng& rng, float* guide_den = nullptr)
{
    const float r0 = rng.next(), r1 = rng.next(), r2 = rng.next(), r3 = rng.next();
    float for_pdf, cos_out, sin_theta;
    unsigned ev;
    f3 dir;
    const f3 f = bsdf_sample(b, r0, r1, r2, dir, for_pdf, cos_out, sc, ev, sin_theta);
    if (black_or_neg(f)) {
        cont = false;
        return S_OK;
    }
    const bool specular = (ev & Bsdf::SPECULAR) != 0;
    float rev_pdf = for_pdf;
    if (!specular) rev_pdf = bsdf_pdf_rev(b, dir, sc);
    const float cp = b.cont_prob;
    if (cp == 0 || (cp < 1.f && r3 > cp)) {
        cont = false;
        return S_OK;
    }
    for_pdf *= cp;
    rev_pdf *= cp;
    if (guide_den) {
        const bool trains = b.in_medium() ? sc.guide.train_media != 0
                                          : (sc.guide.train_surfaces != 0 && surface_guidable(sc.mats[b.id]));
        *guide_den = trains ? luminance(s.throughput * f) * cos_out / cp : 0.f;
    }
    Weights& w = s.sw.w;
    const float bpt_i1 = float(!w.prev_specular && !specular);
    if (specular) {
        s.sw.d_bpt_a = cos_out / for_pdf;
        s.sw.d_bpt_b = cos_out * w.d_bpt / w.rev_inv;
        s.sw.d_pde_a = cos_out / for_pdf;
        s.sw.d_pde_b = cos_out * w.d_pde / w.rev_inv;
    } else {
        s.sw.d_bpt_a = cos_out / for_pdf;
        s.sw.d_bpt_b = (cos_out / for_pdf) * (w.d_shared * bpt_i1 + rev_pdf * w.d_bpt / w.rev_inv);
        s.sw.d_pde_a = cos_out / for_pdf;
        s.sw.d_pde_b = (cos_out / for_pdf) * (w.d_shared * bpt_i1 * p.n_bpt + rev_pdf * w.d_pde / w.rev_inv);
    }
    w.d_shared = 1.f / for_pdf;
    s.specular_path = s.specular_path && specular;
    s.origin = hitp;
    s.dir = dir;
    s.throughput *= f * (cos_out / for_pdf);
    w.prev_in_medium = b.in_medium();
    w.prev_specular = specular;
    s.sw.last_sin = sin_theta;
    if (ev & Bsdf::REFRACT) { // kernel/scene.cuh:75-96
        if (is.enter) {
            if (is.med >= 0)
                if (!s.stack.push(is.mat, is.med, priority_of(sc, is.mat))) return S_ERR_STACK;
        } else if (s.stack.n > 1) {
            s.stack.n--;
        }
    }
    cont = true;
    return S_OK;
}

// state record <-> registers
__device__ __forceinline__ void load_sub(const PathState& st, Sub& s)
{
    s.origin = st.origin;
    s.dir = st.dir;
    s.throughput = st.throughput;
    s.path_length = st.path_length;
    s.infinite = (st.flags & PF_INFINITE) != 0;
    s.specular_path = (st.flags & PF_SPECULAR_PATH) != 0;
    unpack_weights(st, s.sw);
}
__device__ __forceinline__ void save_sub(PathState& st, const Sub& s)
{
    st.origin = s.origin;
    st.dir = s.dir;
    st.throughput = s.throughput;
    st.path_length = s.path_length;
    st.flags = (st.flags & ~(PF_INFINITE | PF_SPECULAR_PATH)) | (s.infinite ? PF_INFINITE : 0u) |
               (s.specular_path ? PF_SPECULAR_PATH : 0u);
    pack_weights(st, s.sw);
}

} // namespace upbp
