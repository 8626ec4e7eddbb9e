// Camera-subpath wavefront: one bounce kernel per path segment runs the medium-aware
// intersector, appends one P-B2D / B-B1D query per scattering medium segment of the ray
// (evaluated by the segment merge kernels), then light hits, next-event estimation,
// vertex connections over the chosen light path's vertex range, surface / P-P3D / B-P2D
// merges (point queries), scattering and wave-aggregated queue compaction. Every
// contribution carries its constant-time MIS weight computed in place.
// Semantics: nolmoonen/gpuupbp gpuupbp/kernel/trace_camera.cu:49-390,
// gpuupbp/kernel/functs_camera.cuh:41-527 and gpuupbp/kernel/pde/functs_*.cuh.
#include "estimators.h"
#include "kernels.h"

namespace upbp {

// kernel/pde/functs_pb2d.cuh:226-342 and functs_bb1d.cuh:273-397: one P-B2D / B-B1D query per
// scattering medium segment of the camera ray, carrying the camera state the merge needs
// (throughput times the attenuation before the segment, the accumulated ray-sampling pdfs
// and the MIS quantities of the segment's start). MODE 0 counts the queries and sums their
// lengths, MODE 1 also counts their pieces, MODE 2 writes queries and pieces.
struct SegCount {
    uint32_t recs = 0, pieces = 0;
    float len = 0.f;
};

template <int MODE>
__device__ __forceinline__ SegCount segment_queries(const DScene& sc, const VolSeg* segs, const LiteSeg* lite, int n,
                                                    unsigned flags, uint32_t qbt, float piece_len, const Sub& s,
                                                    uint32_t pix, SegRec* out, uint8_t* npl)
{
    f3 att = mk(1.f);
    float rsf = 1.f, rsr = 1.f;
    SegCount c;
    for (int i = 0; i < n; i++) {
        const float dmin = segs ? segs[i].dmin : lite[i].dmin;
        const float dmax = segs ? segs[i].dmax : lite[i].dmax;
        const int mid = segs ? segs[i].med : lite[i].med;
        if (mid < 0) continue;
        const DMedium& m = sc.media[mid];
        if (m.has_scattering) {
            const uint32_t np = MODE == 0 ? 0u : piece_count(dmax - dmin, piece_len);
            if (MODE == 2) {
                const SegQ q{s.origin, s.dir, dmin, dmax, mid, rsf, rsr, kEndInMedium | (i == 0 ? flags : 0u),
                             s.path_length};
                store_segrec(out + c.recs, q, s.sw, s.throughput * att, pix);
                npl[c.recs] = uint8_t(np);
            }
            c.recs++;
            c.pieces += np;
            c.len += dmax - dmin;
        }
        if (segs) {
            if (qbt == T_BEAM_SHORT) att *= segs[i].att / segs[i].pf;
            else att *= segs[i].att;
            if (!positive(att)) break;
            rsf *= segs[i].pf;
            rsr *= segs[i].pr;
        } else {
            att *= eval_attenuation(m, dmax - dmin);
            if (!positive(att)) break;
            float srr;
            const float sff = ray_sample_pdf(m, dmin, dmax, i == 0 ? flags : 0u, &srr);
            rsf *= sff;
            rsr *= srr;
        }
    }
    return c;
}

// Wave-aggregated allocation of c < 2^BITS slots per lane: one atomic per wave.
template <int BITS>
__device__ __forceinline__ uint32_t wave_alloc(uint32_t c, uint32_t* counter)
{
    const uint32_t lane = __lane_id();
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    uint32_t pre = 0, tot = 0;
    for (int bit = 0; bit < BITS; bit++) {
        const uint64_t m = __ballot((c >> bit) & 1u);
        pre += uint32_t(__popcll(m & lt)) << bit;
        tot += uint32_t(__popcll(m)) << bit;
    }
    const int leader = __ffsll((unsigned long long)__ballot(1)) - 1;
    uint32_t base = 0;
    if (int(lane) == leader && tot) base = atomicAdd(counter, tot);
    base = __shfl(base, leader);
    return base + pre;
}

// Wave-aggregated allocation of one block for the whole wave (c < 2^BITS per lane): returns
// the block's first slot and its size (the same in every lane).
template <int BITS>
__device__ __forceinline__ uint32_t wave_alloc_block(uint32_t c, uint32_t* counter, uint32_t& tot)
{
    const uint32_t lane = __lane_id();
    tot = 0;
    for (int bit = 0; bit < BITS; bit++) tot += uint32_t(__popcll(__ballot((c >> bit) & 1u))) << bit;
    const int leader = __ffsll((unsigned long long)__ballot(1)) - 1;
    uint32_t base = 0;
    if (int(lane) == leader && tot) base = atomicAdd(counter, tot);
    return __shfl(base, leader);
}

// kernel/functs_camera.cuh:41-103
__device__ __forceinline__ int generate_camera_sample(const DScene& sc, const IterParams& p, uint32_t pix, Sub& s,
                                                      float& sx, float& sy, Rng& rng)
{
    const DCamera& cam = sc.cam;
    const uint32_t rx = uint32_t(cam.res_x);
    const uint32_t x = pix % rx, y = pix / rx;
    const float j0 = rng.next(), j1 = rng.next();
    sx = float(x) + j0;
    sy = float(y) + j1;
    const f3 dir = raster_dir(cam, sx, sy);
    const float cos_at = dot(cam.forward, dir);
    const float ipd = cam.ipd / cos_at;
    const float cam_pdf_w = (ipd * ipd) / cos_at;
    s.origin = cam.origin;
    s.dir = dir;
    s.throughput = mk(1.f);
    s.path_length = 1;
    s.specular_path = true;
    s.infinite = false;
    s.sw.w.d_shared = float(p.n_light) / cam_pdf_w;
    s.sw.w.d_bpt = 0.f;
    s.sw.w.d_pde = 0.f;
    s.sw.w.for_inv = s.sw.w.rev_inv = s.sw.w.for_ratio = s.sw.w.rev_ratio = 0.f;
    s.sw.w.prev_specular = 0;
    s.sw.w.prev_in_medium = 0;
    s.sw.last_sin = s.sw.d_bpt_a = s.sw.d_bpt_b = s.sw.d_pde_a = s.sw.d_pde_b = 0.f;
    return init_stack(sc, s.stack, cam.mat_id, cam.med_id);
}

// Camera paths start in 8x8-pixel tiles per wave (16x16 per block): the queue of live paths, and
// with it the bounce kernels' waves and the segment-query pieces they write, then hold compact
// pixel tiles, whose rays are more coherent than a 64-pixel row. Only the order changes; every
// pixel keeps its own random streams.
// vb: the block's index within its job (blockIdx.x for a single job)
__device__ __forceinline__ void camera_generate_body(const DScene& sc, const Bufs& b, uint32_t res_x, uint32_t tiles_x,
                                                     uint32_t vb)
{
    const IterParams& p = *b.params;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t x = (vb % tiles_x) * 16u + (wave & 1u) * 8u + (lane & 7u);
    const uint32_t y = (vb / tiles_x) * 16u + (wave >> 1) * 8u + (lane >> 3);
    const uint32_t pix = y * res_x + x;
    if (x >= res_x || pix >= p.n_pixels) return;
    Rng rng(kDomainCamera, pix, p.iter_key, 0, P_JITTER);
    Sub s;
    float sx, sy;
    const int ret = generate_camera_sample(sc, p, pix, s, sx, sy, rng);
    PathState st;
    st.flags = ((p.techniques & (T_PREVIOUS | T_COMPATIBLE)) != 0) ? PF_ONLY_SPEC_SURF : 0u;
    st.color = mk(0.f);
    st.sx = sx;
    st.sy = sy;
    save_sub(st, s);
    b.cstate[pix] = st;
    const bool alive = ret == S_OK;
    if (alive) store_stack(b.cstack, pix, s.stack);
    else stat_add(b.stats, STAT_ERRORS, 1);
    queue_append(alive, pix, b.queue[0], &b.qcount[64]);
}

__global__ void __launch_bounds__(256) k_camera_generate(DScene sc, Bufs b, uint32_t res_x, uint32_t tiles_x)
{
    camera_generate_body(sc, b, res_x, tiles_x, blockIdx.x);
}

__global__ void __launch_bounds__(256) k_camera_generate_batch(const BatchJob* jobs, const uint2* bmap)
{
    const uint2 m = bmap[blockIdx.x];
    const BatchJob& J = jobs[m.x];
    camera_generate_body(J.sc, J.b, J.res_x, J.tiles_x, m.y);
}

struct CamCounters {
    uint32_t conn = 0, pq = 0, sq = 0, acc = 0, pieces = 0;
};

// One segment of kernel/trace_camera.cu:79-382. Returns whether the path continues. LOG adds
// the MIS weight log of the sampled camera paths (Bufs::mislog).
template <bool LOG>
__device__ bool camera_step(const DScene& sc, const IterParams& p, const Bufs& b, uint32_t pix, int bounce, Sub& s,
                            uint32_t& pflags, f3& color, int& ret, CamCounters& cc, float& guide_den)
{
    const bool origin_in_medium = (pflags & PF_ORIGIN_IN_MEDIUM) != 0;
    bool only_spec_surf = (pflags & PF_ONLY_SPEC_SURF) != 0;
    const uint32_t seg = s.path_length - 1;
    Rng mrng(kDomainCamera, pix, p.iter_key, seg, P_MEDIUM);
    Hit is;
    hit_reset(is, kInf);
    VolSeg segs[kMaxSegments];
    LiteSeg lite[kMaxSegments];
    int nseg, nlite;
    bool hit;
    const unsigned rf = origin_in_medium ? kOriginInMedium : 0u;
    ret = intersect(sc, hit, s.origin, s.dir, is, s.stack, rf, segs, nseg, lite, nlite, &mrng);
    if (ret != S_OK) return false;
    const bool use_full = is.on_surface() || p.qbt == T_BEAM_SHORT || !hit;
    bool ok = true;
    if ((p.build_pb2d && b.maps->med.n > 0) || (p.build_bb1d && b.maps->beam.n > 0)) {
        const VolSeg* sg = use_full ? segs : nullptr;
        const LiteSeg* lt = use_full ? nullptr : lite;
        const int ns = use_full ? nseg : nlite;
        const SegCount c0 = segment_queries<0>(sc, sg, lt, ns, rf, p.qbt, 0.f, s, pix, nullptr, nullptr);
        // piece length: the tuned length, stretched so that one ray makes at most kPiecesPerRay
        // pieces plus one per segment
        const float pl = bounce < p.bundle_bounces ? p.bundle_piece_len : p.piece_len;
        const float plen = pl > 0.f ? fmaxf(pl, c0.len / float(kPiecesPerRay)) : 0.f;
        const SegCount c1 = segment_queries<1>(sc, sg, lt, ns, rf, p.qbt, plen, s, pix, nullptr, nullptr);
        const uint32_t at = wave_alloc<4>(c1.recs, &b.qcount[128 + bounce]);
        uint32_t ptot;
        const uint32_t pbase = wave_alloc_block<6>(c1.pieces, &b.qcount[320 + bounce], ptot);
        if (c1.recs) {
            // an overflowing wave writes no pieces at all (the merge kernels skip pieces whose
            // record index is out of range)
            if (__ballot(at + c1.recs > b.seg_cap) || pbase + ptot > b.piece_cap) {
                ret = S_ERR_SEGMENTS;
                return false;
            }
            uint8_t npl[16];
            segment_queries<2>(sc, sg, lt, ns, rf, p.qbt, plen, s, pix, b.segq + at, npl);
            // The wave's pieces go to its block piece-index-major: first every lane's first
            // piece, then every lane's second, ... so that the merge kernel, which hands 64
            // consecutive pieces to one wave, gives a wave the same stretch of 64 neighbouring
            // rays instead of the whole length of one or two rays.
            const uint32_t lane = __lane_id();
            const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
            // Bounces walked per bundle also list each j-th run (every lane's j-th piece: 64
            // neighbouring rays at similar depth) as one bundle. Runs are at most the wave's
            // pieces, so the bundle list fits the piece capacity.
            // (Lanes without a query are inactive here: the run count is the largest piece count
            // of the active lanes, found bit by bit with ballots, and the first active lane
            // allocates.)
            uint32_t ubase = 0;
            const bool bundles = b.units != nullptr && bounce < p.bundle_bounces;
            const int uleader = __ffsll((unsigned long long)__ballot(1)) - 1;
            if (bundles) {
                uint32_t ng = 0;
                for (int bit = 5; bit >= 0; bit--)
                    if (__ballot(c1.pieces >= (ng | (1u << bit)))) ng |= 1u << bit;
                if (int(lane) == uleader && ng) ubase = atomicAdd(&b.qcount[384 + bounce], ng);
                ubase = __shfl(ubase, uleader);
            }
            uint32_t start = 0, rec = 0, k = 0;
            for (uint32_t j = 0;; j++) {
                const uint64_t mj = __ballot(c1.pieces > j);
                if (!mj) break;
                if (bundles && int(lane) == uleader && ubase + j < b.piece_cap)
                    b.units[ubase + j] = make_uint2(pbase + start, uint32_t(__popcll(mj)));
                if (c1.pieces > j) {
                    b.pieces[pbase + start + uint32_t(__popcll(mj & below))] =
                        make_uint2(at + rec, k | (uint32_t(npl[rec]) << 16));
                    if (++k == npl[rec]) {
                        k = 0;
                        rec++;
                    }
                }
                start += uint32_t(__popcll(mj));
            }
            cc.sq += c1.recs;
            cc.pieces += c1.pieces;
        }
    }
    if (!hit) {
        if (s.path_length < p.min_len) return false;
        if (sc.background_light < 0) return false;
        if (sc.media[sc.global_medium].has_attenuation) return false;
    }
    float rsf = 1.f, rsr = 1.f;
    if (nseg) {
        rsf = acc_for(segs, nseg);
        rsr = acc_rev(segs, nseg);
        s.throughput *= acc_att(segs, nseg) / rsf;
    }
    if (black_or_neg(s.throughput)) return false;
    Weights& w = s.sw.w;
    if (s.path_length > 1u) {
        const float fi = pde_factor(p, w.prev_in_medium, w.prev_specular, 1.f / rsr, w.rev_ratio, w.for_inv,
                                    w.for_ratio, s.sw.last_sin);
        w.d_bpt = s.sw.d_bpt_a * fi + s.sw.d_bpt_b;
        w.d_pde = s.sw.d_pde_a * fi + s.sw.d_pde_b;
    }
    w.d_shared /= rsf;
    w.d_bpt /= rsf;
    w.d_pde /= rsf;
    if (!hit) {
        color += s.throughput * light_radiance_mis(sc, p, sc.lights[sc.background_light], s, rsr);
        return false;
    }
    Bsdf bs;
    bsdf_setup(bs, s.dir, is, sc, false, relative_ior(sc, is, s.stack));
    if (!bs.valid()) return false;
    float for_ratio = 0.f, rev_ratio = 0.f;
    if (bs.in_medium()) for_ratio = rev_ratio = 1.f / sc.media[bs.id].min_pos_att;
    const f3 hitp = s.origin + s.dir * is.dist;
    bs.pos = hitp;
    pflags = is.in_medium() ? (pflags | PF_ORIGIN_IN_MEDIUM) : (pflags & ~PF_ORIGIN_IN_MEDIUM);
    {
        const float d = bs.d_factor();
        w.d_shared *= is.dist * is.dist;
        w.d_shared /= d;
        w.d_bpt /= d;
        w.d_pde /= d;
        w.for_inv = 1.f / rsf;
        w.rev_inv = 1.f / rsr;
        w.for_ratio = for_ratio;
        w.rev_ratio = rev_ratio;
    }
    const bool logged = LOG && pix % b.mis_stride == 0;
    if (is.light >= 0) {
        if (s.path_length < p.min_len) return false;
        float wl = -1.f;
        color += s.throughput * light_radiance_mis(sc, p, sc.lights[is.light], s, rsr, logged ? &wl : nullptr);
        if (logged && wl >= 0.f && sc.lights[is.light].type == L_AREA)
            mis_log(b, MT_LIGHT_HIT, s.path_length, 0, pix, uint32_t(is.light), wl);
        return false;
    }
    if (s.path_length >= p.max_len) return false;
    if (p.do_bpt && !bs.delta() && s.path_length + 1 >= p.min_len && sc.n_lights > 0 &&
        (bs.in_medium() || !only_spec_surf)) {
        f3 ct;
        Rng nrng(kDomainCamera, pix, p.iter_key, seg, P_NEE);
        float wl = -1.f;
        int lid = 0;
        ret = direct_illumination(sc, p, ct, s, hitp, bs, nrng, logged ? &wl : nullptr, &lid);
        if (ret != S_OK) return false;
        if (logged && wl >= 0.f && sc.lights[lid].type == L_AREA)
            mis_log(b, MT_NEE, s.path_length + 1, 1, pix, uint32_t(lid), wl);
        color += s.throughput * ct;
    }
    if (p.do_bpt && !bs.delta() && (bs.in_medium() || !only_spec_surf)) {
        bool behind = false;
        if (bs.in_medium() && s.stack.n > 0) {
            const int m = s.stack.mat[s.stack.n - 1];
            if (m >= 0 && sc.mats[m].real) behind = true;
        }
        Rng brng(kDomainCamera, pix, p.iter_key, seg, P_BPT_PICK);
        const uint32_t lp = uint32_t(brng.next() * float(p.n_light));
        const uint32_t nv = b.nvert[lp];
        const VertexRec* vr = b.verts + size_t(lp) * p.vstride;
        for (uint32_t k = 0; k < nv; k++) {
            LVertex lv;
            load_vertex(vr + k, lv);
            const uint32_t len = lv.path_length + 1 + s.path_length;
            if (len > p.max_len) break;
            const bool small = len < p.min_len;
            const bool diff_media =
                lv.bsdf.in_medium() && bs.in_medium() && lv.bsdf.id != bs.id && (lv.behind_surf || behind);
            if (!(small || diff_media)) {
                f3 ct;
                cc.conn++;
                float wl = -1.f;
                ret = connect_vertices(sc, p, ct, lv, bs, hitp, s, logged ? &wl : nullptr);
                if (ret != S_OK) return false;
                // logged when the light path's vertices before this one were all stored
                if (logged && wl >= 0.f && lv.path_length == k + 1)
                    mis_log(b, MT_BPT, lv.path_length + s.path_length + 1, lv.path_length + 1, pix, lp, wl);
                color += s.throughput * lv.throughput * ct;
            }
        }
    }
    if (p.build_surf && b.maps->surf.n > 0 && bs.on_surface() && !bs.delta() && !only_spec_surf) {
        f3 acc = mk(0.f);
        cc.pq++;
        ok = bvh_point_query(b.maps->surf, hitp, [&](uint32_t k) {
            const float4 pt = b.maps->surf_pts[k];
            if (!pred_point_point(hitp, mk(pt.x, pt.y, pt.z), p.r2_surf)) return;
            LVertex lv;
            const uint32_t slot = uint32_t(__float_as_int(pt.w));
            load_vertex(&b.verts[slot], lv);
            cc.acc++;
            float wl = -1.f;
            eval_pp(sc, p, acc, s, bs, lv, logged ? &wl : nullptr);
            if (logged && wl >= 0.f && slot % p.vstride + 1 == lv.path_length)
                mis_log(b, MT_SURF, lv.path_length + s.path_length, lv.path_length, pix, slot / p.vstride, wl);
        });
        color += s.throughput * p.norm_surf * acc;
    }
    if (p.build_pp3d && b.maps->med.n > 0 && bs.in_medium() && !bs.delta()) {
        f3 acc = mk(0.f);
        cc.pq++;
        ok &= bvh_point_query(b.maps->med, hitp, [&](uint32_t k) {
            const float4 pt = b.maps->med_pts[k];
            if (!pred_point_point(hitp, mk(pt.x, pt.y, pt.z), p.r2_pp3d)) return;
            LVertex lv;
            const uint32_t slot = uint32_t(__float_as_int(pt.w));
            load_vertex(&b.verts[slot], lv);
            cc.acc++;
            float wl = -1.f;
            eval_pp(sc, p, acc, s, bs, lv, logged ? &wl : nullptr);
            if (logged && wl >= 0.f && slot % p.vstride + 1 == lv.path_length)
                mis_log(b, MT_PP3D, lv.path_length + s.path_length, lv.path_length, pix, slot / p.vstride, wl);
        });
        color += s.throughput * p.norm_pp3d * acc;
    }
    if (p.build_bp2d && b.maps->beam.n > 0 && bs.in_medium()) {
        f3 acc = mk(0.f);
        cc.pq++;
        ok &= bvh_point_query(b.maps->beam, hitp, [&](uint32_t k) {
            const float4 g0 = b.maps->beam_geo[3 * k], g1 = b.maps->beam_geo[3 * k + 1];
            const f3 bo = mk(g0.x, g0.y, g0.z), bd = mk(g1.x, g1.y, g1.z);
            if (!pred_point_beam(bo, bd, g0.w, hitp, p.r2_bp2d)) return; // program_bp2d.cu:31-45
            float idist, ir2;
            if (!pred_bre(idist, ir2, bo, bd, 0.f, g0.w, hitp, p.r2_bp2d)) return;
            const float4 g2 = b.maps->beam_geo[3 * k + 2];
            if (!(idist >= g2.x && idist < g2.y)) return; // another sub-beam of this beam owns the point
            LBeam lb;
            load_beam(&b.beams[__float_as_int(g1.w)], lb);
            cc.acc++;
            float wl = -1.f;
            eval_bp2d(sc, p, acc, s, bs, lb, idist, ir2, logged ? &wl : nullptr);
            if (logged && wl >= 0.f)
                mis_log(b, MT_BP2D, lb.path_length + s.path_length, lb.path_length, pix, lb.path_idx, wl);
        });
        color += s.throughput * p.norm_bp2d * acc;
    }
    if (!ok) {
        ret = S_ERR_STACK;
        return false;
    }
    bool cont;
    Rng srng(kDomainCamera, pix, p.iter_key, seg, P_SCATTER);
    ret = sample_scattering(sc, p, cont, bs, hitp, is, s, srng, b.gsamp ? &guide_den : nullptr);
    if (ret != S_OK) return false;
    if (!cont) return false;
    if (bs.on_surface()) {
        if (!s.sw.w.prev_specular && (only_spec_surf || (p.techniques & T_SPECULAR_ONLY))) return false;
    } else {
        if (p.techniques & T_SPECULAR_ONLY) return false;
        if (only_spec_surf) {
            if (p.techniques & T_COMPATIBLE) pflags &= ~PF_ONLY_SPEC_SURF;
            else return false;
        }
    }
    return true;
}

// Training data of the learned proposal. A scattering of a camera path in a medium or at a
// guidable surface that continues is credited with everything the path collects after it:
// emission, next-event estimation, connections and point merges at later vertices (the path
// colour) and the P-B2D / B-B1D merges along its later segments (the merge colour). With
// throughput T before the scattering, scattering function f, cosine cos, continuation
// probability cp and sampling pdf p, the path carries T f cos / (p cp) onward, so
// lum(collected after) cp / lum(T f cos) estimates L_i / p: the incident radiance along the
// sampled direction (the part the path's later vertices collect under their MIS weights) over
// its pdf, the weight of the KL gradient that fits the proposal's radiance mixture to L_i. The sample is (position, scattered direction) with
// the arrival inputs zero: the mixture depends on the position only.
//
// The scattering at vertex k waits in the path's pending slot until the next bounce, where the
// merges along the segment it sampled have run; it then enters the pool with the snapshot
// lum(path colour after vertex k) + lum(merge colour before segment k + 1's merges).
// k_guide_finalise turns snapshots into weights once the iteration's colours are final.
// Inlined: a call taking the kernel's scene and buffer arguments by reference would copy them
// to scratch memory for the whole bounce kernel.
__device__ __forceinline__ void guide_record_sample(const DScene& sc, const Bufs& b, uint32_t pix, f3 color,
                                                 bool scattered, f3 origin, f3 dir, float den)
{
    float* pend = b.gpend + size_t(pix) * kGuidePend;
    bool emit = false;
    float rec[kGuideRaw];
    if (pend[9] > 0.f) {
        for (int k = 0; k < 10; k++) rec[k] = pend[k];
        rec[10] = pend[10] + luminance(mk(b.mcolor[3 * pix], b.mcolor[3 * pix + 1], b.mcolor[3 * pix + 2]));
        rec[11] = __uint_as_float(pix);
        emit = true;
        pend[9] = 0.f;
    }
    const uint64_t m = __ballot(emit);
    if (m) { // wave-aggregated append
        const uint32_t lane = __lane_id();
        const int leader = __ffsll((unsigned long long)m) - 1;
        uint32_t base = 0;
        if (int(lane) == leader) base = atomicAdd(b.gcount, uint32_t(__popcll(m)));
        base = __shfl(base, leader);
        const uint32_t at = base + uint32_t(__popcll(m & (lane == 0 ? 0ull : (~0ull >> (64 - lane)))));
        if (emit && at < b.gcap)
            for (int k = 0; k < kGuideRaw; k++) b.graw[size_t(at) * kGuideRaw + k] = rec[k];
    }
    // a scattering in a medium or at a guidable surface that continues becomes the pending record
    if (scattered && den > 0.f) {
        const DGuide& g = sc.guide;
        const float inv_res = 1.f / float(g.res);
        pend[0] = (origin.x - g.lo.x) * g.inv_cell.x * inv_res;
        pend[1] = (origin.y - g.lo.y) * g.inv_cell.y * inv_res;
        pend[2] = (origin.z - g.lo.z) * g.inv_cell.z * inv_res;
        pend[3] = pend[4] = pend[5] = 0.f;
        pend[6] = dir.x;
        pend[7] = dir.y;
        pend[8] = dir.z;
        pend[9] = den;
        pend[10] = luminance(color);
    }
}

// pool of raw records -> training samples (10 floats: position, arrival, scattered, weight) with
// a positive weight, written to the sample ring after the splat, when path and merge colours are
// final. Counters: gcount[1] samples kept this iteration, gcount[2] the ring's write position
// (modulo gring), gcount[3] its fill. At most gring samples are written per iteration, so no ring
// slot is written twice in one launch.
__global__ void __launch_bounds__(256) k_guide_finalise(Bufs b)
{
    const uint32_t n = min(b.gcount[0], b.gcap);
    const uint32_t start = b.gcount[2];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float* r = b.graw + size_t(i) * kGuideRaw;
        const uint32_t pix = __float_as_uint(r[11]);
        const float fin = luminance(b.cstate[pix].color) +
                          luminance(mk(b.mcolor[3 * pix], b.mcolor[3 * pix + 1], b.mcolor[3 * pix + 2]));
        const float w = (fin - r[10]) / r[9];
        const bool keep = w > 0.f && isfinite(w);
        const uint64_t m = __ballot(keep);
        if (!m) continue;
        const uint32_t lane = __lane_id();
        const int leader = __ffsll((unsigned long long)m) - 1;
        uint32_t base = 0;
        if (int(lane) == leader) base = atomicAdd(b.gcount + 1, uint32_t(__popcll(m)));
        base = __shfl(base, leader);
        const uint32_t at = base + uint32_t(__popcll(m & (lane == 0 ? 0ull : (~0ull >> (64 - lane)))));
        if (keep && at < b.gring) {
            float* o = b.gsamp + size_t((uint64_t(start) + at) % b.gring) * 10;
            for (int k = 0; k < 9; k++) o[k] = r[k];
            o[9] = w;
        }
    }
}

// advances the ring past this iteration's samples (one thread, after k_guide_finalise)
__global__ void k_guide_ring_advance(Bufs b)
{
    const uint32_t kept = min(b.gcount[1], b.gring);
    b.gcount[2] = uint32_t((uint64_t(b.gcount[2]) + kept) % b.gring);
    b.gcount[3] = min(b.gcount[3] + kept, b.gring);
}

template <bool LOG>
__device__ __forceinline__ void camera_bounce_body(const DScene& sc, const Bufs& b, int bounce, uint32_t vb)
{
    const IterParams& p = *b.params;
    const uint32_t i = vb * blockDim.x + threadIdx.x;
    const uint32_t n = b.qcount[64 + bounce];
    if (i >= n) return;
    const uint32_t pix = b.queue[bounce & 1][i];
    PathState st = b.cstate[pix];
    Sub s;
    load_sub(st, s);
    load_stack(b.cstack, pix, s.stack);
    uint32_t pflags = st.flags;
    f3 color = st.color;
    int ret = S_OK;
    CamCounters cc;
    float guide_den = 0.f;
    bool cont = camera_step<LOG>(sc, p, b, pix, bounce, s, pflags, color, ret, cc, guide_den);
    if (ret != S_OK) {
        cont = false;
        stat_add(b.stats, STAT_ERRORS, 1);
    }
    if (b.gsamp) guide_record_sample(sc, b, pix, color, cont, s.origin, s.dir, guide_den);
    stat_add(b.stats, STAT_BPT_CONN, cc.conn);
    stat_add(b.stats, STAT_POINT_Q, cc.pq);
    stat_add(b.stats, STAT_SEG_Q, cc.sq);
    stat_add(b.stats, STAT_SEG_PIECES, cc.pieces);
    stat_add(b.stats, STAT_ACCEPT, cc.acc);
    st.color = color;
    if (cont) {
        s.path_length++;
        st.flags = pflags;
        save_sub(st, s);
        store_stack(b.cstack, pix, s.stack);
    }
    b.cstate[pix] = st;
    queue_append(cont, pix, b.queue[(bounce + 1) & 1], &b.qcount[64 + bounce + 1]);
}

template <bool LOG>
__global__ void __launch_bounds__(256, 4) k_camera_bounce(DScene sc, Bufs b, int bounce)
{
    camera_bounce_body<LOG>(sc, b, bounce, blockIdx.x);
}

__global__ void __launch_bounds__(256, 4) k_camera_bounce_batch(const BatchJob* jobs, const uint2* bmap, int bounce)
{
    const uint2 m = bmap[blockIdx.x];
    camera_bounce_body<false>(jobs[m.x].sc, jobs[m.x].b, bounce, m.y);
}

// adds every camera path's colour at its jittered raster position (frame_buffer.cuh:35-44)
__device__ __forceinline__ void camera_splat_body(const Bufs& b, float res_x, float res_y, uint32_t vb)
{
    const IterParams& p = *b.params;
    const uint32_t pix = vb * 256u + threadIdx.x;
    if (pix >= p.n_pixels) return;
    const PathState& st = b.cstate[pix];
    if (!(st.sx >= 0.f && st.sx < res_x && st.sy >= 0.f && st.sy < res_y)) return;
    const size_t tgt = size_t(int(st.sx)) + size_t(int(st.sy)) * size_t(res_x);
    const f3 c = st.color + mk(b.mcolor[3 * pix], b.mcolor[3 * pix + 1], b.mcolor[3 * pix + 2]);
    atomicAdd(&b.fb32[3 * tgt], c.x);
    atomicAdd(&b.fb32[3 * tgt + 1], c.y);
    atomicAdd(&b.fb32[3 * tgt + 2], c.z);
}

__global__ void __launch_bounds__(256) k_camera_splat(Bufs b, float res_x, float res_y)
{
    camera_splat_body(b, res_x, res_y, blockIdx.x);
}

// folds the FP32 iteration buffer into the FP64 cross-iteration accumulator, unless the iteration
// overflowed a photon list (then it is discarded and rendered again)
__global__ void __launch_bounds__(256) k_accumulate(const float* fb32, double* accum, uint32_t n, const uint32_t* ovf)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && *ovf == 0u) accum[i] += double(fb32[i]);
}

// splat (phase 0, a block per 256 pixels of each job) and accumulate (phase 1, a block per 256
// values of each job) of every job; bmap gives each block its job and its block within the job.
// A job's accumulate must see all of its splat, so the phases are two launches.
__global__ void __launch_bounds__(256) k_finish_batch(const BatchJob* jobs, const uint2* bmap, int phase)
{
    const uint2 m = bmap[blockIdx.x];
    const BatchJob& J = jobs[m.x];
    if (phase == 0) {
        camera_splat_body(J.b, J.fres_x, J.fres_y, m.y);
    } else {
        const uint32_t i = m.y * 256u + threadIdx.x;
        const uint32_t n = 3u * J.b.params->n_pixels;
        if (i < n && *J.ovf == 0u) J.b.accum[i] += double(J.b.fb32[i]);
    }
}

void launch_camera_generate(const DScene& sc, const IterParams& p, const Bufs& b, hipStream_t s)
{
    const uint32_t res_x = uint32_t(sc.cam.res_x);
    const uint32_t res_y = (p.n_pixels + res_x - 1) / res_x;
    const uint32_t tiles_x = (res_x + 15) / 16, tiles_y = (res_y + 15) / 16;
    hipLaunchKernelGGL(k_camera_generate, dim3(tiles_x * tiles_y), dim3(256), 0, s, sc, b, res_x, tiles_x);
}

void launch_camera_generate_batch(const BatchJob* jobs, const uint2* bmap, uint32_t blocks, hipStream_t s)
{
    if (blocks) hipLaunchKernelGGL(k_camera_generate_batch, dim3(blocks), dim3(256), 0, s, jobs, bmap);
}

void launch_camera_bounce_batch(const BatchJob* jobs, const uint2* bmap, uint32_t blocks, int bounce, hipStream_t s)
{
    if (blocks) hipLaunchKernelGGL(k_camera_bounce_batch, dim3(blocks), dim3(256), 0, s, jobs, bmap, bounce);
}

void launch_camera_finish_batch(const BatchJob* jobs, const uint2* bmap_pix, uint32_t blocks_pix, const uint2* bmap_acc,
                                uint32_t blocks_acc, hipStream_t s)
{
    if (blocks_pix) hipLaunchKernelGGL(k_finish_batch, dim3(blocks_pix), dim3(256), 0, s, jobs, bmap_pix, 0);
    if (blocks_acc) hipLaunchKernelGGL(k_finish_batch, dim3(blocks_acc), dim3(256), 0, s, jobs, bmap_acc, 1);
}

void launch_camera_bounce(const DScene& sc, const IterParams& p, const Bufs& b, int bounce, hipStream_t s)
{
    if (b.mislog)
        hipLaunchKernelGGL(k_camera_bounce<true>, dim3((p.n_pixels + p.block - 1) / p.block), dim3(p.block), 0, s, sc,
                           b, bounce);
    else
        hipLaunchKernelGGL(k_camera_bounce<false>, dim3((p.n_pixels + p.block - 1) / p.block), dim3(p.block), 0, s, sc,
                           b, bounce);
}

void launch_camera_finish(const IterParams& p, const Bufs& b, float res_x, float res_y, const uint32_t* ovf,
                          hipStream_t s)
{
    hipLaunchKernelGGL(k_camera_splat, dim3((p.n_pixels + 255) / 256), dim3(256), 0, s, b, res_x, res_y);
    const uint32_t n = 3 * p.n_pixels;
    hipLaunchKernelGGL(k_accumulate, dim3((n + 255) / 256), dim3(256), 0, s, b.fb32, b.accum, n, ovf);
    if (b.gsamp) {
        hipLaunchKernelGGL(k_guide_finalise, dim3(1024), dim3(256), 0, s, b);
        hipLaunchKernelGGL(k_guide_ring_advance, dim3(1), dim3(1), 0, s, b);
    }
}

} // namespace upbp
