// Segment merge kernels: P-B2D (camera segment x medium photon point) and B-B1D (camera
// segment x photon sub-beam) for the segment queries one camera bounce appended.
//
// Semantics: nolmoonen/gpuupbp gpuupbp/kernel/pde/functs_pb2d.cuh:57-342 with
// program_pb2d.cu:31-46, and functs_bb1d.cuh:57-397 with program_bb1d.cu:31-48.
//
// Execution: a persistent grid over query pieces (a long camera segment is walked as several
// independent pieces, so no single query serialises the launch). Each lane owns one piece at
// a time and walks the LBVH one node per loop trip; a lane whose piece is done takes the
// next one from a global work counter (one atomic per wave), so lanes stay busy while
// pieces differ in cost by orders of magnitude. The work is split in three stages joined
// by per-wave LDS queues, so that each stage runs with the whole wave converged: the node
// step (box tests, stack) appends the leaves it hits as candidates; 64 candidates at a time
// are tested with the exact predicate and piece ownership; 64 accepted (query, photon) pairs
// at a time are evaluated (attenuation, phase function, both-side MIS updates) and their colour
// gathered per query in LDS, then added to the query's camera path with float atomics.
#include "estimators.h"
#include "kernels.h"

#include <algorithm>

namespace upbp {

namespace {

constexpr int kMergeBlock = 256;
constexpr int kMergeWaves = kMergeBlock / 64;
constexpr uint32_t kCandCap = 64 + 64 * kWide; // < 64 pending + up to kWide leaves per lane in one node step
constexpr uint32_t kPairCap = 128; // < 64 pending + at most one accept per lane in a test batch
// B-B1D bundles list leaves as node << 3 | child: trees over fewer primitives than this (so fewer
// wide nodes) walk bundles, larger ones walk per piece
constexpr uint32_t kBundleMaxPrims = 1u << 28;

// The piece of a candidate or stolen subtree travels as (k, n) packed in 16 bits (k | n << 6,
// k < n <= 32); the camera range it owns is recomputed from the query record, which the
// consumer loads anyway. Small entries keep the LDS of a 4-wave block under 32 KB, so that
// five blocks fit a CU.
__device__ __forceinline__ uint32_t pack_kn(uint32_t k, uint32_t n) { return k | (n << 6); }

struct CandQueue {
    uint32_t q[kCandCap];    // segment query index
    uint32_t leaf[kCandCap]; // sorted LBVH leaf
    uint16_t kn[kCandCap];   // piece of the query
};

// subtrees handed from busy lanes to idle lanes of the same wave
struct Steal {
    uint32_t q[64], node[64];
    uint16_t kn[64];
};

struct PairQueue {
    uint32_t q[kPairCap];    // segment query index
    uint32_t slot[kPairCap]; // light vertex slot (P-B2D) or beam slot (B-B1D)
    float x[kPairCap], y[kPairCap], z[kPairCap], w[kPairCap];
};

// Colour of accepted pairs, gathered per segment query before it goes to the camera path's
// merge colour: 64 direct-mapped slots per wave keyed by the query. A query's accepted pairs
// mostly fall in the few batches while its piece is walked, so most of its float atomics to
// HBM become LDS adds; a slot taken by another query is written back first.
struct ColorCache {
    uint32_t tag[64];   // segment query index + 1 (0: empty)
    uint32_t pix[64];   // its camera path
    uint32_t claim[64]; // lane that takes the slot in this batch
    float r[64], g[64], b[64];
};

__device__ __forceinline__ uint32_t cache_slot(uint32_t q) { return (q * 2654435761u) >> 26; }

__device__ __forceinline__ void add_color(float* mcolor, uint32_t pix, float r, float g, float b)
{
    if (r != 0.f) atomicAdd(&mcolor[3 * pix], r);
    if (g != 0.f) atomicAdd(&mcolor[3 * pix + 1], g);
    if (b != 0.f) atomicAdd(&mcolor[3 * pix + 2], b);
}

// orders this wave's LDS writes before the reads that follow (all lanes of one wave)
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ uint64_t lanes_below()
{
    const uint32_t lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__device__ __forceinline__ void pq_push(PairQueue& Q, uint32_t& qn, bool keep, uint32_t q, uint32_t slot, float x,
                                        float y, float z, float w)
{
    const uint64_t m = __ballot(keep);
    if (keep) {
        const uint32_t at = qn + uint32_t(__popcll(m & lanes_below()));
        Q.q[at] = q;
        Q.slot[at] = slot;
        Q.x[at] = x;
        Q.y[at] = y;
        Q.z[at] = z;
        Q.w[at] = w;
    }
    qn += uint32_t(__popcll(m));
}

__device__ __forceinline__ void cq_push(CandQueue& C, uint32_t& cn, bool keep, uint32_t q, uint32_t leaf, uint32_t kn)
{
    const uint64_t m = __ballot(keep);
    if (keep) {
        const uint32_t at = cn + uint32_t(__popcll(m & lanes_below()));
        C.q[at] = q;
        C.leaf[at] = leaf;
        C.kn[at] = uint16_t(kn);
    }
    cn += uint32_t(__popcll(m));
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// B-B1D test of one query piece against one sub-beam record (g0: beam origin, length; g1:
// direction, beam slot; g2: owned beam range, medium, flags): program_bb1d.cu:31-48 on the whole
// beam, then ownership of the camera parameter by the piece and of the beam parameter by the
// sub-beam.
__device__ __forceinline__ bool bb1d_accept(const IterParams& p, float4 g0, float4 g1, float4 g2, f3 ro, f3 rd,
                                            float dmin, float dmax, float lo, float hi, int med, uint32_t& slot,
                                            float& x, float& y, float& z, float& w)
{
    if (!pred_beam_beam(ro, rd, dmin, dmax, mk(g0.x, g0.y, g0.z), mk(g1.x, g1.y, g1.z), 0.f, g0.w, p.r2_bb1d, z, w, x,
                        y))
        return false;
    if (!(x >= lo && x < hi)) return false;
    if (!(y >= g2.x && y < g2.y)) return false;
    // B-B1D subset of the light paths and the same-medium test (functs_bb1d.cuh:65)
    if (!(__float_as_int(g2.w) & SB_BB1D) || __float_as_int(g2.z) != med) return false;
    slot = uint32_t(__float_as_int(g1.w));
    return true;
}

// Leaf test of one query against one photon (sorted leaf k). Returns whether the pair is
// accepted and its parameters: P-B2D (idist, ir2); B-B1D (t camera, t beam, distance, sin).
// Only the query piece owning the pair's camera parameter [lo, hi) accepts it.
template <int KIND>
__device__ __forceinline__ bool leaf_accept(const IterParams& p, const PhotonMaps& M, const BeamRec* beams, uint32_t k,
                                            f3 ro, f3 rd, float dmin, float dmax, float lo, float hi, int med,
                                            uint32_t& slot, float& x, float& y, float& z, float& w)
{
    if (KIND == 0) { // program_pb2d.cu:31-46
        const float4 pt = M.med_pts[k];
        const f3 ce = mk(pt.x, pt.y, pt.z);
        if (!pred_point_beam(ro, rd, dmax, ce, p.r2_pb2d)) return false;
        if (!pred_bre(x, y, ro, rd, dmin, dmax, ce, p.r2_pb2d)) return false;
        if (!(x >= lo && x < hi)) return false;
        slot = uint32_t(__float_as_int(pt.w));
        z = w = 0.f;
        return true;
    } else if (p.beam_leaf) { // the same test from the sub-beam's reference and its beam's record
        const uint2 rf = M.beam_ref[k];
        const BeamRec& br = beams[rf.x];
        const float4 a = br.a, d = br.b;
        if (!pred_beam_beam(ro, rd, dmin, dmax, mk(a.x, a.y, a.z), mk(d.x, d.y, d.z), 0.f, a.w, p.r2_bb1d, z, w, x,
                            y))
            return false;
        if (!(x >= lo && x < hi)) return false;
        float blo, bhi, ta, tb;
        subbeam_range(a.w, rf.y & 0xffffu, rf.y >> 16, blo, bhi, ta, tb);
        if (!(y >= blo && y < bhi)) return false;
        if (!(uint32_t(__float_as_int(br.d.z)) < p.bb1d_used) || __float_as_int(d.w) != med) return false;
        slot = rf.x;
        return true;
    } else {
        return bb1d_accept(p, M.beam_geo[3 * k], M.beam_geo[3 * k + 1], M.beam_geo[3 * k + 2], ro, rd, dmin, dmax, lo,
                           hi, med, slot, x, y, z, w);
    }
}

// Evaluates the last min(qn, 64) queued pairs with the whole wave: attenuation, phase function,
// both-side MIS updates (and the MIS weight log with LOG), then gathers their colour per query in
// the wave's colour cache. One lane per contended cache slot takes it (writing back the query it
// held); every lane whose query holds its slot then adds in LDS, the others go to HBM.
template <int KIND, bool LOG>
__device__ __forceinline__ void eval_pairs(const DScene& sc, const IterParams& p, const Bufs& b, PairQueue& Q,
                                           ColorCache& K, uint32_t& qn, float norm)
{
    const uint32_t lane = __lane_id();
    const uint32_t m = qn < 64u ? qn : 64u;
    wave_lds_sync();
    bool has = false;
    uint32_t key = 0, cslot = 0, cpix = 0;
    f3 c = mk(0.f);
    if (lane < m) {
        const uint32_t e = qn - m + lane;
        const uint32_t pqi = Q.q[e], pslot = Q.slot[e];
        const float px = Q.x[e], py = Q.y[e], pz = Q.z[e], pw = Q.w[e];
        SegQ q;
        StateWeights cw;
        f3 tatt;
        uint32_t pix;
        load_segrec(b.segq + pqi, q, cw, tatt, pix);
        f3 sum = mk(0.f);
        const bool logged = LOG && pix % b.mis_stride == 0;
        float wl = -1.f;
        if (KIND == 0) {
            LVertex lv;
            load_vertex(&b.verts[pslot], lv);
            eval_pb2d(sc, p, sum, q, cw, lv, px, py, logged ? &wl : nullptr);
            if (logged && wl >= 0.f && pslot % p.vstride + 1 == lv.path_length)
                mis_log(b, MT_PB2D, lv.path_length + q.cam_len, lv.path_length, pix, pslot / p.vstride, wl);
        } else {
            LBeam lb;
            load_beam(&b.beams[pslot], lb);
            eval_bb1d(sc, p, sum, q, cw, lb, pz, pw, px, py, logged ? &wl : nullptr);
            if (logged && wl >= 0.f)
                mis_log(b, MT_BB1D, lb.path_length + q.cam_len, lb.path_length, pix, lb.path_idx, wl);
        }
        c = (tatt * norm) * sum;
        has = c.x != 0.f || c.y != 0.f || c.z != 0.f;
        key = pqi + 1u;
        cslot = cache_slot(pqi);
        cpix = pix;
    }
    wave_lds_sync();
    if (has) K.claim[cslot] = lane;
    wave_lds_sync();
    if (has && K.claim[cslot] == lane && K.tag[cslot] != key) {
        if (K.tag[cslot] != 0u) add_color(b.mcolor, K.pix[cslot], K.r[cslot], K.g[cslot], K.b[cslot]);
        K.tag[cslot] = key;
        K.pix[cslot] = cpix;
        K.r[cslot] = K.g[cslot] = K.b[cslot] = 0.f;
    }
    wave_lds_sync();
    if (has) {
        if (K.tag[cslot] == key) {
            atomicAdd(&K.r[cslot], c.x);
            atomicAdd(&K.g[cslot], c.y);
            atomicAdd(&K.b[cslot], c.z);
        } else {
            add_color(b.mcolor, cpix, c.x, c.y, c.z);
        }
    }
    wave_lds_sync();
    qn -= m;
}

// One job's segment merges of one bounce, by the wave the queues belong to. bid: the block's
// index (its XCD is bid % 8); pass_order: take the pieces as the camera pass wrote them whatever
// the job's piece order (batched launches, which sort nothing); leftovers: take the B-B1D pieces
// the bundle kernel left (pieces_x). LOG adds the MIS weight log of the sampled camera paths
// (Bufs::mislog).
template <int KIND, bool LOG>
__device__ __forceinline__ void seg_merge_body(const DScene& sc, const Bufs& b, int bounce, uint32_t bid,
                                               bool pass_order, bool leftovers, CandQueue& C, PairQueue& Q, Steal& S,
                                               ColorCache& K)
{
    const IterParams& p = *b.params;
    const uint32_t lane = __lane_id();
    K.tag[lane] = 0u;
    K.r[lane] = K.g[lane] = K.b[lane] = 0.f;
    wave_lds_sync();
    // B-B1D walks query pieces from its XCD's region first (then helps the others);
    // P-B2D (few candidates per query) walks whole segments
    const uint32_t n_recs = min(b.qcount[128 + bounce], b.seg_cap);
    const uint32_t total = KIND == 1 ? min(b.qcount[leftovers ? 448 + bounce : 320 + bounce], b.piece_cap) : n_recs;
    uint32_t* head = &b.qcount[192 + 2 * bounce + KIND];
    uint32_t* xw = b.xwork + size_t(bounce) * kXWork;
    const int order = pass_order || leftovers ? 0 : p.piece_order;
    const bool split = KIND == 1 && (order == 1 || order >= 3);
    const uint2* plist = KIND == 1 && (order != 0 || leftovers) ? b.pieces_x : b.pieces;
    uint32_t src = bid % kXcds, tried = 0; // work range being consumed (same in every lane)
    const WBvh& bvh = KIND == 0 ? b.maps->wmed : b.maps->wbeam;
    const float norm = KIND == 0 ? p.norm_pb2d : p.norm_bb1d;

    uint32_t cn = 0, qn = 0; // candidates / pairs queued (same value in every lane)
    bool exhausted = false;  // no pieces left to take (same value in every lane)
    int cur = -1;            // query of this lane's current piece
    uint32_t kn = 0;         // and which piece of it (pack_kn)
    RayBox rb = make_raybox(mk(0.f), mk(1.f), 0.f, 0.f);
    int node = 0, sb = 0, sp = 0; // pending subtrees: stack[sb, sp), oldest (largest) at sb
    int stack[kBvhStack];
    uint32_t n_nodes = 0, n_cands = 0, n_acc = 0, n_err = 0, n_trips = 0, n_cb = 0, n_pb = 0;

    for (;;) {
        if (!exhausted) { // lanes without work take the next query pieces
            const bool need = cur < 0;
            const uint64_t m = __ballot(need);
            if (m) {
                const int leader = __ffsll((unsigned long long)m) - 1;
                uint32_t first = 0, lim = 0;
                bool got = false;
                while (!got && !exhausted) {
                    uint32_t base = 0;
                    uint32_t* h = split ? &xw[208 + src] : head;
                    const uint32_t rb0 = split ? xw[192 + src] : 0u, re0 = split ? xw[200 + src] : total;
                    if (int(lane) == leader) base = atomicAdd(h, uint32_t(__popcll(m)));
                    base = __shfl(base, leader);
                    if (rb0 + base < re0) {
                        got = true;
                        first = rb0 + base;
                        lim = re0;
                    } else if (split && ++tried < uint32_t(kXcds)) {
                        src = (src + 1) % kXcds;
                    } else {
                        exhausted = true;
                    }
                }
                if (got && need) {
                    const uint32_t wi = first + uint32_t(__popcll(m & lanes_below()));
                    const uint2 pc = wi < lim ? (KIND == 1 ? plist[wi] : make_uint2(wi, 1u << 16)) : make_uint2(~0u, 0u);
                    // a piece left unwritten by an overflowing camera wave is skipped
                    if (pc.x < n_recs && (pc.y & 0xffffu) < (pc.y >> 16) && (pc.y >> 16) <= kPiecesPerRay) {
                        const SegRec* r = b.segq + pc.x;
                        const float4 a = r->a, bb = r->b;
                        float lo, hi, ta, tb;
                        piece_range(a.w, bb.w, pc.y & 0xffffu, pc.y >> 16, lo, hi, ta, tb);
                        rb = make_raybox(mk(a.x, a.y, a.z), mk(bb.x, bb.y, bb.z), ta, tb);
                        kn = pack_kn(pc.y & 0xffffu, pc.y >> 16);
                        node = 0;
                        sb = sp = 0;
                        cur = int(pc.x);
                    }
                }
            }
        } else { // no pieces left: idle lanes take the oldest pending subtree of a busy lane
            const bool idle = cur < 0, donor = cur >= 0 && sp > sb;
            const uint64_t mi = __ballot(idle), md = __ballot(donor);
            const uint32_t t = min(uint32_t(__popcll(mi)), uint32_t(__popcll(md)));
            if (t) {
                if (donor) {
                    const uint32_t k = uint32_t(__popcll(md & lanes_below()));
                    if (k < t) {
                        S.q[k] = uint32_t(cur);
                        S.node[k] = uint32_t(stack[sb++]);
                        S.kn[k] = uint16_t(kn);
                    }
                }
                wave_lds_sync();
                if (idle) {
                    const uint32_t k = uint32_t(__popcll(mi & lanes_below()));
                    if (k < t) {
                        cur = int(S.q[k]);
                        node = int(S.node[k]);
                        kn = S.kn[k];
                        const SegRec* r = b.segq + cur;
                        const float4 a = r->a, bb = r->b;
                        float lo, hi, ta, tb;
                        piece_range(a.w, bb.w, kn & 63u, kn >> 6, lo, hi, ta, tb);
                        rb = make_raybox(mk(a.x, a.y, a.z), mk(bb.x, bb.y, bb.z), ta, tb);
                        sb = sp = 0;
                    }
                }
                wave_lds_sync();
            }
        }
        const bool active = cur >= 0;
        const bool any_active = __ballot(active) != 0;
        n_trips++;
        // one wide node of this lane's traversal; leaves hit become candidates
        int lf[kWide];
#pragma unroll
        for (int k = 0; k < kWide; k++) lf[k] = -1;
        uint32_t qnow = 0;
        if (active) {
            const WNode nd = bvh.nodes[node];
            n_nodes++;
            const float sx = wnode_step(nd.meta, 0), sy = wnode_step(nd.meta, 1), sz = wnode_step(nd.meta, 2);
            const int nc = wnode_count(nd.meta);
            int next = -1;
#pragma unroll
            for (int k = 0; k < kWide; k++) {
                if (k < nc) {
                    float e, lx, ly, lz, ux, uy, uz;
                    wnode_box(nd, k, sx, sy, sz, lx, ly, lz, ux, uy, uz);
                    const bool h = slab(rb, lx, ly, lz, ux, uy, uz, rb.tmax, e);
                    const int ch = nd.child[k];
                    if (h) {
                        if (ch < 0) {
                            lf[k] = ~ch;
                        } else if (next < 0) {
                            next = ch;
                        } else {
                            if (sp == kBvhStack && sb > 0) { // slide the pending entries down
                                for (int t = sb; t < sp; t++) stack[t - sb] = stack[t];
                                sp -= sb;
                                sb = 0;
                            }
                            if (sp < kBvhStack) {
                                stack[sp++] = ch;
                            } else {
                                n_err++; // traversal stack exhausted: the piece is abandoned and counted
                                sb = sp = 0;
                                for (int t = 0; t < kWide; t++) lf[t] = -1;
                                next = -1;
                                break;
                            }
                        }
                    }
                }
            }
            if (next < 0 && sp > sb) next = stack[--sp];
            node = next;
            qnow = uint32_t(cur);
            if (next < 0) cur = -1;
        }
        // all lanes take part in the queue appends (ballots over the whole wave)
#pragma unroll
        for (int k = 0; k < kWide; k++) cq_push(C, cn, lf[k] >= 0, qnow, uint32_t(lf[k]), kn);
        // Evaluate 64 accepted pairs at a time whenever 64 are queued (this keeps fewer than 64
        // pending before each candidate test), else test 64 candidates at a time. Once no lane
        // has work left, drain both queues: candidates first, then pairs.
        for (;;) {
            const bool eval_p = qn >= 64u || (!any_active && cn == 0u && qn > 0u);
            const bool test_c = !eval_p && (cn >= 64u || (!any_active && cn > 0u));
            if (!eval_p && !test_c) break;
            if (eval_p) {
                n_pb++;
                eval_pairs<KIND, LOG>(sc, p, b, Q, K, qn, norm);
            } else {
                const uint32_t n = cn < 64u ? cn : 64u;
                n_cb++;
                wave_lds_sync();
                bool acc = false;
                uint32_t qi = 0, slot = 0;
                float x = 0.f, y = 0.f, z = 0.f, w = 0.f;
                if (lane < n) {
                    const uint32_t e = cn - n + lane;
                    qi = C.q[e];
                    const uint32_t leaf = C.leaf[e], ckn = C.kn[e];
                    const SegRec* r = b.segq + qi;
                    const float4 a = r->a, bb = r->b, d = r->d;
                    float clo, chi, cta, ctb;
                    piece_range(a.w, bb.w, ckn & 63u, ckn >> 6, clo, chi, cta, ctb);
                    acc = leaf_accept<KIND>(p, *b.maps, b.beams, leaf, mk(a.x, a.y, a.z), mk(bb.x, bb.y, bb.z), a.w,
                                            bb.w, clo, chi, __float_as_int(d.z), slot, x, y, z, w);
                }
                wave_lds_sync();
                cn -= n;
                n_cands += lane < n ? 1u : 0u;
                n_acc += acc ? 1u : 0u;
                pq_push(Q, qn, acc, qi, slot, x, y, z, w);
            }
        }
        if (!any_active && exhausted) break;
    }
    // write back the queries still held in the colour cache
    wave_lds_sync();
    if (K.tag[lane] != 0u) add_color(b.mcolor, K.pix[lane], K.r[lane], K.g[lane], K.b[lane]);
    const uint32_t s_nodes = wave_sum(n_nodes), s_cands = wave_sum(n_cands), s_acc = wave_sum(n_acc),
                   s_err = wave_sum(n_err);
    // loop trips and batches are the same in every lane: lane 0 reports them for the wave
    if (lane == 0) {
        stat_add(b.stats, STAT_SEG_NODES, s_nodes);
        stat_add(b.stats, STAT_SEG_CANDS, s_cands);
        stat_add(b.stats, STAT_ACCEPT, s_acc);
        stat_add(b.stats, KIND == 0 ? STAT_ACC_PB2D : STAT_ACC_BB1D, s_acc);
        stat_add(b.stats, STAT_ERRORS, s_err);
        stat_add(b.stats, STAT_MERGE_STACK, s_err);
        stat_add(b.stats, STAT_SEG_TRIPS, n_trips);
        stat_add(b.stats, STAT_CAND_BATCHES, n_cb);
        stat_add(b.stats, STAT_PAIR_BATCHES, n_pb);
    }
}

template <int KIND, bool LOG>
__global__ void __launch_bounds__(kMergeBlock) k_seg_merge(DScene sc, Bufs b, int bounce, int leftovers)
{
    __shared__ CandQueue cq[kMergeWaves];
    __shared__ PairQueue pq[kMergeWaves];
    __shared__ Steal stl[kMergeWaves];
    __shared__ ColorCache ccache[kMergeWaves];
    const uint32_t w = threadIdx.x >> 6;
    seg_merge_body<KIND, LOG>(sc, b, bounce, blockIdx.x, false, leftovers != 0, cq[w], pq[w], stl[w], ccache[w]);
}

// The segment merges of one bounce for every job of a batch in one persistent launch: each block
// walks the jobs in turn (starting at its own, so that the blocks spread over the jobs), and its
// waves take pieces of a job until that job has none left, so that all resident waves help with
// whatever work is left, whatever the jobs' sizes.
template <int KIND>
__global__ void __launch_bounds__(kMergeBlock) k_seg_merge_batch(const BatchJob* jobs, uint32_t njobs, int bounce)
{
    __shared__ CandQueue cq[kMergeWaves];
    __shared__ PairQueue pq[kMergeWaves];
    __shared__ Steal stl[kMergeWaves];
    __shared__ ColorCache ccache[kMergeWaves];
    const uint32_t w = threadIdx.x >> 6;
    for (uint32_t t = 0; t < njobs; t++) {
        const BatchJob& J = jobs[(blockIdx.x + t) % njobs];
        const IterParams& p = *J.b.params;
        const bool on = KIND == 0 ? (p.build_pb2d && J.b.maps->wmed.n > 0) : (p.build_bb1d && J.b.maps->wbeam.n > 0);
        if (!on || bounce >= int(p.max_len)) continue;
        // B-B1D on a bundled bounce: the pieces the bundle kernel left
        const bool left = KIND == 1 && bounce < p.bundle_bounces && J.b.units != nullptr &&
                          J.b.maps->wbeam.n < kBundleMaxPrims;
        seg_merge_body<KIND, false>(J.sc, J.b, bounce, blockIdx.x, true, left, cq[w], pq[w], stl[w], ccache[w]);
    }
}

// ---- B-B1D bundles -------------------------------------------------------------------------
// A bundle is one camera wave's run of j-th pieces: up to 64 neighbouring rays' stretches at
// similar depth (on the first bounce, the primary rays of an 8x8-pixel tile). One wave walks the
// beam tree once for the whole bundle. Every piece lies in a capsule around one axis segment (the
// pieces' mean direction through their mean start, over their extent along it; radius: the
// farthest piece end from that line), and a child is visited when its box, grown by the radius,
// meets the axis segment. The wave takes up to 64 frontier nodes per step, one per lane, so one
// fetch of a node serves every ray of the bundle. The leaves the capsule reaches are staged in
// LDS kStage at a time (the quantised leaf box and the sub-beam record), and every ray tests each
// of them as the per-piece walk does at a leaf: its own slab test of the leaf box, then the exact
// predicate with both ownership rules. The capsule meets every box a piece meets, so the bundle
// lists every leaf the per-piece walk reaches, and the pairs accepted are the same (a leaf box
// can overhang its parent's by under one quantisation step; a pair the exact predicate places in
// that sliver is the only possible difference). A bundle wider than p.bundle_radius, or whose
// capsule (grown by a sub-beam box) is larger than its pieces' own grown segments together (rays
// that diverge, or stretches at different depths), hands its pieces to the per-piece walk
// (pieces_x, counted in qcount[448 + bounce]), which runs next. Three 4-wave blocks fit a CU
// (LDS and VGPRs), the per-piece walk's occupancy.
constexpr uint32_t kFrontCap = 256;                      // frontier size up to which 64 nodes go per step
constexpr uint32_t kFrontAlloc = kFrontCap + 7 * 64 + 8; // + growth of one-node steps down the deepest tree
constexpr uint32_t kStage = 32;                          // sub-beams staged in LDS at a time
constexpr uint32_t kBundleLeafCap = kStage + 64 * kWide; // < kStage pending + up to kWide leaves per lane

struct BundleLds {
    uint32_t front[kFrontAlloc];   // frontier nodes (taken from the top)
    uint32_t leaf[kBundleLeafCap]; // leaves reached, not yet tested: node << 3 | child
    float4 g0[kStage], g1[kStage], g2[kStage]; // the sub-beam records being tested
    float lx[kStage], ly[kStage], lz[kStage], ux[kStage], uy[kStage], uz[kStage]; // and their quantised boxes
    uint16_t cand[128];            // (piece lane, staged sub-beam) pairs whose box test passed
};

__device__ __forceinline__ float wave_fsum(float v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ float wave_fmax(float v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ float wave_fmin(float v)
{
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}

// One job's B-B1D bundles of one bounce, by the wave the LDS belongs to.
template <bool LOG>
__device__ __forceinline__ void seg_bundle_body(const DScene& sc, const Bufs& b, int bounce, BundleLds& U,
                                                PairQueue& Q, ColorCache& K)
{
    const IterParams& p = *b.params;
    const uint32_t lane = __lane_id();
    K.tag[lane] = 0u;
    K.r[lane] = K.g[lane] = K.b[lane] = 0.f;
    wave_lds_sync();
    const uint32_t n_recs = min(b.qcount[128 + bounce], b.seg_cap);
    const uint32_t n_units = min(b.qcount[384 + bounce], b.piece_cap);
    uint32_t* uhead = b.xwork + size_t(bounce) * kXWork + 216;
    uint32_t* left = &b.qcount[448 + bounce];
    const WNode* nodes = b.maps->wbeam.nodes;
    const float4* geo = b.maps->beam_geo;
    const float norm = p.norm_bb1d;
    uint32_t qn = 0; // pairs queued (same value in every lane)
    uint32_t n_bundles = 0, n_left = 0, n_nodes = 0, n_leaves = 0, n_tests = 0, n_acc = 0, n_err = 0;

    for (;;) {
        uint32_t u = 0;
        if (lane == 0) u = atomicAdd(uhead, 1u);
        u = __shfl(u, 0);
        if (u >= n_units) break;
        const uint2 un = b.units[u];
        const uint32_t cnt = min(un.y, 64u);
        // this lane's piece (the per-piece walk's validity rule: a piece left unwritten by an
        // overflowing camera wave is skipped)
        bool has = false;
        uint2 pc = make_uint2(~0u, 0u);
        f3 ro = mk(0.f), rd = mk(0.f);
        float dmin = 0.f, dmax = 0.f, lo = 0.f, hi = 0.f, ta = 0.f, tb = 0.f;
        int med = 0;
        if (lane < cnt && un.x + lane < b.piece_cap) {
            pc = b.pieces[un.x + lane];
            if (pc.x < n_recs && (pc.y & 0xffffu) < (pc.y >> 16) && (pc.y >> 16) <= kPiecesPerRay) {
                const SegRec* r = b.segq + pc.x;
                const float4 a = r->a, bb = r->b, d = r->d;
                ro = mk(a.x, a.y, a.z);
                rd = mk(bb.x, bb.y, bb.z);
                dmin = a.w;
                dmax = bb.w;
                med = __float_as_int(d.z);
                piece_range(dmin, dmax, pc.y & 0xffffu, pc.y >> 16, lo, hi, ta, tb);
                has = true;
            }
        }
        const uint64_t hm = __ballot(has);
        if (!hm) continue;
        const uint32_t nh = uint32_t(__popcll(hm));
        // the capsule
        const f3 e0 = has ? ro + rd * ta : mk(0.f), e1 = has ? ro + rd * tb : mk(0.f);
        const float inv_n = 1.f / float(nh);
        const f3 c = mk(wave_fsum(e0.x), wave_fsum(e0.y), wave_fsum(e0.z)) * inv_n;
        f3 ax = mk(wave_fsum(rd.x), wave_fsum(rd.y), wave_fsum(rd.z));
        const float al = sqrtf(dot(ax, ax));
        bool fit = al > 1e-3f * float(nh);
        ax = fit ? ax / al : mk(1.f, 0.f, 0.f);
        float smin = kInf, smax = -kInf, r2 = 0.f, mag = 0.f;
        if (has) {
            const f3 v0 = e0 - c, v1 = e1 - c;
            const float s0 = dot(v0, ax), s1 = dot(v1, ax);
            const f3 q0 = v0 - ax * s0, q1 = v1 - ax * s1;
            r2 = fmaxf(dot(q0, q0), dot(q1, q1));
            smin = fminf(s0, s1);
            smax = fmaxf(s0, s1);
            mag = fmaxf(fmaxf(fmaxf(fabsf(e0.x), fabsf(e0.y)), fabsf(e0.z)),
                        fmaxf(fmaxf(fabsf(e1.x), fabsf(e1.y)), fabsf(e1.z)));
        }
        smin = wave_fmin(smin);
        smax = wave_fmax(smax);
        const float rad = sqrtf(wave_fmax(r2));
        mag = wave_fmax(mag);
        // Cost test: the capsule, grown by a sub-beam box (half size hb), holds no more volume
        // than the pieces' own grown segments together. Staggered or diverging pieces (a long or
        // wide capsule) would make every piece test sub-beams far from it.
        const float hb = p.bundle_box;
        const float plen = wave_fsum(has ? tb - ta : 0.f);
        const float cap_vol = (rad + hb) * (rad + hb) * (smax - smin + 2.f * hb);
        fit = fit && rad <= p.bundle_radius && cap_vol <= hb * hb * (plen + 2.f * hb * float(nh));
        // the capsule test only culls: FP32 slack on the radius keeps it conservative. A piece
        // reaching far outside the scene (a huge coordinate) would make the slack, not the pieces,
        // set the capsule: such a bundle goes to the per-piece walk too.
        const float rg = rad * (1.f + 1e-4f) + 1e-5f * mag + 1e-6f;
        fit = fit && rg <= rad + hb;
        if (!fit) { // to the per-piece walk
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(left, nh);
            base = __shfl(base, 0);
            const uint32_t at = base + uint32_t(__popcll(hm & lanes_below()));
            if (has && at < b.piece_cap) b.pieces_x[at] = pc;
            n_left += lane == 0 ? nh : 0u;
            continue;
        }
        n_bundles += lane == 0 ? 1u : 0u;
        const f3 a0 = c + ax * smin, a1 = c + ax * smax;
        const RayBox cap = make_raybox(a0, a1 - a0, 0.f, 1.f);
        const RayBox prb = make_raybox(ro, rd, ta, tb); // this piece, as the per-piece walk tests it
        uint32_t fs = 1, ln = 0; // frontier / listed leaves (same value in every lane)
        if (lane == 0) U.front[0] = 0u;
        wave_lds_sync();
        while (fs > 0 || ln > 0) {
            if (fs > 0) {
                const uint32_t room = fs < kFrontCap ? (kFrontCap - fs) / 7u : 0u;
                const uint32_t take = min(min(64u, fs), max(1u, room));
                if (fs + 7u * take > kFrontAlloc) { // unreachable for trees of at most 64 levels
                    n_err += lane == 0 ? nh : 0u;
                    fs = ln = 0;
                    break;
                }
                int node = -1;
                if (lane < take) node = int(U.front[fs - 1u - lane]);
                wave_lds_sync();
                fs -= take;
                int lf[kWide], in[kWide];
#pragma unroll
                for (int k = 0; k < kWide; k++) lf[k] = in[k] = -1;
                if (node >= 0) {
                    const WNode nd = nodes[node];
                    n_nodes++;
                    const float sx = wnode_step(nd.meta, 0), sy = wnode_step(nd.meta, 1), sz = wnode_step(nd.meta, 2);
                    const int nc = wnode_count(nd.meta);
#pragma unroll
                    for (int k = 0; k < kWide; k++) {
                        if (k < nc) {
                            float e, lx, ly, lz, ux, uy, uz;
                            wnode_box(nd, k, sx, sy, sz, lx, ly, lz, ux, uy, uz);
                            if (slab(cap, lx - rg, ly - rg, lz - rg, ux + rg, uy + rg, uz + rg, 1.f, e)) {
                                const int ch = nd.child[k];
                                if (ch < 0) lf[k] = (node << 3) | k;
                                else in[k] = ch;
                            }
                        }
                    }
                }
#pragma unroll
                for (int k = 0; k < kWide; k++) {
                    const uint64_t mi = __ballot(in[k] >= 0), ml = __ballot(lf[k] >= 0);
                    if (in[k] >= 0) U.front[fs + uint32_t(__popcll(mi & lanes_below()))] = uint32_t(in[k]);
                    if (lf[k] >= 0) U.leaf[ln + uint32_t(__popcll(ml & lanes_below()))] = uint32_t(lf[k]);
                    fs += uint32_t(__popcll(mi));
                    ln += uint32_t(__popcll(ml));
                }
                wave_lds_sync();
            }
            // test kStage listed leaves at a time, and the rest once the frontier is empty
            while (ln >= kStage || (fs == 0u && ln > 0u)) {
                const uint32_t n = min(ln, kStage);
                if (lane < n) { // the leaf's box from its node (fetched moments ago) and its record
                    n_leaves++;
                    const uint32_t e = U.leaf[ln - n + lane], k = e & 7u;
                    const WNode& nd = nodes[e >> 3];
                    const uint32_t meta = nd.meta;
                    float lx, ly, lz, ux, uy, uz;
                    wnode_box(nd, int(k), wnode_step(meta, 0), wnode_step(meta, 1), wnode_step(meta, 2), lx, ly, lz, ux,
                              uy, uz);
                    const uint32_t kl = uint32_t(~nd.child[k]);
                    U.lx[lane] = lx;
                    U.ly[lane] = ly;
                    U.lz[lane] = lz;
                    U.ux[lane] = ux;
                    U.uy[lane] = uy;
                    U.uz[lane] = uz;
                    U.g0[lane] = geo[3 * kl];
                    U.g1[lane] = geo[3 * kl + 1];
                    U.g2[lane] = geo[3 * kl + 2];
                }
                wave_lds_sync();
                ln -= n;
                // Each piece's box test of each staged sub-beam; the pairs that pass are queued
                // and take the exact test 64 at a time, one pair per lane (the piece's values
                // come from its lane by shuffles), so that the exact tests run on full waves.
                uint32_t cn = 0;
                for (uint32_t j = 0; j <= n; j++) {
                    if (j < n) {
                        float e;
                        const bool h =
                            has && slab(prb, U.lx[j], U.ly[j], U.lz[j], U.ux[j], U.uy[j], U.uz[j], prb.tmax, e);
                        const uint64_t mh = __ballot(h);
                        if (h) U.cand[cn + uint32_t(__popcll(mh & lanes_below()))] = uint16_t((j << 6) | lane);
                        cn += uint32_t(__popcll(mh));
                        n_tests += h ? 1u : 0u;
                        if (cn < 64u) continue;
                    }
                    if (cn == 0u) continue;
                    const uint32_t m = min(cn, 64u);
                    wave_lds_sync();
                    const uint32_t ce = lane < m ? U.cand[cn - m + lane] : lane;
                    const uint32_t src = ce & 63u, sj = (ce >> 6) & 63u;
                    const f3 pro = mk(__shfl(ro.x, src), __shfl(ro.y, src), __shfl(ro.z, src));
                    const f3 prd = mk(__shfl(rd.x, src), __shfl(rd.y, src), __shfl(rd.z, src));
                    const float pdmin = __shfl(dmin, src), pdmax = __shfl(dmax, src);
                    const float plo = __shfl(lo, src), phi = __shfl(hi, src);
                    const int pmed = __shfl(med, src);
                    const uint32_t pq_q = uint32_t(__shfl(int(pc.x), src));
                    bool acc = false;
                    uint32_t slot = 0;
                    float x = 0.f, y = 0.f, z = 0.f, w = 0.f;
                    if (lane < m)
                        acc = bb1d_accept(p, U.g0[sj], U.g1[sj], U.g2[sj], pro, prd, pdmin, pdmax, plo, phi, pmed, slot,
                                          x, y, z, w);
                    wave_lds_sync();
                    cn -= m;
                    n_acc += acc ? 1u : 0u;
                    pq_push(Q, qn, acc, pq_q, slot, x, y, z, w);
                    if (qn >= 64u) eval_pairs<1, LOG>(sc, p, b, Q, K, qn, norm);
                    if (cn > 0u && j == n) j--; // drain what is left after the last sub-beam
                }
                wave_lds_sync();
            }
        }
    }
    while (qn > 0u) eval_pairs<1, LOG>(sc, p, b, Q, K, qn, norm);
    // write back the queries still held in the colour cache
    wave_lds_sync();
    if (K.tag[lane] != 0u) add_color(b.mcolor, K.pix[lane], K.r[lane], K.g[lane], K.b[lane]);
    const uint32_t s_nodes = wave_sum(n_nodes), s_leaves = wave_sum(n_leaves), s_tests = wave_sum(n_tests),
                   s_acc = wave_sum(n_acc);
    if (lane == 0) {
        stat_add(b.stats, STAT_BUNDLES, n_bundles);
        stat_add(b.stats, STAT_BUNDLE_LEFT, n_left);
        stat_add(b.stats, STAT_BUNDLE_NODES, s_nodes);
        stat_add(b.stats, STAT_BUNDLE_LEAVES, s_leaves);
        stat_add(b.stats, STAT_BUNDLE_TESTS, s_tests);
        stat_add(b.stats, STAT_ACCEPT, s_acc);
        stat_add(b.stats, STAT_ACC_BB1D, s_acc);
        stat_add(b.stats, STAT_ERRORS, n_err);
        stat_add(b.stats, STAT_MERGE_STACK, n_err);
    }
}

template <bool LOG>
__global__ void __launch_bounds__(kMergeBlock, 3) k_seg_bundle(DScene sc, Bufs b, int bounce)
{
    __shared__ BundleLds blds[kMergeWaves];
    __shared__ PairQueue pq[kMergeWaves];
    __shared__ ColorCache ccache[kMergeWaves];
    const uint32_t w = threadIdx.x >> 6;
    seg_bundle_body<LOG>(sc, b, bounce, blds[w], pq[w], ccache[w]);
}

// The bundles of one bounce for every job of a batch that walks bundles on it, in one persistent
// launch (the jobs walked in turn as in k_seg_merge_batch).
__global__ void __launch_bounds__(kMergeBlock, 3) k_seg_bundle_batch(const BatchJob* jobs, uint32_t njobs, int bounce)
{
    __shared__ BundleLds blds[kMergeWaves];
    __shared__ PairQueue pq[kMergeWaves];
    __shared__ ColorCache ccache[kMergeWaves];
    const uint32_t w = threadIdx.x >> 6;
    for (uint32_t t = 0; t < njobs; t++) {
        const BatchJob& J = jobs[(blockIdx.x + t) % njobs];
        const IterParams& p = *J.b.params;
        if (!(p.build_bb1d && J.b.maps->wbeam.n > 0) || bounce >= int(p.max_len) || bounce >= p.bundle_bounces ||
            !J.b.units || J.b.maps->wbeam.n >= kBundleMaxPrims)
            continue;
        seg_bundle_body<false>(J.sc, J.b, bounce, blds[w], pq[w], ccache[w]);
    }
}

// ---- B-B1D packet traversal -------------------------------------------------------------
// One wave walks the wide beam tree for 64 query pieces at once: the node index is the same in
// every lane, so each node is fetched once per wave (a scalar load) instead of once per lane, and
// each lane tests its own piece against the node's child boxes. A child that any lane hits is
// visited by the whole wave; at a leaf, the lanes whose box test passed run the exact predicate.
// The traversal stack is per wave in LDS. This pays when the wave's pieces are close together --
// the same stretch of 64 neighbouring camera rays, which is how the camera pass writes them -- and
// is used for the bounces up to IterParams::packet_bounces.
constexpr int kPacketStack = 192;

template <bool LOG>
__global__ void __launch_bounds__(kMergeBlock) k_seg_merge_packet(DScene sc, Bufs b, int bounce)
{
    const IterParams& p = *b.params;
    __shared__ int stk_all[kMergeWaves][kPacketStack];
    __shared__ PairQueue pq[kMergeWaves];
    __shared__ ColorCache ccache[kMergeWaves];
    int* stk = stk_all[threadIdx.x >> 6];
    PairQueue& Q = pq[threadIdx.x >> 6];
    ColorCache& K = ccache[threadIdx.x >> 6];
    const uint32_t lane = __lane_id();
    K.tag[lane] = 0u;
    K.r[lane] = K.g[lane] = K.b[lane] = 0.f;
    wave_lds_sync();
    const uint32_t n_recs = min(b.qcount[128 + bounce], b.seg_cap);
    const uint32_t total = min(b.qcount[320 + bounce], b.piece_cap);
    uint32_t* head = &b.qcount[192 + 2 * bounce + 1];
    const WNode* nodes = b.maps->wbeam.nodes;
    uint32_t qn = 0; // pairs queued (same value in every lane)
    uint32_t n_nodes = 0, n_cands = 0, n_acc = 0, n_err = 0, n_pb = 0;
    for (;;) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(head, 64u);
        base = __shfl(base, 0);
        if (base >= total) break;
        // this lane's piece: its ray, segment, owned camera range and box-test range
        const uint32_t wi = base + lane;
        const uint2 pc = wi < total ? (p.piece_order == 2 ? b.pieces_x : b.pieces)[wi] : make_uint2(~0u, 0u);
        const bool act = pc.x < n_recs && (pc.y & 0xffffu) < (pc.y >> 16) && (pc.y >> 16) <= kPiecesPerRay;
        f3 ro = mk(0.f), rd = mk(1.f);
        float dmin = 0.f, dmax = 0.f, lo = 0.f, hi = 0.f;
        int med = -1;
        RayBox rb = make_raybox(ro, rd, 1.f, 0.f);
        if (act) {
            const SegRec* r = b.segq + pc.x;
            const float4 a = r->a, bb = r->b, d = r->d;
            ro = mk(a.x, a.y, a.z);
            rd = mk(bb.x, bb.y, bb.z);
            dmin = a.w;
            dmax = bb.w;
            med = __float_as_int(d.z);
            float ta, tb;
            piece_range(dmin, dmax, pc.y & 0xffffu, pc.y >> 16, lo, hi, ta, tb);
            rb = make_raybox(ro, rd, ta, tb);
        }
        if (!__ballot(act)) continue;
        int sp = 0; // stack depth (same in every lane)
        int node = 0;
        bool overflow = false;
        for (;;) {
            const int nu = __builtin_amdgcn_readfirstlane(node);
            const WNode nd = nodes[nu];
            n_nodes++;
            const float sx = wnode_step(nd.meta, 0), sy = wnode_step(nd.meta, 1), sz = wnode_step(nd.meta, 2);
            const int nc = wnode_count(nd.meta);
            uint32_t hits = 0;
            if (act) {
#pragma unroll
                for (int k = 0; k < kWide; k++) {
                    float e, lx, ly, lz, ux, uy, uz;
                    if (k >= nc) continue;
                    wnode_box(nd, k, sx, sy, sz, lx, ly, lz, ux, uy, uz);
                    if (slab(rb, lx, ly, lz, ux, uy, uz, rb.tmax, e)) hits |= 1u << k;
                }
            }
            int next = -1;
            for (int k = 0; k < nc; k++) {
                const bool h = (hits >> k) & 1u;
                if (!__ballot(h)) continue;
                const int ch = nd.child[k];
                if (ch < 0) { // leaf: the exact predicate in the lanes that hit its box
                    uint32_t slot = 0;
                    float x = 0.f, y = 0.f, z = 0.f, w = 0.f;
                    bool acc = false;
                    if (h) {
                        n_cands++;
                        acc = leaf_accept<1>(p, *b.maps, b.beams, uint32_t(~ch), ro, rd, dmin, dmax, lo, hi, med, slot,
                                             x, y, z, w);
                    }
                    n_acc += acc ? 1u : 0u;
                    pq_push(Q, qn, acc, pc.x, slot, x, y, z, w);
                    if (qn >= 64u) {
                        n_pb++;
                        eval_pairs<1, LOG>(sc, p, b, Q, K, qn, p.norm_bb1d);
                    }
                } else if (next < 0) {
                    next = ch;
                } else if (sp < kPacketStack) {
                    if (lane == 0) stk[sp] = ch;
                    sp++;
                } else {
                    overflow = true;
                }
            }
            if (overflow) break;
            if (next < 0) {
                if (sp == 0) break;
                wave_lds_sync();
                next = stk[--sp];
                wave_lds_sync();
            }
            node = next;
        }
        // a full stack abandons the wave's pieces; they are counted (zero on every workload here)
        if (overflow) n_err += act ? 1u : 0u;
    }
    while (qn > 0u) {
        n_pb++;
        eval_pairs<1, LOG>(sc, p, b, Q, K, qn, p.norm_bb1d);
    }
    wave_lds_sync();
    if (K.tag[lane] != 0u) add_color(b.mcolor, K.pix[lane], K.r[lane], K.g[lane], K.b[lane]);
    const uint32_t s_cands = wave_sum(n_cands), s_acc = wave_sum(n_acc), s_err = wave_sum(n_err);
    if (lane == 0) {
        stat_add(b.stats, STAT_SEG_NODES, n_nodes);
        stat_add(b.stats, STAT_SEG_CANDS, s_cands);
        stat_add(b.stats, STAT_ACCEPT, s_acc);
        stat_add(b.stats, STAT_ACC_BB1D, s_acc);
        stat_add(b.stats, STAT_ERRORS, s_err);
        stat_add(b.stats, STAT_MERGE_STACK, s_err);
        stat_add(b.stats, STAT_SEG_TRIPS, n_nodes);
        stat_add(b.stats, STAT_PAIR_BATCHES, n_pb);
    }
}

// ---- XCD-aware split of the B-B1D query pieces ------------------------------------
// Each piece is put in one of 4x4x4 cells of the beam map's bounds by its midpoint; the
// cells are numbered octant-major, so the eight cells of octant x (its pieces stored
// contiguously) form the work range of XCD x. Workgroups are dispatched to the XCDs
// round-robin, so block i works mostly on octant i % 8 and that XCD's L2 holds the
// subtrees of one region instead of the whole tree.

__device__ __forceinline__ uint32_t piece_cell(const Bufs& b, const WNode& root, uint2 pc)
{
    const SegRec* r = b.segq + pc.x;
    const float4 a = r->a, bb = r->b;
    float lo, hi, ta, tb;
    piece_range(a.w, bb.w, pc.y & 0xffffu, pc.y >> 16, lo, hi, ta, tb);
    const float t = 0.5f * (ta + tb);
    const float m[3] = {a.x + bb.x * t, a.y + bb.y * t, a.z + bb.z * t};
    const float o[3] = {root.ox, root.oy, root.oz};
    int c[3];
    for (int k = 0; k < 3; k++) {
        const float ext = 255.f * wnode_step(root.meta, k);
        const float u = ext > 0.f ? (m[k] - o[k]) / ext : 0.f;
        const int v = int(u * 4.f);
        c[k] = v < 0 ? 0 : (v > 3 ? 3 : v);
    }
    const uint32_t octant = uint32_t((c[0] >> 1) | ((c[1] >> 1) << 1) | ((c[2] >> 1) << 2));
    const uint32_t sub = uint32_t((c[0] & 1) | ((c[1] & 1) << 1) | ((c[2] & 1) << 2));
    return octant * 8u + sub;
}

__global__ void __launch_bounds__(256) k_piece_cells(Bufs b, int bounce)
{
    __shared__ uint32_t hist[kCells];
    if (threadIdx.x < kCells) hist[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t n = min(b.qcount[320 + bounce], b.piece_cap);
    uint32_t* xw = b.xwork + size_t(bounce) * kXWork;
    const WNode root = b.maps->wbeam.nodes[0];
    const uint32_t n_recs = min(b.qcount[128 + bounce], b.seg_cap);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint2 pc = b.pieces[i];
        const uint32_t c = pc.x < n_recs ? piece_cell(b, root, pc) : 0u;
        b.piece_cell[i] = uint8_t(c);
        atomicAdd(&hist[c], 1u);
    }
    __syncthreads();
    if (threadIdx.x < kCells && hist[threadIdx.x]) atomicAdd(&xw[threadIdx.x], hist[threadIdx.x]);
}

__global__ void k_piece_offsets(Bufs b, int bounce)
{
    uint32_t* xw = b.xwork + size_t(bounce) * kXWork;
    if (threadIdx.x != 0) return;
    uint32_t run = 0;
    for (int c = 0; c < kCells; c++) {
        if (c % 8 == 0) xw[192 + c / 8] = run;
        xw[64 + c] = run;
        xw[128 + c] = run;
        run += xw[c];
        if (c % 8 == 7) xw[200 + c / 8] = run;
    }
    for (int x = 0; x < kXcds; x++) xw[208 + x] = 0;
}

__global__ void __launch_bounds__(256) k_piece_scatter(Bufs b, int bounce)
{
    __shared__ uint32_t cnt[kCells], base[kCells];
    const uint32_t n = min(b.qcount[320 + bounce], b.piece_cap);
    uint32_t* xw = b.xwork + size_t(bounce) * kXWork;
    const uint32_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const uint32_t c0 = blockIdx.x * chunk, c1 = min(n, c0 + chunk);
    if (threadIdx.x < kCells) cnt[threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t i = c0 + threadIdx.x; i < c1; i += blockDim.x) atomicAdd(&cnt[b.piece_cell[i]], 1u);
    __syncthreads();
    if (threadIdx.x < kCells) {
        base[threadIdx.x] = cnt[threadIdx.x] ? atomicAdd(&xw[128 + threadIdx.x], cnt[threadIdx.x]) : 0u;
        cnt[threadIdx.x] = 0;
    }
    __syncthreads();
    for (uint32_t i = c0 + threadIdx.x; i < c1; i += blockDim.x) {
        const uint32_t c = b.piece_cell[i];
        b.pieces_x[base[c] + atomicAdd(&cnt[c], 1u)] = b.pieces[i];
    }
}

} // namespace

// blocks of the persistent merge grid: CUs x resident blocks per CU (at most the occupancy
// limit; bpc > 0 asks for fewer)
int merge_grid_blocks(int bpc)
{
    static int cus = 0, per = 0;
    if (cus == 0) {
        int dev = 0, per0 = 0, per1 = 0, per2 = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per0, k_seg_merge<0, false>, kMergeBlock, 0);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per1, k_seg_merge<1, false>, kMergeBlock, 0);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per2, k_seg_merge_packet<false>, kMergeBlock, 0);
        per = std::max(1, std::min(std::min(per0, per1), per2));
        cus = std::max(1, cus);
    }
    return cus * (bpc > 0 ? std::min(bpc, per) : per);
}

// blocks of the persistent bundle grid: CUs x its occupancy limit
int bundle_grid_blocks()
{
    static int blocks = 0;
    if (blocks == 0) {
        int dev = 0, cus = 0, per = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        int per1 = 0, per2 = 0;
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_seg_bundle<false>, kMergeBlock, 0);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per1, k_seg_bundle<true>, kMergeBlock, 0);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per2, k_seg_bundle_batch, kMergeBlock, 0);
        blocks = std::max(1, cus) * std::max(1, std::min(per, std::min(per1, per2)));
    }
    return blocks;
}

// ---- Morton order of the B-B1D query pieces --------------------------------------------
__device__ __forceinline__ uint32_t spread_bits10(uint32_t v)
{
    v &= 1023u;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// 30-bit Morton code of a piece's midpoint in the beam map's root box (0 for an unwritten piece)
__device__ __forceinline__ uint32_t piece_morton(const Bufs& b, uint32_t n_recs, uint2 pc)
{
    if (!(pc.x < n_recs && (pc.y & 0xffffu) < (pc.y >> 16))) return 0u;
    const WNode root = b.maps->wbeam.nodes[0];
    const SegRec* r = b.segq + pc.x;
    const float4 a = r->a, bb = r->b;
    float lo, hi, ta, tb;
    piece_range(a.w, bb.w, pc.y & 0xffffu, pc.y >> 16, lo, hi, ta, tb);
    const float t = 0.5f * (ta + tb);
    const float m[3] = {a.x + bb.x * t, a.y + bb.y * t, a.z + bb.z * t};
    const float o[3] = {root.ox, root.oy, root.oz};
    uint32_t q[3];
    for (int k = 0; k < 3; k++) {
        const float ext = 256.f * wnode_step(root.meta, k);
        const float u = ext > 0.f ? (m[k] - o[k]) / ext : 0.f;
        q[k] = uint32_t(fminf(fmaxf(u * 1024.f, 0.f), 1023.f));
    }
    return (spread_bits10(q[0]) << 2) | (spread_bits10(q[1]) << 1) | spread_bits10(q[2]);
}

__global__ void __launch_bounds__(256) k_piece_keys(Bufs b, int bounce, uint32_t n, uint32_t* keys, uint32_t* vals)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t n_recs = min(b.qcount[128 + bounce], b.seg_cap);
    keys[i] = piece_morton(b, n_recs, b.pieces[i]);
    vals[i] = i;
}

// ---- spatial order of whole units of 64 pieces ------------------------------------------
// The camera pass writes a wave's pieces as one run: 64 consecutive pieces are neighbouring
// rays' stretches of similar depth, which is what keeps one merge wave's node and leaf fetches
// shared. Units of 64 consecutive pieces are kept whole and ordered by the Morton code of their
// middle piece, so that the waves in flight at one time work on one region of the scene and the
// caches hold the part of the beam tree they walk. The sorted units are then cut into eight
// contiguous ranges, one per XCD (workgroups are dispatched to the XCDs round-robin, so block i
// starts on range i % 8 and each XCD's L2 holds one region's subtrees), or left as one range.
__global__ void __launch_bounds__(256) k_unit_keys(Bufs b, int bounce, uint32_t cap_units, uint32_t* keys,
                                                    uint32_t* vals, uint32_t* nu_dev)
{
    const uint32_t n = min(b.qcount[320 + bounce], b.piece_cap);
    const uint32_t nu = (n + 63u) / 64u;
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u == 0) *nu_dev = nu;
    if (u >= nu || u >= cap_units) return;
    const uint32_t n_recs = min(b.qcount[128 + bounce], b.seg_cap);
    keys[u] = piece_morton(b, n_recs, b.pieces[min(u * 64u + 32u, n - 1u)]);
    vals[u] = u;
}

__global__ void __launch_bounds__(256) k_unit_gather(Bufs b, int bounce, uint32_t cap_units, const uint32_t* vals,
                                                      int per_xcd)
{
    const uint32_t n = min(b.qcount[320 + bounce], b.piece_cap);
    const uint32_t nu = min((n + 63u) / 64u, cap_units);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        uint32_t* xw = b.xwork + size_t(bounce) * kXWork;
        for (uint32_t x = 0; x < uint32_t(kXcds); x++) {
            xw[192 + x] = per_xcd ? (x * nu / kXcds) * 64u : 0u;
            xw[200 + x] = per_xcd ? ((x + 1) * nu / kXcds) * 64u : (x == 0 ? nu * 64u : 0u);
            xw[208 + x] = 0u;
        }
    }
    if (i >= nu * 64u) return;
    const uint32_t src = vals[i >> 6] * 64u + (i & 63u);
    b.pieces_x[i] = src < n ? b.pieces[src] : make_uint2(~0u, 0u);
}

__global__ void __launch_bounds__(256) k_piece_gather(Bufs b, uint32_t n, const uint32_t* vals)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b.pieces_x[i] = b.pieces[vals[i]];
}

void launch_piece_keys(const Bufs& b, int bounce, uint32_t n, uint32_t* keys, uint32_t* vals, hipStream_t s)
{
    if (n) hipLaunchKernelGGL(k_piece_keys, dim3((n + 255) / 256), dim3(256), 0, s, b, bounce, n, keys, vals);
}

void launch_piece_gather(const Bufs& b, uint32_t n, const uint32_t* vals, hipStream_t s)
{
    if (n) hipLaunchKernelGGL(k_piece_gather, dim3((n + 255) / 256), dim3(256), 0, s, b, n, vals);
}

void launch_unit_keys(const Bufs& b, int bounce, uint32_t cap_units, uint32_t* keys, uint32_t* vals, uint32_t* nu_dev,
                      hipStream_t s)
{
    hipLaunchKernelGGL(k_unit_keys, dim3((cap_units + 255) / 256), dim3(256), 0, s, b, bounce, cap_units, keys, vals,
                       nu_dev);
}

void launch_unit_gather(const Bufs& b, int bounce, uint32_t cap_units, const uint32_t* vals, int per_xcd,
                        hipStream_t s)
{
    const size_t n = size_t(cap_units) * 64u;
    hipLaunchKernelGGL(k_unit_gather, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, s, b, bounce, cap_units, vals,
                       per_xcd);
}

void launch_seg_merge_batch(const BatchJob* jobs, uint32_t njobs, int kind, int bounce, hipStream_t s)
{
    const int blocks = merge_grid_blocks(0);
    if (kind == 0) {
        hipLaunchKernelGGL(k_seg_merge_batch<0>, dim3(blocks), dim3(kMergeBlock), 0, s, jobs, njobs, bounce);
    } else { // the jobs' bundles first (a job that walks none on this bounce is skipped)
        hipLaunchKernelGGL(k_seg_bundle_batch, dim3(bundle_grid_blocks()), dim3(kMergeBlock), 0, s, jobs, njobs,
                           bounce);
        hipLaunchKernelGGL(k_seg_merge_batch<1>, dim3(blocks), dim3(kMergeBlock), 0, s, jobs, njobs, bounce);
    }
}

void launch_seg_merge(const DScene& sc, const IterParams& p, const Bufs& b, int kind, int bounce, hipStream_t s)
{
    const int blocks = merge_grid_blocks(p.merge_bpc);
    if (kind == 0 && p.build_pb2d && p.cap_med > 0) {
        if (b.mislog) hipLaunchKernelGGL((k_seg_merge<0, true>), dim3(blocks), dim3(kMergeBlock), 0, s, sc, b, bounce, 0);
        else hipLaunchKernelGGL((k_seg_merge<0, false>), dim3(blocks), dim3(kMergeBlock), 0, s, sc, b, bounce, 0);
    }
    if (kind == 1 && p.build_bb1d && p.cap_beam > 0) {
        if (p.piece_order == 1) {
            hipLaunchKernelGGL(k_piece_cells, dim3(1024), dim3(256), 0, s, b, bounce);
            hipLaunchKernelGGL(k_piece_offsets, dim3(1), dim3(64), 0, s, b, bounce);
            hipLaunchKernelGGL(k_piece_scatter, dim3(1024), dim3(256), 0, s, b, bounce);
        }
        if (bounce < p.packet_bounces) {
            if (b.mislog)
                hipLaunchKernelGGL((k_seg_merge_packet<true>), dim3(blocks), dim3(kMergeBlock), 0, s, sc, b, bounce);
            else
                hipLaunchKernelGGL((k_seg_merge_packet<false>), dim3(blocks), dim3(kMergeBlock), 0, s, sc, b, bounce);
        } else {
            // bundles first on the first bundle_bounces bounces; the per-piece walk then takes
            // the pieces they left
            const int left = bounce < p.bundle_bounces && p.cap_beam < kBundleMaxPrims ? 1 : 0;
            if (left) {
                const int bb = bundle_grid_blocks();
                if (b.mislog)
                    hipLaunchKernelGGL((k_seg_bundle<true>), dim3(bb), dim3(kMergeBlock), 0, s, sc, b, bounce);
                else
                    hipLaunchKernelGGL((k_seg_bundle<false>), dim3(bb), dim3(kMergeBlock), 0, s, sc, b, bounce);
            }
            if (b.mislog)
                hipLaunchKernelGGL((k_seg_merge<1, true>), dim3(blocks), dim3(kMergeBlock), 0, s, sc, b, bounce, left);
            else
                hipLaunchKernelGGL((k_seg_merge<1, false>), dim3(blocks), dim3(kMergeBlock), 0, s, sc, b, bounce, left);
        }
    }
}

} // namespace upbp
