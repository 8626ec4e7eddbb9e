// Segment merge kernels: P-B2D (camera segment x medium photon point) and B-B1D (camera
// segment x photon sub-beam) for the segment queries one camera bounce appended.
//
// Semantics: nolmoonen/gpuupbp gpuupbp/kernel/pde/functs_pb2d.cuh:57-342 with
// program_pb2d.cu:31-46, and functs_bb1d.cuh:57-397 with program_bb1d.cu:31-48.
//
// Execution: a persistent grid. Each lane owns one query at a time and walks the LBVH one
// node per loop trip; a lane whose query is done takes the next one from a global work
// counter (one atomic per wave), so lanes stay busy while queries differ in cost by orders
// of magnitude. Accepted (query, photon) pairs go to a per-wave LDS queue; whenever 64 are
// queued the whole wave evaluates them together, so the long estimator body (attenuation,
// phase function, both-side MIS updates) runs without divergence. Each evaluated pair adds
// its weighted contribution to the query's camera path with float atomics.
#include "estimators.h"
#include "kernels.h"

#include <algorithm>

namespace upbp {

namespace {

constexpr int kMergeBlock = 256;
constexpr int kMergeWaves = kMergeBlock / 64;
constexpr uint32_t kPairCap = 192; // 64 pending + up to 2 accepts per lane in one trip

struct PairQueue {
    uint32_t q[kPairCap];    // segment query index
    uint32_t slot[kPairCap]; // light vertex slot (P-B2D) or beam slot (B-B1D)
    float x[kPairCap], y[kPairCap], z[kPairCap], w[kPairCap];
};

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

__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Leaf test of one query against one photon (sorted leaf k). Returns whether the pair is
// accepted and its parameters: P-B2D (idist, ir2); B-B1D (t camera, t beam, distance, sin).
template <int KIND>
__device__ __forceinline__ bool leaf_accept(const IterParams& p, const PhotonMaps& M, uint32_t k, f3 ro, f3 rd,
                                            float dmin, float dmax, int med, uint32_t& slot, float& x, float& y,
                                            float& z, float& w)
{
    if (KIND == 0) { // program_pb2d.cu:31-46
        const float4 pt = M.med_pts[k];
        const f3 ce = mk(pt.x, pt.y, pt.z);
        if (!pred_point_beam(ro, rd, dmax, ce, p.r2_pb2d)) return false;
        if (!pred_bre(x, y, ro, rd, dmin, dmax, ce, p.r2_pb2d)) return false;
        slot = uint32_t(__float_as_int(pt.w));
        z = w = 0.f;
        return true;
    } else { // program_bb1d.cu:31-48 on the whole beam, then sub-beam ownership of t beam
        const float4 g0 = M.beam_geo[3 * k], g1 = M.beam_geo[3 * k + 1];
        if (!pred_beam_beam(ro, rd, dmin, dmax, mk(g0.x, g0.y, g0.z), mk(g1.x, g1.y, g1.z), 0.f, g0.w, p.r2_bb1d, z,
                            w, x, y))
            return false;
        const float4 g2 = M.beam_geo[3 * k + 2];
        if (!(y >= g2.x && y < g2.y)) return false;
        // B-B1D subset of the light paths and the same-medium test (functs_bb1d.cuh:65)
        if (!(__float_as_int(g2.w) & SB_BB1D) || __float_as_int(g2.z) != med) return false;
        slot = uint32_t(__float_as_int(g1.w));
        return true;
    }
}

template <int KIND>
__global__ void __launch_bounds__(kMergeBlock) k_seg_merge(DScene sc, IterParams p, Bufs b, int bounce)
{
    __shared__ PairQueue pq[kMergeWaves];
    PairQueue& Q = pq[threadIdx.x >> 6];
    const uint32_t lane = __lane_id();
    const uint32_t total = min(b.qcount[128 + bounce], b.seg_cap);
    uint32_t* head = &b.qcount[192 + 2 * bounce + KIND];
    const Bvh& bvh = KIND == 0 ? b.maps.med : b.maps.beam;
    const float norm = KIND == 0 ? p.norm_pb2d : p.norm_bb1d;

    uint32_t qn = 0;        // pairs queued in Q (same value in every lane)
    bool exhausted = false; // no queries left to take (same value in every lane)
    int cur = -1;           // this lane's query
    f3 ro = mk(0.f), rd = mk(0.f);
    float dmin = 0.f, dmax = 0.f;
    int med = -1;
    RayBox rb = make_raybox(mk(0.f), mk(1.f), 0.f, 0.f);
    int node = 0, sp = 0;
    int stack[kBvhStack];
    uint32_t n_nodes = 0, n_cands = 0, n_acc = 0, n_err = 0;

    for (;;) {
        if (!exhausted) { // lanes without a query take the next ones
            const bool need = cur < 0;
            const uint64_t m = __ballot(need);
            if (m) {
                const int leader = __ffsll((unsigned long long)m) - 1;
                uint32_t base = 0;
                if (int(lane) == leader) base = atomicAdd(head, uint32_t(__popcll(m)));
                base = __shfl(base, leader);
                if (need) {
                    const uint32_t qi = base + uint32_t(__popcll(m & lanes_below()));
                    if (qi < total) {
                        const SegRec* r = b.segq + qi;
                        const float4 a = r->a, bb = r->b, d = r->d;
                        ro = mk(a.x, a.y, a.z);
                        dmin = a.w;
                        rd = mk(bb.x, bb.y, bb.z);
                        dmax = bb.w;
                        med = __float_as_int(d.z);
                        rb = make_raybox(ro, rd, dmin, dmax);
                        node = 0;
                        sp = 0;
                        cur = int(qi);
                    }
                }
                if (base + uint32_t(__popcll(m)) >= total) exhausted = true;
            }
        }
        const bool active = cur >= 0;
        const bool any_active = __ballot(active) != 0;
        bool acc0 = false, acc1 = false;
        uint32_t s0 = 0, s1 = 0, qnow = 0;
        float x0 = 0.f, y0 = 0.f, z0 = 0.f, w0 = 0.f, x1 = 0.f, y1 = 0.f, z1 = 0.f, w1 = 0.f;
        if (active) { // one node of this lane's traversal
            const BNode nd = bvh.nodes[node];
            n_nodes++;
            float e0, e1;
            const bool h0 = slab(rb, nd.lx0, nd.ly0, nd.lz0, nd.ux0, nd.uy0, nd.uz0, rb.tmax, e0);
            const bool h1 = slab(rb, nd.lx1, nd.ly1, nd.lz1, nd.ux1, nd.uy1, nd.uz1, rb.tmax, e1);
            int next = -1;
            if (h0) {
                if (nd.c0 < 0) {
                    n_cands++;
                    acc0 = leaf_accept<KIND>(p, b.maps, uint32_t(~nd.c0), ro, rd, dmin, dmax, med, s0, x0, y0, z0, w0);
                } else {
                    next = nd.c0;
                }
            }
            if (h1) {
                if (nd.c1 < 0) {
                    n_cands++;
                    acc1 = leaf_accept<KIND>(p, b.maps, uint32_t(~nd.c1), ro, rd, dmin, dmax, med, s1, x1, y1, z1, w1);
                } else if (next < 0) {
                    next = nd.c1;
                } else if (sp < kBvhStack) {
                    stack[sp++] = nd.c1;
                } else {
                    n_err++; // traversal stack exhausted: the query is abandoned and counted
                    next = -1;
                    sp = 0;
                }
            }
            if (next < 0 && sp > 0) next = stack[--sp];
            node = next;
            qnow = uint32_t(cur);
            if (next < 0) cur = -1;
            n_acc += uint32_t(acc0) + uint32_t(acc1);
        }
        // all lanes take part in the queue appends (ballots over the whole wave)
        pq_push(Q, qn, acc0, qnow, s0, x0, y0, z0, w0);
        pq_push(Q, qn, acc1, qnow, s1, x1, y1, z1, w1);
        // evaluate 64 queued pairs at a time; at the end, whatever is left
        while (qn >= 64u || (!any_active && qn > 0u)) {
            const uint32_t n = qn < 64u ? qn : 64u;
            wave_lds_sync();
            if (lane < n) {
                const uint32_t e = qn - n + lane;
                const uint32_t qi = Q.q[e], slot = Q.slot[e];
                const float x = Q.x[e], y = Q.y[e], z = Q.z[e], w = Q.w[e];
                SegQ q;
                StateWeights cw;
                f3 tatt;
                uint32_t pix;
                load_segrec(b.segq + qi, q, cw, tatt, pix);
                f3 acc = mk(0.f);
                if (KIND == 0) {
                    LVertex lv;
                    load_vertex(&b.verts[slot], lv);
                    eval_pb2d(sc, p, acc, q, cw, lv, x, y);
                } else {
                    LBeam lb;
                    load_beam(&b.beams[slot], lb);
                    eval_bb1d(sc, p, acc, q, cw, lb, z, w, x, y);
                }
                const f3 c = (tatt * norm) * acc;
                if (c.x != 0.f) atomicAdd(&b.mcolor[3 * pix], c.x);
                if (c.y != 0.f) atomicAdd(&b.mcolor[3 * pix + 1], c.y);
                if (c.z != 0.f) atomicAdd(&b.mcolor[3 * pix + 2], c.z);
            }
            wave_lds_sync();
            qn -= n;
        }
        if (!any_active && exhausted) break;
    }
    const uint32_t tn = wave_sum(n_nodes), tc = wave_sum(n_cands), ta = wave_sum(n_acc), te = wave_sum(n_err);
    if (lane == 0) {
        stat_add(b.stats, STAT_SEG_NODES, tn);
        stat_add(b.stats, STAT_SEG_CANDS, tc);
        stat_add(b.stats, STAT_ACCEPT, ta);
        stat_add(b.stats, KIND == 0 ? STAT_ACC_PB2D : STAT_ACC_BB1D, ta);
        stat_add(b.stats, STAT_ERRORS, te);
    }
}

} // namespace

int merge_grid_blocks()
{
    static int blocks = 0;
    if (blocks == 0) {
        int dev = 0, cus = 0, per0 = 0, per1 = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per0, k_seg_merge<0>, kMergeBlock, 0);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per1, k_seg_merge<1>, kMergeBlock, 0);
        const int per = std::max(1, std::min(per0, per1));
        blocks = std::max(1, cus) * per;
    }
    return blocks;
}

void launch_seg_merge(const DScene& sc, const IterParams& p, const Bufs& b, int kind, int bounce, hipStream_t s)
{
    const int blocks = merge_grid_blocks();
    if (kind == 0 && p.build_pb2d && b.maps.med.n > 0)
        hipLaunchKernelGGL(k_seg_merge<0>, dim3(blocks), dim3(kMergeBlock), 0, s, sc, p, b, bounce);
    if (kind == 1 && p.build_bb1d && b.maps.beam.n > 0)
        hipLaunchKernelGGL(k_seg_merge<1>, dim3(blocks), dim3(kMergeBlock), 0, s, sc, p, b, bounce);
}

} // namespace upbp
