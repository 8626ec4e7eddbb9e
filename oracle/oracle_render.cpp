// Oracle renderer: CPU BVH, intersector with the medium priority stack, light and
// camera subpaths, the five photon-density estimators with constant-time MIS.
// TEST INFRASTRUCTURE ONLY (see oracle.h). Citations: nolmoonen/gpuupbp.
#include "oracle_core.h"
#include "oracle_mis.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <thread>

namespace orc {

// ---- CPU BVH over boxes (median split) --------------------------------------
struct Box {
    V lo{kInf, kInf, kInf}, hi{-kInf, -kInf, -kInf};
    void grow(V p)
    {
        lo = {std::fmin(lo.x, p.x), std::fmin(lo.y, p.y), std::fmin(lo.z, p.z)};
        hi = {std::fmax(hi.x, p.x), std::fmax(hi.y, p.y), std::fmax(hi.z, p.z)};
    }
    void grow(const Box& b)
    {
        grow(b.lo);
        grow(b.hi);
    }
};

struct Bvh {
    struct Node {
        Box b;
        int left = -1, right = -1, start = 0, count = 0;
    };
    std::vector<Node> nodes;
    std::vector<uint32_t> order;
    std::vector<Box> boxes;

    void build(std::vector<Box> in)
    {
        boxes = std::move(in);
        nodes.clear();
        order.resize(boxes.size());
        for (size_t i = 0; i < order.size(); i++) order[i] = uint32_t(i);
        if (boxes.empty()) return;
        std::vector<V> cent(boxes.size());
        for (size_t i = 0; i < boxes.size(); i++) cent[i] = (boxes[i].lo + boxes[i].hi) * .5f;
        nodes.reserve(2 * boxes.size());
        build_rec(0, int(boxes.size()), cent);
    }
    int build_rec(int start, int end, const std::vector<V>& cent)
    {
        const int id = int(nodes.size());
        nodes.emplace_back();
        Box b, cb;
        for (int i = start; i < end; i++) {
            b.grow(boxes[order[size_t(i)]]);
            cb.grow(cent[order[size_t(i)]]);
        }
        // conservative inflation so that box tests never reject what a predicate accepts
        const V ext = b.hi - b.lo;
        const float pad = 1e-5f * std::fmax(maxc(ext), 1e-3f) + 1e-6f * std::fmax(maxc(V{std::fabs(b.lo.x), std::fabs(b.lo.y), std::fabs(b.lo.z)}),
                                                                                   maxc(V{std::fabs(b.hi.x), std::fabs(b.hi.y), std::fabs(b.hi.z)}));
        b.lo = b.lo - v3(pad);
        b.hi = b.hi + v3(pad);
        nodes[size_t(id)].b = b;
        if (end - start <= 4) {
            nodes[size_t(id)].start = start;
            nodes[size_t(id)].count = end - start;
            return id;
        }
        const V ce = cb.hi - cb.lo;
        const int axis = ce.x > ce.y ? (ce.x > ce.z ? 0 : 2) : (ce.y > ce.z ? 1 : 2);
        const int mid = (start + end) / 2;
        std::nth_element(order.begin() + start, order.begin() + mid, order.begin() + end,
                         [&](uint32_t a, uint32_t c) { return comp(cent[a], axis) < comp(cent[c], axis); });
        const int l = build_rec(start, mid, cent);
        const int r = build_rec(mid, end, cent);
        nodes[size_t(id)].left = l;
        nodes[size_t(id)].right = r;
        return id;
    }
    static bool slab(const Box& b, V o, V inv, float tmin, float tmax)
    {
        float t0 = tmin, t1 = tmax;
        const float lo[3] = {b.lo.x, b.lo.y, b.lo.z}, hi[3] = {b.hi.x, b.hi.y, b.hi.z};
        const float oo[3] = {o.x, o.y, o.z}, id[3] = {inv.x, inv.y, inv.z};
        for (int a = 0; a < 3; a++) {
            float tn = (lo[a] - oo[a]) * id[a], tf = (hi[a] - oo[a]) * id[a];
            if (std::isnan(tn) || std::isnan(tf)) { // origin on a slab plane with zero direction
                if (oo[a] < lo[a] || oo[a] > hi[a]) return false;
                continue;
            }
            if (tn > tf) std::swap(tn, tf);
            t0 = std::fmax(t0, tn);
            t1 = std::fmin(t1, tf);
            if (t0 > t1) return false;
        }
        return true;
    }
    template <typename F>
    void query_point(V q, F&& fn) const
    {
        if (nodes.empty()) return;
        int stack[128];
        int sp = 0;
        stack[sp++] = 0;
        while (sp) {
            const Node& n = nodes[size_t(stack[--sp])];
            if (q.x < n.b.lo.x || q.y < n.b.lo.y || q.z < n.b.lo.z || q.x > n.b.hi.x || q.y > n.b.hi.y ||
                q.z > n.b.hi.z)
                continue;
            if (n.count) {
                for (int i = 0; i < n.count; i++) fn(order[size_t(n.start + i)]);
            } else {
                stack[sp++] = n.left;
                stack[sp++] = n.right;
            }
        }
    }
    // fn(prim, tmax&) may shrink tmax (closest hit) or leave it (all hits)
    template <typename F>
    void query_segment(V o, V d, float tmin, float tmax, F&& fn) const
    {
        if (nodes.empty()) return;
        const V inv{1.f / d.x, 1.f / d.y, 1.f / d.z};
        int stack[128];
        int sp = 0;
        stack[sp++] = 0;
        float t_hi = tmax;
        while (sp) {
            const Node& n = nodes[size_t(stack[--sp])];
            if (!slab(n.b, o, inv, std::fmin(tmin, 0.f) - 1e-4f, t_hi * 1.0001f + 1e-4f)) continue;
            if (n.count) {
                for (int i = 0; i < n.count; i++) fn(order[size_t(n.start + i)], t_hi);
            } else {
                stack[sp++] = n.left;
                stack[sp++] = n.right;
            }
        }
    }
};

// Moller-Trumbore ray/triangle test (Moller & Trumbore 1997); replaces OptiX's built-in triangles.
static bool tri_hit(const Tri& t, V o, V d, float tmin, float tmax, float& th, float& bu, float& bv)
{
    const V e1 = t.p1 - t.p0, e2 = t.p2 - t.p0;
    const V pv = cross(d, e2);
    const float det = dot(e1, pv);
    if (det == 0.f) return false;
    const float inv = 1.f / det;
    const V tv = o - t.p0;
    const float u = dot(tv, pv) * inv;
    if (u < 0.f || u > 1.f) return false;
    const V qv = cross(tv, e1);
    const float v = dot(d, qv) * inv;
    if (v < 0.f || u + v > 1.f) return false;
    const float tt = dot(e2, qv) * inv;
    if (!(tt > tmin && tt < tmax)) return false;
    th = tt;
    bu = u;
    bv = v;
    return true;
}

static Box tri_box(const Tri& t)
{
    Box b;
    b.grow(t.p0);
    b.grow(t.p1);
    b.grow(t.p2);
    return b;
}

// ---- range-query predicates: kernel/intersection.cuh:33-173 (contraction off) ----
// kernel/intersection.cuh:55-68
static bool point_beam(V ro, V rd, float rt, V ce, float ra2)
{
    const V eo = ce - ro;
    const float t = dot(eo, rd);
    if (t <= 0.f || t >= rt) return false;
    const V tmp = ro + rd * t - ce;
    return dot(tmp, tmp) <= ra2;
}
// kernel/intersection.cuh:71-75
static bool point_point(V p0, V p1, float ra2)
{
    const V d = p0 - p1;
    return dot(d, d) <= ra2;
}
// kernel/intersection.cuh:88-110
static bool bre(float& isect_dist, float& isect_rad2, V ro, V rd, float min_t, float max_t, V p, float r2)
{
    const V op = p - ro;
    const float dist = dot(op, rd);
    if (dist > min_t && dist < max_t) {
        const V ir = ro + rd * dist - p;
        const float ir2 = dot(ir, ir);
        if (ir2 <= r2) {
            isect_dist = dist;
            isect_rad2 = ir2;
            return true;
        }
    }
    return false;
}
// kernel/intersection.cuh:124-173
static bool beam_beam(V o1, V d1, float min1, float max1, V o2, V d2, float min2, float max2, float maxd2,
                      float& distance, float& sin_theta, float& t1, float& t2)
{
    const V c = cross(d1, d2);
    const float s2 = dot(c, c);
    const float ad = dot((o2 - o1), c);
    if (ad * ad >= maxd2 * s2) return false;
    const float d1d2 = dot(d1, d2);
    const float d1d2_sqr = d1d2 * d1d2;
    const float d1d2_sqr_m1 = d1d2_sqr - 1.f;
    if (d1d2_sqr_m1 < 1e-5f && d1d2_sqr_m1 > -1e-5f) return false;
    const float d1o1 = dot(d1, o1);
    const float d1o2 = dot(d1, o2);
    t1 = (d1o1 - d1o2 - d1d2 * (dot(d2, o1) - dot(d2, o2))) / d1d2_sqr_m1;
    if (t1 <= min1 || t1 >= max1) return false;
    t2 = (t1 + d1o1 - d1o2) / d1d2;
    if (t2 <= min2 || t2 >= max2 || std::isnan(t2)) return false;
    sin_theta = std::sqrt(s2);
    distance = std::fabs(ad) / sin_theta;
    return true;
}

// ---- renderer data ----------------------------------------------------------
struct VertexWeights {
    float d_shared = 0, d_bpt = 0, d_pde = 0;
    float for_inv = 0, rev_inv = 0, for_ratio = 0, rev_ratio = 0;
    bool prev_in_medium = false, prev_specular = false;
};
struct StateWeights {
    VertexWeights w;
    float last_sin_theta = 0, d_bpt_a = 0, d_bpt_b = 0, d_pde_a = 0, d_pde_b = 0;
};
struct SubpathState {
    BStack stack;
    V origin, dir, throughput;
    uint32_t path_length;
    bool infinite, specular_path;
    StateWeights sw;
};
struct LightVertex {
    V hit, throughput;
    uint32_t path_length;
    uint32_t path = 0; // light path index
    BSDF bsdf;
    VertexWeights w;
    bool in_medium, behind_surf;
};
struct LightBeam {
    V origin, dir, throughput;
    int med;
    float length, for_pdf, rev_pdf;
    unsigned flags;
    uint32_t path_idx;
    StateWeights sw;
    uint32_t path_length;
    bool infinite;
};
struct VolSeg {
    V att, emi;
    float dmin, dmax, pdf_for, pdf_rev;
    int med;
};
struct VolLite {
    float dmin, dmax;
    int med;
};
struct LiteIsect {
    float dist;
    int mat, med;
    bool enter;
    bool operator==(const LiteIsect& o) const { return mat == o.mat && med == o.med; }
};

// Per-iteration parameters (shared/launch_params.h:44-198 subset)
struct Params {
    uint32_t min_len, max_len, pbt, qbt;
    float n_bpt, n_bb1d, n_bp2d, n_pb2d, n_pp3d, n_surf;
    bool do_bpt, do_bb1d, do_bp2d, do_pb2d, do_pp3d, do_surf;
    bool build_bb1d, build_bp2d, build_pb2d, build_pp3d, build_surf;
    bool trace_light, trace_camera, connect_from_surf, ignore_spec, use_shading_normal;
    uint32_t n_pixels, n_light, techniques, bb1d_used;
    float norm_bb1d, norm_bp2d, norm_pb2d, norm_pp3d, norm_surf;
    float mf_bb1d, mf_bp2d, mf_pb2d, mf_pp3d, mf_surf;
    float r_bb1d, r_bp2d, r_pb2d, r_pp3d, r_surf, rad_point_med, rad_beam;
    float r2_bb1d, r2_bp2d, r2_pb2d, r2_pp3d, r2_surf;
    uint64_t iter_key;
    uint32_t iteration;
};

struct Ctx {
    const Scene* sc;
    Params p;
    Bvh bvh_real, bvh_imag;
    // light pass outputs
    std::vector<LightVertex> verts;
    std::vector<uint32_t> path_begin, path_end;
    std::vector<uint32_t> surf_idx, med_idx;
    std::vector<LightBeam> beams;
    Bvh bvh_surf, bvh_med, bvh_beam;
    std::atomic<uint64_t> errors{0};
    // MIS logging (FP64 full-path oracle): emission points of the light paths, records per
    // logged camera path (pixel -> slot) and per logged light path
    bool mis_on = false;
    std::vector<V> light_origin;
    std::vector<int> light_pick;
    std::vector<int32_t> mis_slot;
    std::vector<std::vector<MisRec>>* mis_cam = nullptr;
    uint32_t mis_lt_paths = 0;
    std::vector<std::vector<MisRec>>* mis_lt = nullptr;
    // learned-proposal training samples (per pixel) and which scatterings feed them
    std::vector<std::vector<std::array<float, 7>>>* train = nullptr;
    bool train_media = false, train_surfaces = false;
};

// camera-subpath vertex kept for the MIS log: its BSDF (incoming from the next vertex toward
// the camera) and the distance pdfs of the edge it was reached by, in camera direction
struct CamVert {
    V pos;
    BSDF bsdf;
    float rsf_c, rsr_c;
};

struct MisLogger {
    std::vector<MisRec>* out = nullptr;
    std::vector<CamVert> cam;                  // current camera path, first hit first
    const std::vector<LightVertex>* lverts = nullptr; // light tracing: vertices of the current light path
    size_t lbegin = 0;
};
static thread_local MisLogger* g_mis = nullptr;

static int priority_of(const Scene& sc, int mat) { return mat >= 0 ? sc.mats[size_t(mat)].priority : kGlobalPriority; }
static bool real_geom(const Scene& sc, int mat) { return mat >= 0 && sc.mats[size_t(mat)].real; }
static const Medium& medium_at(const Scene& sc, int id)
{
    return sc.media[size_t(std::fmin(float(id), float(sc.media.size() - 1)))];
}

// kernel/intersector.cu:103-144 (fill_isect) + closest-hit over real triangles
static bool trace_real(const Ctx& c, V o, V d, float tmin, float tmax, Isect& out)
{
    const Scene& sc = *c.sc;
    int best = -1;
    float bt = 0, bu = 0, bv = 0;
    c.bvh_real.query_segment(o, d, tmin, tmax, [&](uint32_t i, float& thi) {
        float t, u, v;
        if (tri_hit(sc.real[i], o, d, tmin, thi, t, u, v)) {
            if (best < 0 || t < bt || (t == bt && int(i) < best)) {
                best = int(i);
                bt = t;
                bu = u;
                bv = v;
                thi = t;
            }
        }
    });
    // analytic spheres after the triangles, exhaustively (a handful per predefined scene)
    int best_s = -1;
    V sn{0.f, 0.f, 0.f};
    for (size_t i = 0; i < sc.sreal.size(); i++) {
        float th[2];
        V nh[2];
        if (sphere_hit(sc.sreal[i], o, d, tmin, best >= 0 || best_s >= 0 ? bt : tmax, th, nh)) {
            best_s = int(i);
            bt = th[0];
            sn = nh[0];
        }
    }
    if (best_s >= 0) {
        const Sph& q = sc.sreal[size_t(best_s)];
        out.dist = bt;
        out.mat = q.mat;
        out.med = sc.mats[size_t(q.mat)].med_idx;
        out.light = q.light;
        out.normal = sn;
        out.shading = sn;
        out.enter = dot(sn, d) < 0.f;
        return true;
    }
    if (best < 0) return false;
    const Tri& t = sc.real[size_t(best)];
    out.dist = bt;
    out.mat = t.mat;
    out.med = sc.mats[size_t(t.mat)].med_idx;
    out.light = t.light;
    out.normal = t.face;
    out.shading = c.p.use_shading_normal ? normalize(t.n0 * (1.f - bu - bv) + t.n1 * bu + t.n2 * bv) : t.face;
    out.enter = dot(t.face, d) < 0.f;
    return true;
}

// all imaginary hits in (tmin, tmax) ordered by distance (intersector.cu:152-167)
static bool trace_imag(const Ctx& c, V o, V d, float tmin, float tmax, PStack<LiteIsect, float, kMaxImaginary>& st)
{
    const Scene& sc = *c.sc;
    bool overflow = false;
    struct Cand {
        float t;
        uint32_t order; // triangles by index, then sphere crossings in test order
        int mat;
        bool enter;
        bool operator<(const Cand& o) const { return t < o.t || (t == o.t && order < o.order); }
    };
    std::vector<Cand> hits;
    c.bvh_imag.query_segment(o, d, tmin, tmax, [&](uint32_t i, float&) {
        float t, u, v;
        if (tri_hit(sc.imag[i], o, d, tmin, tmax, t, u, v))
            hits.push_back({t, i, sc.imag[i].mat, dot(sc.imag[i].face, d) < 0.f});
    });
    // imaginary spheres report both crossings (intersector.cu:78-93 under an any-hit program)
    uint32_t order = uint32_t(sc.imag.size());
    for (const Sph& q : sc.simag) {
        float th[2];
        V nh[2];
        const int k = sphere_hit(q, o, d, tmin, tmax, th, nh);
        for (int j = 0; j < k; j++) hits.push_back({th[j], order++, q.mat, dot(nh[j], d) < 0.f});
    }
    std::sort(hits.begin(), hits.end());
    for (auto& h : hits) {
        LiteIsect li{h.t, h.mat, sc.mats[size_t(h.mat)].med_idx, h.enter};
        if (!st.push(li, h.t)) {
            overflow = true;
            break;
        }
    }
    return !overflow;
}

// kernel/intersector.cuh:45-393
static int intersect_(const Ctx& c, bool& hit, V ro, V rd, Isect& result, BStack& bstack, bool occlusion,
                      unsigned flags, std::vector<VolSeg>& segs, std::vector<VolLite>& lite, Stream* rng)
{
    const Scene& sc = *c.sc;
    hit = false;
    const bool sample_media = !occlusion;
    Isect tmp = result;
    float dist = 0.f;
    PStack<LiteIsect, float, kMaxIntersections> hwm_unused;
    (void)hwm_unused;
    std::vector<LiteIsect> hits_with_material;
    const int top_priority = bstack.top_priority();
    const float first_eps = (flags & ORIGIN_IN_MEDIUM) ? 0.f : kEpsRay;
    float t_max, miss_dist;
    if (occlusion) {
        t_max = tmp.dist - kEpsRay;
        miss_dist = tmp.dist;
    } else {
        t_max = kInf;
        miss_dist = kInf;
    }
    BStack stack_copy = bstack;
    float t_min = first_eps;
    bool not_yet_real;
    do {
        not_yet_real = false;
        Isect h;
        hit = trace_real(c, ro, rd, t_min, t_max, h);
        if (hit) tmp = h;
        else tmp.dist = kInf;
        if (hit) {
            dist = tmp.dist;
            const int priority = priority_of(sc, tmp.mat);
            const Boundary top = stack_copy.top();
            if (tmp.med < 0) tmp.med = top.med;
            if (priority < top_priority && priority != kThinWall) {
                const Boundary el{tmp.mat, tmp.med};
                if (tmp.enter) {
                    if (hits_with_material.size() == size_t(kMaxIntersections)) return ERR_INTERSECTION_ARRAY;
                    hits_with_material.push_back({tmp.dist, tmp.mat, tmp.med, tmp.enter});
                    if (!stack_copy.push(el, priority)) return ERR_BOUNDARY_STACK;
                } else if (stack_copy.pop(el, priority)) {
                    if (hits_with_material.size() == size_t(kMaxIntersections)) return ERR_INTERSECTION_ARRAY;
                    hits_with_material.push_back({tmp.dist, tmp.mat, tmp.med, tmp.enter});
                }
                not_yet_real = true;
                t_min = dist + kEpsRay;
                // the reference shortens the remaining occlusion range by the hit distance
                // again (intersector.cuh:175-177); kept as is
                t_max = occlusion ? miss_dist - dist - kEpsRay : kInf;
            }
        }
    } while (not_yet_real && t_max > t_min);
    if (!hit || not_yet_real) {
        hit = false;
        dist = miss_dist;
    }
    result = tmp;
    result.dist = dist;

    bool scattered = false;
    PStack<LiteIsect, float, kMaxImaginary> imag;
    bool have_imag = false;
    if (!real_geom(sc, bstack.top().mat) && dist > first_eps + kEpsRay) {
        have_imag = true;
        if (!trace_imag(c, ro, rd, first_eps, dist - kEpsRay, imag)) return ERR_INTERSECTION_STACK;
        stack_copy = bstack;
        float dmin = 0.f;
        for (int i = 0; i < imag.n; i++) {
            const LiteIsect& is = imag.el[i];
            const int priority = priority_of(sc, is.mat);
            const Boundary el{is.mat, is.med};
            if (is.enter) {
                if (priority >= stack_copy.top_priority()) {
                    if (lite.size() == size_t(kMaxSegments)) return ERR_VOL_SEGMENT_ARRAY;
                    lite.push_back({dmin, is.dist, stack_copy.top().med});
                    dmin = is.dist;
                }
                if (!stack_copy.push(el, priority)) return ERR_BOUNDARY_STACK;
            } else {
                if (priority == stack_copy.top_priority() && el == stack_copy.top()) {
                    if (lite.size() == size_t(kMaxSegments)) return ERR_VOL_SEGMENT_ARRAY;
                    lite.push_back({dmin, is.dist, stack_copy.top().med});
                    dmin = is.dist;
                }
                stack_copy.pop(el, priority);
            }
        }
        if (dmin < dist) {
            if (lite.size() == size_t(kMaxSegments)) return ERR_VOL_SEGMENT_ARRAY;
            lite.push_back({dmin, dist, stack_copy.top().med});
        }
    } else {
        if (lite.size() == size_t(kMaxSegments)) return ERR_VOL_SEGMENT_ARRAY;
        lite.push_back({0.f, dist, bstack.top().med});
    }

    const bool origin_in_medium = flags & ORIGIN_IN_MEDIUM;
    const bool end_in_medium = flags & END_IN_MEDIUM;
    int popped = 0;
    for (size_t i = 0; i < lite.size(); i++) {
        const VolLite& sg = lite[i];
        const int mid = sg.med;
        if (mid < 0) continue;
        const Medium& m = medium_at(sc, mid);
        unsigned f = 0;
        if (origin_in_medium && sg.dmin == 0) f |= ORIGIN_IN_MEDIUM;
        if (end_in_medium && sg.dmax == dist) f |= END_IN_MEDIUM;
        float fp = 1.f, rp = 1.f;
        float dmax = sg.dmax;
        if (m.has_scattering) {
            if (sample_media) {
                float r;
                do {
                    r = rng->next();
                } while (r == 0.f);
                const float to_next = sg.dmax - sg.dmin;
                const float dm = sample_ray(m, to_next, r, &fp, f, &rp);
                if (dm < to_next) {
                    result.dist = sg.dmin + dm;
                    result.mat = -1;
                    result.med = mid;
                    result.light = -1;
                    result.normal = v3(0.f);
                    result.enter = false;
                    hit = true;
                    scattered = true;
                    dmax = result.dist;
                }
            } else {
                fp = ray_sample_pdf(m, sg.dmin, sg.dmax, f, &rp);
            }
        }
        if (segs.size() == size_t(kMaxSegments)) return ERR_VOL_SEGMENT_ARRAY;
        const float len = dmax - sg.dmin;
        segs.push_back({eval_attenuation(m, len), m.emission * len, sg.dmin, dmax, fp, rp, mid});
        if (have_imag) {
            for (int j = popped; j < imag.n; j++) {
                const LiteIsect& is = imag.el[j];
                if (is.dist > dmax) break;
                if (is.enter) {
                    if (!bstack.push({is.mat, is.med}, priority_of(sc, is.mat))) return ERR_BOUNDARY_STACK;
                } else if (bstack.n > 1) {
                    bstack.pop({is.mat, is.med}, priority_of(sc, is.mat));
                }
                popped++;
            }
        }
        if (scattered) break;
    }
    for (const LiteIsect& h : hits_with_material) {
        if (h.dist > result.dist) break;
        if (h.enter) {
            if (!bstack.push({h.mat, h.med}, priority_of(sc, h.mat))) return ERR_BOUNDARY_STACK;
        } else if (bstack.n > 1) {
            bstack.pop({h.mat, h.med}, priority_of(sc, h.mat));
        }
    }
    return OK;
}

// kernel/intersector.cuh:396-418
static int occluded(const Ctx& c, bool& hit, V p, V d, float tmax, const BStack& bs, unsigned flags,
                    std::vector<VolSeg>& segs)
{
    BStack copy = bs;
    Isect tmp;
    tmp.dist = tmax;
    std::vector<VolLite> lite;
    return intersect_(c, hit, p, d, tmp, copy, true, flags, segs, lite, nullptr);
}

// kernel/scene.cuh:113-152
static float relative_ior(const Scene& sc, const Isect& is, const BStack& st)
{
    if (!is.on_surface()) return -1.f;
    const float ior1 = sc.mats[size_t(is.mat)].ior;
    if (ior1 < 0.f) return -1.f;
    float ior2 = 1.f;
    const int m2 = is.enter ? st.top().mat : (st.n > 1 ? st.second().mat : st.top().mat);
    if (m2 >= 0 && sc.mats[size_t(m2)].real) ior2 = sc.mats[size_t(m2)].ior;
    if (ior2 < 0.f) return -1.f;
    return is.enter ? ior2 / ior1 : ior1 / ior2;
}

// kernel/path_weight.cuh:35-51
static float pde_factor(const Params& p, bool in_med, bool spec, float for_inv, float for_ratio, float rev_inv,
                        float rev_ratio, float sin_theta)
{
    const float fp = p.pbt == UPBP_BEAM_LONG ? for_inv : for_ratio;
    const float rp = p.qbt == UPBP_BEAM_LONG ? rev_inv : rev_ratio;
    return float(!in_med && !spec) * p.mf_surf +
           float(in_med) * (p.mf_pp3d + p.mf_pb2d * rp + p.mf_bp2d * fp + p.mf_bb1d * sin_theta * fp * rp);
}

static float acc_for(const std::vector<VolSeg>& s)
{
    float r = 1.f;
    for (auto& x : s) r *= x.pdf_for;
    return r;
}
static float acc_rev(const std::vector<VolSeg>& s)
{
    float r = 1.f;
    for (auto& x : s) r *= x.pdf_rev;
    return r;
}
static V acc_att(const std::vector<VolSeg>& s)
{
    V r = v3(1.f);
    for (auto& x : s) r *= x.att;
    return r;
}

static int init_stack(const Scene& sc, BStack& st, int mat, int med)
{
    st.n = 0;
    st.push({-1, int(sc.global_medium)}, kGlobalPriority);
    if (mat != -1 && med != -1)
        if (!st.push({mat, med}, priority_of(sc, mat))) return ERR_BOUNDARY_STACK;
    return OK;
}

// kernel/functs_shared.cuh:37-154
// guide_den (camera paths with training samples on): lum(throughput f) cos / cp of a continuing
// scattering in a medium or at a guidable surface whose training is on, else 0 -- the
// denominator of the training weight (the GPU's sample_scattering, passes.h)
static int sample_scattering(const Ctx& c, bool& cont, const BSDF& b, V hit, const Isect& is, SubpathState& s,
                             Stream& rng, float* guide_den = nullptr)
{
    const Params& p = c.p;
    const float r0 = rng.next(), r1 = rng.next(), r2 = rng.next(), r3 = rng.next();
    float for_pdf, cos_out, sin_theta;
    unsigned ev;
    V dir;
    const V f = bsdf_sample(b, r0, r1, r2, dir, for_pdf, cos_out, *c.sc, ev, sin_theta);
    if (black_or_neg(f)) {
        cont = false;
        return OK;
    }
    const bool specular = (ev & BSDF::SPECULAR) != 0;
    float rev_pdf = for_pdf;
    if (!specular) rev_pdf = bsdf_pdf(b, dir, *c.sc, true);
    const float cp = b.cont_prob;
    if (cp == 0 || (cp < 1.f && r3 > cp)) {
        cont = false;
        return OK;
    }
    for_pdf *= cp;
    rev_pdf *= cp;
    if (guide_den) {
        const bool trains = b.in_medium() ? c.train_media
                                          : (c.train_surfaces && surface_guidable(c.sc->mats[size_t(b.mat)]));
        *guide_den = trains ? luminance(s.throughput * f) * cos_out / cp : 0.f;
    }
    VertexWeights& w = s.sw.w;
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
    s.origin = hit;
    s.dir = dir;
    s.throughput *= f * (cos_out / for_pdf);
    w.prev_in_medium = b.in_medium();
    w.prev_specular = specular;
    s.sw.last_sin_theta = sin_theta;
    if (ev & BSDF::REFRACT) { // kernel/scene.cuh:75-96
        if (is.enter) {
            if (is.med >= 0)
                if (!s.stack.push({is.mat, is.med}, priority_of(*c.sc, is.mat))) return ERR_BOUNDARY_STACK;
        } else if (s.stack.n > 1) {
            s.stack.n--;
        }
    }
    cont = true;
    return OK;
}

// ---- MIS log --------------------------------------------------------------------
// Primitive densities of the full path x_0 = light point, x_1..x_nl = light vertices lv,
// then camera vertices cam[nc-1] .. cam[0], then the camera. With `merge` the last light
// vertex and the last camera vertex are the same path vertex (the merge); otherwise the
// edge between the two sides is a connection whose distance pdfs are conn_rsf (toward
// the camera side) and conn_rsr (toward the light side).
static void mis_record(const Ctx& c, int tech, float weight, V x0, int light_id, const LightVertex* lv, int nl,
                       const CamVert* cam, int nc, bool merge, float conn_rsf, float conn_rsr, float ref_light_dist = 0.f,
                       int lpath = -1)
{
    if (!g_mis || !g_mis->out) return;
    const Scene& sc = *c.sc;
    MisRec r;
    r.tech = tech;
    r.weight = weight;
    r.ref_light_dist = ref_light_dist;
    r.lpath = lpath >= 0 ? lpath : light_id;
    const int k = nl + nc + 1 - (merge ? 1 : 0);
    if (k < 2) return;
    r.k = k;
    r.s = merge ? nl : nl + 1;
    if (tech == MT_LIGHT_HIT) r.s = 0;
    const size_t nv = size_t(k) + 1;
    r.type.assign(nv, MV_SURFACE);
    r.fpdf.assign(nv, 0.0);
    r.rpdf.assign(nv, 0.0);
    r.cin.assign(nv, 1.0);
    r.cout.assign(nv, 1.0);
    r.sin_t.assign(nv, 0.0);
    r.ratio.assign(nv, 0.0);
    r.dist.assign(size_t(k), 0.0);
    r.rsf.assign(size_t(k), 1.0);
    r.rsr.assign(size_t(k), 1.0);
    // positions: left (arriving edge) and right (leaving edge) differ only at a merge vertex
    std::vector<V> pl(nv), pr(nv);
    pl[0] = pr[0] = x0;
    for (int j = 1; j <= nl; j++) pl[size_t(j)] = pr[size_t(j)] = lv[j - 1].hit;
    auto cam_index = [&](int i) { return k - 1 - i; }; // path index of cam[i]
    for (int i = 0; i < nc; i++) {
        const int q = cam_index(i);
        if (merge && q == nl) pr[size_t(q)] = cam[i].pos;
        else pl[size_t(q)] = pr[size_t(q)] = cam[i].pos;
    }
    pl[size_t(k)] = pr[size_t(k)] = sc.cam.origin;
    r.type[0] = MV_LIGHT;
    r.type[size_t(k)] = MV_CAMERA;
    // edges
    for (int j = 0; j < k; j++) {
        const V d = pl[size_t(j + 1)] - pr[size_t(j)];
        r.dist[size_t(j)] = std::sqrt(double(d.x) * d.x + double(d.y) * d.y + double(d.z) * d.z);
    }
    for (int j = 1; j <= nl; j++) { // arriving edge of light vertex x_j
        r.rsf[size_t(j - 1)] = 1.0 / double(lv[j - 1].w.for_inv);
        r.rsr[size_t(j - 1)] = 1.0 / double(lv[j - 1].w.rev_inv);
    }
    for (int i = 0; i < nc; i++) { // arriving camera edge of cam[i]: x_q <- x_{q+1}
        const int q = cam_index(i);
        r.rsf[size_t(q)] = cam[i].rsr_c;
        r.rsr[size_t(q)] = cam[i].rsf_c;
    }
    if (!merge) { // the connection edge x_nl -> x_{nl+1}
        r.rsf[size_t(nl)] = conn_rsf;
        r.rsr[size_t(nl)] = conn_rsr;
    }
    // interior vertices
    auto norm_dir = [](V a, V b) { const V d = b - a; return d / std::sqrt(dot(d, d)); };
    for (int q = 1; q < k; q++) {
        float cosg, fp, rp, st;
        const int li = q - 1;         // light vertex index when q <= nl
        const int ci = k - 1 - q;     // camera vertex index when q > nl (or q == nl at a merge)
        if (merge && q == nl) {
            const BSDF& cb = cam[ci].bsdf;
            const BSDF& lb = lv[li].bsdf;
            bsdf_evaluate(cb, lb.dir_fix_world(), sc, cosg, &fp, &rp, &st);
            r.rpdf[size_t(q)] = double(fp) * cb.cont_prob;
            r.fpdf[size_t(q)] = double(rp) * lb.cont_prob;
            r.cin[size_t(q)] = lb.d_factor();
            r.cout[size_t(q)] = cb.d_factor();
            r.sin_t[size_t(q)] = st;
            r.type[size_t(q)] = lb.in_medium() ? MV_MEDIUM : MV_SURFACE;
            if (lb.in_medium()) r.ratio[size_t(q)] = 1.0 / sc.media[size_t(lb.med)].min_pos_att;
        } else if (q <= nl) {
            const BSDF& b = lv[li].bsdf;
            bsdf_evaluate(b, norm_dir(pr[size_t(q)], pl[size_t(q + 1)]), sc, cosg, &fp, &rp, &st);
            r.fpdf[size_t(q)] = double(fp) * b.cont_prob;
            r.rpdf[size_t(q)] = double(rp) * b.cont_prob;
            r.cin[size_t(q)] = b.d_factor();
            r.cout[size_t(q)] = std::fabs(cosg);
            r.sin_t[size_t(q)] = st;
            r.type[size_t(q)] = b.in_medium() ? MV_MEDIUM : MV_SURFACE;
            if (b.in_medium()) r.ratio[size_t(q)] = 1.0 / sc.media[size_t(b.med)].min_pos_att;
        } else {
            const BSDF& b = cam[ci].bsdf;
            bsdf_evaluate(b, norm_dir(pl[size_t(q)], pr[size_t(q - 1)]), sc, cosg, &fp, &rp, &st);
            r.rpdf[size_t(q)] = double(fp) * b.cont_prob;
            r.fpdf[size_t(q)] = double(rp) * b.cont_prob;
            r.cin[size_t(q)] = std::fabs(cosg);
            r.cout[size_t(q)] = b.d_factor();
            r.sin_t[size_t(q)] = st;
            r.type[size_t(q)] = b.in_medium() ? MV_MEDIUM : MV_SURFACE;
            if (b.in_medium()) r.ratio[size_t(q)] = 1.0 / sc.media[size_t(b.med)].min_pos_att;
        }
    }
    // light (area lights only) and camera
    const Light& l = sc.lights[size_t(light_id)];
    const V to1 = norm_dir(pr[0], pl[1]);
    r.cos_light = std::fabs(double(dot(l.frame.z, to1)));
    r.dpdf = double(l.inv_area) / double(sc.lights.size());
    r.eps = r.cos_light / 3.14159265358979323846;
    const V tok = norm_dir(pl[size_t(k)], pr[size_t(k - 1)]);
    const double ca = double(dot(sc.cam.direction, tok));
    const double ipd = double(sc.cam.image_plane_dist) / ca;
    r.cam_pdf = ipd * ipd / ca;
    g_mis->out->push_back(std::move(r));
}

// medium "BSDF" (phase function) at a point of a ray with direction dir
static BSDF medium_bsdf(const Ctx& c, V dir, int med, bool from_light, V pos)
{
    Isect is;
    is.dist = 1.f;
    is.mat = -1;
    is.med = med;
    BSDF b;
    bsdf_setup(b, dir, is, *c.sc, from_light, -1.f);
    b.pos = pos;
    return b;
}

// light vertices x_1 .. x_{n} of light path `path` (false unless every one was stored)
static bool light_prefix(const Ctx& c, uint32_t path, uint32_t n, std::vector<LightVertex>& out)
{
    out.clear();
    const uint32_t b = c.path_begin[path], e = c.path_end[path];
    if (e - b < n) return false;
    for (uint32_t j = 0; j < n; j++) {
        if (c.verts[b + j].path_length != j + 1) return false;
        out.push_back(c.verts[b + j]);
    }
    return true;
}

// a point on a photon beam as the light vertex of a beam merge
static LightVertex beam_point(const Ctx& c, const LightBeam& lb, float t, float rsf, float rsr)
{
    LightVertex v;
    v.hit = lb.origin + lb.dir * t;
    v.path_length = lb.path_length;
    v.path = lb.path_idx;
    v.bsdf = medium_bsdf(c, lb.dir, lb.med, true, v.hit);
    v.w.for_inv = 1.f / rsf;
    v.w.rev_inv = 1.f / rsr;
    v.in_medium = true;
    return v;
}

// ---- light subpath ----------------------------------------------------------
struct LightOut {
    std::vector<LightVertex> verts;
    std::vector<LightBeam> beams;
    std::vector<float> fb; // splats, W*H*3
};

// kernel/functs_light.cuh:159-247
static void add_beams(const Ctx& c, V ro, V rd, V thr, uint32_t plen, unsigned flags, const std::vector<VolSeg>& segs,
                      const std::vector<VolLite>& lite, uint32_t path_idx, const SubpathState& s,
                      std::vector<LightBeam>& out)
{
    const Scene& sc = *c.sc;
    V t = thr;
    float fpdf = 1.f, rpdf = 1.f;
    auto emit_beam = [&](float dmin, float dmax, const Medium& m, int mid, size_t i) {
        LightBeam b;
        b.med = mid;
        b.origin = ro + rd * dmin;
        b.dir = rd;
        b.length = std::fmin(dmax - dmin, m.max_beam_length);
        b.for_pdf = fpdf;
        b.rev_pdf = rpdf;
        b.flags = END_IN_MEDIUM | (i == 0 ? flags : 0u);
        b.throughput = t;
        b.path_idx = path_idx;
        b.sw = s.sw;
        b.path_length = plen;
        b.infinite = s.infinite;
        out.push_back(b);
    };
    if (c.p.pbt == UPBP_BEAM_SHORT) {
        for (size_t i = 0; i < segs.size(); i++) {
            const VolSeg& sg = segs[i];
            const Medium& m = sc.media[size_t(sg.med)];
            if (m.has_scattering) emit_beam(sg.dmin, sg.dmax, m, sg.med, i);
            t *= sg.att / sg.pdf_for;
            fpdf *= sg.pdf_for;
            rpdf *= sg.pdf_rev;
        }
    } else {
        for (size_t i = 0; i < lite.size(); i++) {
            const VolLite& sg = lite[i];
            if (sg.med < 0) continue;
            const Medium& m = sc.media[size_t(sg.med)];
            if (m.has_scattering) emit_beam(sg.dmin, sg.dmax, m, sg.med, i);
            t *= eval_attenuation(m, sg.dmax - sg.dmin);
            float srp;
            const float sfp = ray_sample_pdf(m, sg.dmin, sg.dmax, i == 0 ? flags : 0u, &srp);
            fpdf *= sfp;
            rpdf *= srp;
        }
    }
}

// camera mapping helpers (kernel/camera.cuh:37-76)
static bool check_raster(const Camera& cam, float x, float y) { return x >= 0 && y >= 0 && x < cam.res_x && y < cam.res_y; }

static void splat(std::vector<float>& fb, const Camera& cam, float x, float y, V c)
{
    if (x < 0.f || x >= cam.res_x || y < 0.f || y >= cam.res_y) return;
    const size_t px = size_t(int(x)) + size_t(int(y)) * size_t(cam.res_x);
    fb[3 * px] += c.x;
    fb[3 * px + 1] += c.y;
    fb[3 * px + 2] += c.z;
}

// kernel/functs_light.cuh:251-388
static int connect_to_camera(const Ctx& c, const SubpathState& s, V hit, const BSDF& b, std::vector<float>& fb)
{
    const Scene& sc = *c.sc;
    const Camera& cam = sc.cam;
    V to_cam = cam.origin - hit;
    if (dot(cam.direction, -to_cam) <= 0.f) return OK;
    const V ip = transform_point(cam.world_to_raster, hit);
    if (!check_raster(cam, ip.x, ip.y)) return OK;
    const float d2 = dot(to_cam, to_cam);
    const float dist = std::sqrt(d2);
    to_cam = to_cam / dist;
    float cos_to_cam, fpdf, rpdf, sin_theta;
    const V f = bsdf_evaluate(b, to_cam, sc, cos_to_cam, &fpdf, &rpdf, &sin_theta);
    if (black_or_neg(f)) return OK;
    rpdf *= b.cont_prob;
    const float cos_at_cam = dot(cam.direction, -to_cam);
    const float ip_dist = cam.image_plane_dist / cos_at_cam;
    const float i2sa = ip_dist * ip_dist / cos_at_cam;
    const float i2s = i2sa * std::fabs(cos_to_cam) / (dist * dist);
    const float cam_pdf_a = i2s;
    std::vector<VolSeg> segs;
    bool occ;
    int ret = occluded(c, occ, hit, to_cam, dist, s.stack, b.in_medium() ? ORIGIN_IN_MEDIUM : 0u, segs);
    if (ret != OK) return ret;
    if (occ) return OK;
    float rsr = 1.f;
    V att = v3(1.f);
    if (!segs.empty()) {
        rsr = acc_rev(segs);
        att = acc_att(segs);
        if (!positive(att)) return OK;
    }
    const VertexWeights& wl = s.sw.w;
    const float f_s = pde_factor(c.p, b.in_medium(), false, wl.for_inv, wl.for_ratio, 1.f / rsr, wl.rev_ratio, sin_theta);
    const float bpt_s1 = float(!wl.prev_specular);
    const float w_light = (cam_pdf_a / float(c.p.n_light)) * rsr *
                          (f_s + bpt_s1 * wl.d_shared + (rpdf / wl.rev_inv) * wl.d_bpt);
    const float mis = 1.f / (w_light + 1.f);
    if (g_mis && g_mis->lverts && g_mis->lverts->size() > g_mis->lbegin) {
        const LightVertex* lv = g_mis->lverts->data() + g_mis->lbegin;
        const int nl = int(g_mis->lverts->size() - g_mis->lbegin);
        if (lv[nl - 1].path_length == uint32_t(nl)) // every vertex of the path was stored
            mis_record(c, MT_LT, mis, c.light_origin[lv->path], c.light_pick[lv->path], lv, nl, nullptr, 0, false,
                       segs.empty() ? 1.f : acc_for(segs), rsr, 0.f, int(lv->path));
    }
    const float s2i = 1.f / i2s;
    const V contrib = mis * s.throughput * f * att / (float(c.p.n_light) * s2i);
    if (black_or_neg(contrib)) return OK;
    splat(fb, cam, ip.x, ip.y, contrib);
    return OK;
}

// kernel/functs_light.cuh:86-156
static int generate_light_sample(const Ctx& c, SubpathState& s, Stream& rng, int* pick = nullptr)
{
    const Scene& sc = *c.sc;
    const float lc = float(sc.lights.size());
    const int id = int(rng.next() * lc);
    if (pick) *pick = id;
    const float d0 = rng.next(), d1 = rng.next(), p0 = rng.next(), p1 = rng.next();
    const Light& l = sc.lights[size_t(id)];
    float epdf, dpdf, cosl;
    s.throughput = light_emit(l, sc, d0, d1, p0, p1, s.origin, s.dir, epdf, &dpdf, &cosl);
    epdf /= lc;
    dpdf /= lc;
    s.throughput /= epdf;
    s.path_length = 1;
    s.infinite = !light_finite(l);
    s.specular_path = false;
    s.sw = StateWeights();
    s.sw.w.d_shared = dpdf / epdf;
    s.sw.w.d_bpt = !light_delta(l) ? (light_finite(l) ? cosl : 1.f) / epdf : 0.f;
    s.sw.w.d_pde = s.sw.w.d_bpt * c.p.n_bpt;
    s.sw.w.prev_specular = false;
    s.sw.w.prev_in_medium = false;
    return init_stack(sc, s.stack, l.mat_id, l.med_id);
}

// kernel/trace_light.cu:42-325
static void trace_light_path(const Ctx& c, uint32_t path_idx, LightOut& out)
{
    const Scene& sc = *c.sc;
    const Params& p = c.p;
    const bool store_beams = (p.do_bb1d && path_idx < p.bb1d_used) || p.do_bp2d;
    SubpathState s;
    Stream emit(kDomainLight, path_idx, p.iter_key, 0, P_EMIT);
    int pick = 0;
    int ret = generate_light_sample(c, s, emit, &pick);
    if (s.infinite && sc.media[sc.global_medium].has_attenuation) return;
    MisLogger logger;
    if (c.mis_on) {
        const_cast<Ctx&>(c).light_origin[path_idx] = s.origin;
        const_cast<Ctx&>(c).light_pick[path_idx] = pick;
        if (path_idx < c.mis_lt_paths) {
            logger.out = &(*c.mis_lt)[path_idx];
            logger.lverts = &out.verts;
            logger.lbegin = out.verts.size();
            g_mis = &logger;
        }
    }
    bool origin_in_medium = false;
    std::vector<VolSeg> segs;
    std::vector<VolLite> lite;
    for (; ret == OK; s.path_length++) {
        const uint32_t seg = s.path_length - 1;
        Stream mrng(kDomainLight, path_idx, p.iter_key, seg, P_MEDIUM);
        Isect is;
        segs.clear();
        lite.clear();
        bool hit;
        const unsigned rf = origin_in_medium ? ORIGIN_IN_MEDIUM : 0u;
        ret = intersect_(c, hit, s.origin, s.dir, is, s.stack, false, rf, segs, lite, &mrng);
        if (ret != OK) break;
        if (store_beams) add_beams(c, s.origin, s.dir, s.throughput, s.path_length, rf, segs, lite, path_idx, s, out.beams);
        if (!hit) break;
        float rsf = 1.f, rsr = 1.f;
        if (!segs.empty()) {
            rsf = acc_for(segs);
            rsr = acc_rev(segs);
            s.throughput *= acc_att(segs) / rsf;
        }
        if (black_or_neg(s.throughput)) break;
        VertexWeights& w = s.sw.w;
        if (s.path_length > 1u) {
            const float fi = pde_factor(p, w.prev_in_medium, w.prev_specular, w.for_inv, w.for_ratio, 1.f / rsr,
                                        w.rev_ratio, s.sw.last_sin_theta);
            w.d_bpt = s.sw.d_bpt_a * fi + s.sw.d_bpt_b;
            w.d_pde = s.sw.d_pde_a * fi + s.sw.d_pde_b;
        }
        w.d_shared /= rsf;
        w.d_bpt /= rsf;
        w.d_pde /= rsf;
        BSDF b;
        bsdf_setup(b, s.dir, is, sc, true, relative_ior(sc, is, s.stack));
        if (!b.valid()) break;
        float for_ratio = 0.f, rev_ratio = 0.f;
        if (b.in_medium()) for_ratio = rev_ratio = 1.f / sc.media[size_t(b.med)].min_pos_att;
        const V hitp = s.origin + s.dir * is.dist;
        b.pos = hitp;
        origin_in_medium = is.in_medium();
        {
            const float d = b.d_factor();
            if (!(s.path_length == 1 && s.infinite)) w.d_shared *= is.dist * is.dist;
            w.d_shared /= d;
            w.d_bpt /= d;
            w.d_pde /= d;
            w.for_inv = 1.f / rsf;
            w.rev_inv = 1.f / rsr;
            w.for_ratio = for_ratio;
            w.rev_ratio = rev_ratio;
        }
        if (!b.delta()) {
            LightVertex lv;
            lv.hit = hitp;
            lv.throughput = s.throughput;
            lv.path_length = s.path_length;
            lv.path = path_idx;
            lv.bsdf = b;
            lv.in_medium = origin_in_medium;
            lv.behind_surf = false;
            if (origin_in_medium && s.stack.n > 0) {
                const int m = s.stack.top().mat;
                if (m >= 0 && sc.mats[size_t(m)].real) lv.behind_surf = true;
            }
            lv.w = w;
            out.verts.push_back(lv);
        }
        if (p.do_bpt && !b.delta() && (b.in_medium() || p.connect_from_surf)) {
            if (s.path_length + 1 >= p.min_len) {
                ret = connect_to_camera(c, s, hitp, b, out.fb);
                if (ret != OK) break;
            }
        }
        if (s.path_length + 2 > p.max_len) break;
        bool cont;
        Stream srng(kDomainLight, path_idx, p.iter_key, seg, P_SCATTER);
        ret = sample_scattering(c, cont, b, hitp, is, s, srng);
        if (ret != OK) break;
        if (!cont) break;
    }
    g_mis = nullptr;
    if (ret != OK) const_cast<Ctx&>(c).errors++;
}

// ---- camera subpath estimators ---------------------------------------------
// kernel/functs_camera.cuh:107-165
static V light_radiance_mis(const Ctx& c, const Light& l, const SubpathState& s, V hit, float rsf, float rsr)
{
    const Params& p = c.p;
    const Scene& sc = *c.sc;
    (void)rsf;
    if (s.specular_path && p.ignore_spec) return v3(0.f);
    const float pick = 1.f / float(sc.lights.size());
    float dpdf_a = 0.f, epdf_w = 0.f;
    const V rad = light_radiance(l, sc, s.dir, hit, &dpdf_a, &epdf_w);
    if (black_or_neg(rad)) return v3(0.f);
    if (s.path_length == 1) return rad;
    if (p.techniques && !(p.techniques & UPBP_BPT)) return s.specular_path ? rad : v3(0.f);
    dpdf_a *= pick;
    epdf_w *= pick;
    const VertexWeights& w = s.sw.w;
    const float bpt_t1 = float(!w.prev_specular);
    const float w_cam = dpdf_a * bpt_t1 * w.d_shared + epdf_w * rsr * w.d_bpt;
    if (g_mis && l.type == L_AREA)
        mis_record(c, MT_LIGHT_HIT, 1.f / (1.f + w_cam), hit, int(&l - sc.lights.data()), nullptr, 0,
                   g_mis->cam.data(), int(g_mis->cam.size()), false, rsr, rsf);
    return rad * (1.f / (1.f + w_cam));
}

// kernel/functs_camera.cuh:170-344
static int direct_illumination(const Ctx& c, V& contrib, const SubpathState& s, V hit, const BSDF& b, Stream& rng)
{
    const Params& p = c.p;
    const Scene& sc = *c.sc;
    const float pick = 1.f / float(sc.lights.size());
    const int id = int(rng.next() * float(sc.lights.size()));
    const float u0 = rng.next(), u1 = rng.next();
    const Light& l = sc.lights[size_t(id)];
    contrib = v3(0.f);
    if (!light_finite(l) && sc.media[sc.global_medium].has_attenuation) return OK;
    V dir;
    float dist, dpdf_w, epdf_w, cos_at;
    const V rad = light_illuminate(l, sc, hit, u0, u1, dir, dist, dpdf_w, &epdf_w, &cos_at);
    if (black_or_neg(rad)) return OK;
    float fpdf, rpdf, cos_to, sin_theta;
    const V f = bsdf_evaluate(b, dir, sc, cos_to, &fpdf, &rpdf, &sin_theta);
    if (black_or_neg(f)) return OK;
    const float cp = b.cont_prob;
    fpdf *= light_delta(l) ? 0.f : cp;
    rpdf *= cp;
    std::vector<VolSeg> segs;
    bool occ;
    int ret = occluded(c, occ, hit, dir, dist, s.stack, b.in_medium() ? ORIGIN_IN_MEDIUM : 0u, segs);
    if (ret != OK) return ret;
    if (occ) return OK;
    float nrf = 1.f, nrr = 1.f;
    V natt = v3(1.f);
    if (!segs.empty()) {
        nrf = acc_for(segs);
        nrr = acc_rev(segs);
        natt = acc_att(segs);
        if (!positive(natt)) return OK;
    }
    (void)nrf;
    const float w_light = fpdf / (pick * dpdf_w);
    const VertexWeights& w = s.sw.w;
    const float ft = pde_factor(p, b.in_medium(), false, 1.f / nrr, w.rev_ratio, w.for_inv, w.for_ratio, sin_theta);
    const float bt1 = float(!w.prev_specular);
    const float w_cam = (epdf_w * cos_to / (dpdf_w * cos_at)) * nrr * (ft + bt1 * w.d_shared + (rpdf / w.rev_inv) * w.d_bpt);
    const float mis = 1.f / (w_light + 1.f + w_cam);
    if (g_mis && l.type == L_AREA)
        mis_record(c, MT_NEE, mis, hit + dir * dist, id, nullptr, 0, g_mis->cam.data(), int(g_mis->cam.size()), false,
                   nrr, nrf);
    contrib = ((cos_to / (pick * dpdf_w)) * (rad * natt * f)) * mis;
    if (black_or_neg(contrib)) contrib = v3(0.f);
    return OK;
}

// kernel/functs_camera.cuh:349-527
static int connect_vertices(const Ctx& c, V& contrib, const LightVertex& lv, const BSDF& bc, V chit,
                            const SubpathState& s)
{
    const Params& p = c.p;
    const Scene& sc = *c.sc;
    contrib = v3(0.f);
    V dir = lv.hit - chit;
    const float d2 = dot(dir, dir);
    const float dist = std::sqrt(d2);
    dir = dir / dist;
    float cos_c, cfp, crp, sin_c;
    const V fc = bsdf_evaluate(bc, dir, sc, cos_c, &cfp, &crp, &sin_c);
    if (black_or_neg(fc)) return OK;
    cfp *= bc.cont_prob;
    crp *= bc.cont_prob;
    float cos_l, lfp, lrp, sin_l;
    const V fl = bsdf_evaluate(lv.bsdf, -dir, sc, cos_l, &lfp, &lrp, &sin_l);
    if (black_or_neg(fl)) return OK;
    lfp *= lv.bsdf.cont_prob;
    lrp *= lv.bsdf.cont_prob;
    const float g = cos_l * cos_c / d2;
    if (g < 0.f) return OK;
    const float c_pdf_a = pdf_w_to_a(cfp, dist, cos_l);
    const float l_pdf_a = pdf_w_to_a(lfp, dist, cos_c);
    unsigned rf = 0;
    if (bc.in_medium()) rf |= ORIGIN_IN_MEDIUM;
    if (lv.in_medium) rf |= END_IN_MEDIUM;
    std::vector<VolSeg> segs;
    bool occ;
    int ret = occluded(c, occ, chit, dir, dist, s.stack, rf, segs);
    if (ret != OK) return ret;
    if (occ) return OK;
    float rsf = 1.f, rsr = 1.f;
    V att = v3(1.f);
    if (!segs.empty()) {
        rsf = acc_for(segs);
        rsr = acc_rev(segs);
        att = acc_att(segs);
        if (!positive(att)) return OK;
    }
    const VertexWeights& wl = lv.w;
    const VertexWeights& wc = s.sw.w;
    const float fs = pde_factor(p, lv.bsdf.in_medium(), false, wl.for_inv, wl.for_ratio, 1.f / rsf, wl.rev_ratio, sin_l);
    const float bs1 = float(!wl.prev_specular);
    const float w_light = c_pdf_a * rsf * (fs + bs1 * wl.d_shared + (lrp / wl.rev_inv) * wl.d_bpt);
    const float ft = pde_factor(p, bc.in_medium(), false, 1.f / rsr, wc.rev_ratio, wc.for_inv, wc.for_ratio, sin_c);
    const float bt1 = float(!wc.prev_specular);
    const float w_cam = l_pdf_a * rsr * (ft + bt1 * wc.d_shared + (crp / wc.rev_inv) * wc.d_bpt);
    const float mis = 1.f / (w_light + 1.f + w_cam);
    if (g_mis) {
        const LightVertex* first = c.verts.data() + c.path_begin[lv.path];
        const int nl = int(lv.path_length);
        if (first + (nl - 1) == &lv)
            mis_record(c, MT_BPT, mis, c.light_origin[lv.path], c.light_pick[lv.path], first, nl, g_mis->cam.data(),
                       int(g_mis->cam.size()), false, rsr, rsf, 0.f, int(lv.path));
    }
    contrib = (g * fc * fl * att) * mis;
    if (black_or_neg(contrib)) contrib = v3(0.f);
    return OK;
}

// kernel/pde/functs_pp.cuh:44-136
static void eval_pp(const Ctx& c, V& acc, const SubpathState& s, const BSDF& bc, const LightVertex& lv)
{
    const Params& p = c.p;
    const bool medium = lv.in_medium;
    const uint32_t len = lv.path_length + s.path_length;
    if (len > p.max_len || len < p.min_len) return;
    const V pdir = lv.bsdf.dir_fix_world();
    float cos_c, fpdf, rpdf, sin_theta;
    const V f = bsdf_evaluate(bc, pdir, *c.sc, cos_c, &fpdf, &rpdf, &sin_theta);
    if (black_or_neg(f)) return;
    fpdf *= bc.cont_prob;
    rpdf *= lv.bsdf.cont_prob;
    const VertexWeights& wl = lv.w;
    const VertexWeights& wc = s.sw.w;
    const float fpf = p.pbt == UPBP_BEAM_LONG ? wl.for_inv : wl.for_ratio;
    const float rpf = p.qbt == UPBP_BEAM_LONG ? wc.for_inv : wc.for_ratio;
    const float val = medium ? 1.f / p.mf_pp3d : 1.f / p.mf_surf;
    const float bs1 = float(!wl.prev_specular);
    const float w_light = val * (wl.d_shared * bs1 * p.n_bpt + fpdf * wl.d_pde / wl.rev_inv);
    const float w_local = 1.f + float(medium) * val *
                                    (p.mf_bb1d * sin_theta * fpf * rpf + p.mf_bp2d * fpf + p.mf_pb2d * rpf);
    const float bt1 = float(!wc.prev_specular);
    const float w_cam = val * (wc.d_shared * bt1 * p.n_bpt + rpdf * wc.d_pde / wc.rev_inv);
    const float mis = 1.f / (w_light + w_local + w_cam);
    if (g_mis) {
        const LightVertex* first = c.verts.data() + c.path_begin[lv.path];
        const int nl = int(lv.path_length);
        if (first + (nl - 1) == &lv)
            mis_record(c, medium ? MT_PP3D : MT_SURF, mis, c.light_origin[lv.path], c.light_pick[lv.path], first, nl,
                       g_mis->cam.data(), int(g_mis->cam.size()), true, 1.f, 1.f, 0.f, int(lv.path));
    }
    acc += mis * f * lv.throughput;
}

struct SegQuery {
    V ro, rd;
    float dmin, dmax;
    const Medium* med;
    int med_id;
    float rsf, rsr;
    unsigned flags;
    uint32_t cam_len;
    const StateWeights* cw;
};

// kernel/pde/functs_pb2d.cuh:57-201
static void eval_pb2d(const Ctx& c, V& acc, const SegQuery& q, const LightVertex& lv)
{
    const Params& p = c.p;
    float idist, ir2;
    if (!bre(idist, ir2, q.ro, q.rd, q.dmin, q.dmax, lv.hit, p.r2_pb2d)) return;
    const uint32_t len = lv.path_length + q.cam_len;
    if (len > p.max_len || len < p.min_len) return;
    const Medium& m = *q.med;
    V att = eval_attenuation(m, idist - q.dmin);
    const float pdf = comp(att, m.min_pos_idx);
    if (p.qbt == UPBP_BEAM_SHORT) att /= pdf;
    float rsf = m.min_pos_att * pdf;
    float rsr = (q.flags & ORIGIN_IN_MEDIUM) ? rsf : pdf;
    const float for_ratio = 1.f / m.min_pos_att, rev_ratio = for_ratio;
    if (!positive(att)) return;
    rsf *= q.rsf;
    rsr *= q.rsr;
    VertexWeights wc = q.cw->w;
    if (q.cam_len > 1u) {
        const float fi = pde_factor(p, wc.prev_in_medium, wc.prev_specular, 1.f / rsr, wc.rev_ratio, wc.for_inv,
                                    wc.for_ratio, q.cw->last_sin_theta);
        wc.d_pde = q.cw->d_pde_a * fi + q.cw->d_pde_b;
        wc.d_bpt = q.cw->d_bpt_a * fi + q.cw->d_bpt_b;
    }
    const V pdir = lv.bsdf.dir_fix_world();
    float sin_theta;
    const float hg = hg_pdf(-q.rd, pdir, m.mean_cosine, &sin_theta);
    float fpdf = cam_medium_pdf(*c.sc, m.mean_cosine, q.ro + q.rd * idist, q.rd, pdir); // camera technique
    float rpdf = hg;                                                      // light technique
    const V f = v3(hg);
    if (black_or_neg(f)) return;
    fpdf *= m.cont_prob;
    rpdf *= lv.bsdf.cont_prob;
    wc.d_shared *= idist * idist;
    wc.d_shared /= rsf;
    wc.d_bpt /= rsf;
    wc.d_pde /= rsf;
    wc.for_inv = 1.f / rsf;
    wc.rev_inv = 1.f / rsr;
    wc.for_ratio = for_ratio;
    wc.rev_ratio = rev_ratio;
    const float kernel = (1.f - ir2 / p.r2_pb2d) / (p.r2_pb2d * kPi * .5f);
    if (!positive(kernel)) return;
    const V unweighted = lv.throughput * att * m.scattering * f * kernel;
    if (black_or_neg(unweighted)) return;
    const VertexWeights& wl = lv.w;
    const float fpf = p.pbt == UPBP_BEAM_LONG ? wl.for_inv : wl.for_ratio;
    const float rpf = p.qbt == UPBP_BEAM_LONG ? wc.for_inv : wc.for_ratio;
    const float val = 1.f / (p.mf_pb2d * rpf);
    const float bs1 = float(!wl.prev_specular);
    const float w_light = val * (wl.d_shared * bs1 * p.n_bpt + fpdf * wl.d_pde / wl.rev_inv);
    const float w_local = 1.f + val * (p.mf_bb1d * sin_theta * fpf * rpf + p.mf_bp2d * fpf + p.mf_pp3d);
    const float bt1 = float(!wc.prev_specular);
    const float w_cam = val * (wc.d_shared * bt1 * p.n_bpt + rpdf * wc.d_pde / wc.rev_inv);
    if (g_mis) {
        std::vector<LightVertex> lp;
        if (light_prefix(c, lv.path, lv.path_length, lp)) {
            std::vector<CamVert> cam = g_mis->cam;
            cam.push_back({q.ro + q.rd * idist, medium_bsdf(c, q.rd, q.med_id, false, q.ro + q.rd * idist), rsf, rsr});
            mis_record(c, MT_PB2D, 1.f / (w_light + w_local + w_cam), c.light_origin[lv.path], c.light_pick[lv.path],
                       lp.data(), int(lp.size()), cam.data(), int(cam.size()), true, 1.f, 1.f, 0.f, int(lv.path));
        }
    }
    acc += (1.f / (w_light + w_local + w_cam)) * unweighted;
}

// kernel/pde/functs_bb1d.cuh:57-248
static void eval_bb1d(const Ctx& c, V& acc, const SegQuery& q, const LightBeam& lb)
{
    const Params& p = c.p;
    if (lb.med != q.med_id) return;
    float bbd, sin_theta, tc, tp;
    if (!beam_beam(q.ro, q.rd, q.dmin, q.dmax, lb.origin, lb.dir, 0.f, lb.length, p.r2_bb1d, bbd, sin_theta, tc, tp))
        return;
    const uint32_t len = lb.path_length + q.cam_len;
    if (len > p.max_len || len < p.min_len) return;
    const Medium& m = *q.med;
    V attc = eval_attenuation(m, tc - q.dmin);
    const float pdfc = comp(attc, m.min_pos_idx);
    if (p.qbt == UPBP_BEAM_SHORT) attc /= pdfc;
    float rsfc = m.min_pos_att * pdfc;
    float rsrc = (q.flags & ORIGIN_IN_MEDIUM) ? rsfc : pdfc;
    const float ratio = 1.f / m.min_pos_att;
    V attp = eval_attenuation(m, tp);
    const float pdfp = comp(attp, m.min_pos_idx);
    if (p.pbt == UPBP_BEAM_SHORT) attp /= pdfp;
    float rsfp = m.min_pos_att * pdfp;
    float rsrp = (lb.flags & ORIGIN_IN_MEDIUM) ? rsfp : pdfp;
    const V att = attc * attp;
    if (!positive(att)) return;
    rsfc *= q.rsf;
    rsrc *= q.rsr;
    VertexWeights wc = q.cw->w;
    if (q.cam_len > 1u) {
        const float fi = pde_factor(p, wc.prev_in_medium, wc.prev_specular, 1.f / rsrc, wc.rev_ratio, wc.for_inv,
                                    wc.for_ratio, q.cw->last_sin_theta);
        wc.d_pde = q.cw->d_pde_a * fi + q.cw->d_pde_b;
        wc.d_bpt = q.cw->d_bpt_a * fi + q.cw->d_bpt_b;
    }
    rsfp *= lb.for_pdf;
    rsrp *= lb.rev_pdf;
    VertexWeights wl = lb.sw.w;
    if (lb.path_length > 1u) {
        const float fi = pde_factor(p, wl.prev_in_medium, wl.prev_specular, wl.for_inv, wl.for_ratio, 1.f / rsrp,
                                    wl.rev_ratio, lb.sw.last_sin_theta);
        wl.d_pde = lb.sw.d_pde_a * fi + lb.sw.d_pde_b;
        wl.d_bpt = lb.sw.d_bpt_a * fi + lb.sw.d_bpt_b;
    }
    const float hg = hg_pdf(-q.rd, -lb.dir, m.mean_cosine, nullptr);
    float fpdf = cam_medium_pdf(*c.sc, m.mean_cosine, q.ro + q.rd * tc, q.rd, -lb.dir);
    float rpdf = hg;
    const V f = v3(hg);
    if (black_or_neg(f)) return;
    fpdf *= m.cont_prob;
    rpdf *= m.cont_prob;
    wc.d_shared *= tc * tc;
    wc.d_shared /= rsfc;
    wc.d_bpt /= rsfc;
    wc.d_pde /= rsfc;
    wc.for_inv = 1.f / rsfc;
    wc.rev_inv = 1.f / rsrc;
    wc.for_ratio = ratio;
    wc.rev_ratio = ratio;
    // the reference's light-side condition is !(len == 1 && !infinite) here
    // (functs_bb1d.cuh:191); kept as is
    if (!(lb.path_length == 1 && !lb.infinite)) wl.d_shared *= tp * tp;
    wl.d_shared /= rsfp;
    wl.d_bpt /= rsfp;
    wl.d_pde /= rsfp;
    wl.for_inv = 1.f / rsfp;
    wl.rev_inv = 1.f / rsrp;
    wl.for_ratio = ratio;
    wl.rev_ratio = ratio;
    const float kernel = (1.f - bbd * bbd / p.r2_bb1d) * 3.f / (4.f * p.r_bb1d * sin_theta);
    if (!positive(kernel)) return;
    const V unweighted = lb.throughput * m.scattering * att * f * kernel;
    if (black_or_neg(unweighted)) return;
    const float fpf = p.pbt == UPBP_BEAM_LONG ? wl.for_inv : wl.for_ratio;
    const float rpf = p.qbt == UPBP_BEAM_LONG ? wc.for_inv : wc.for_ratio;
    const float val = 1.f / (p.mf_bb1d * sin_theta * fpf * rpf);
    const float bs1 = float(!wl.prev_specular);
    const float w_light = val * (wl.d_shared * bs1 * p.n_bpt + fpdf * wl.d_pde / wl.rev_inv);
    const float w_local = 1.f + val * (p.mf_bp2d * fpf + p.mf_pb2d * rpf + p.mf_pp3d);
    const float bt1 = float(!wc.prev_specular);
    const float w_cam = val * (wc.d_shared * bt1 * p.n_bpt + rpdf * wc.d_pde / wc.rev_inv);
    if (g_mis) {
        std::vector<LightVertex> lp;
        if (light_prefix(c, lb.path_idx, lb.path_length - 1, lp)) {
            lp.push_back(beam_point(c, lb, tp, rsfp, rsrp));
            std::vector<CamVert> cam = g_mis->cam;
            cam.push_back({q.ro + q.rd * tc, medium_bsdf(c, q.rd, q.med_id, false, q.ro + q.rd * tc), rsfc, rsrc});
            const float ref_d = (lb.path_length == 1 && !lb.infinite) ? 1.f : tp;
            mis_record(c, MT_BB1D, 1.f / (w_light + w_local + w_cam), c.light_origin[lb.path_idx],
                       c.light_pick[lb.path_idx], lp.data(), int(lp.size()), cam.data(), int(cam.size()), true, 1.f,
                       1.f, ref_d, int(lb.path_idx));
        }
    }
    acc += (1.f / (w_light + w_local + w_cam)) * unweighted;
}

// kernel/pde/functs_bp2d.cuh:51-198
static void eval_bp2d(const Ctx& c, V& acc, const SubpathState& s, const BSDF& bc, V hit, const LightBeam& lb)
{
    const Params& p = c.p;
    float idist, ir2;
    if (!bre(idist, ir2, lb.origin, lb.dir, 0.f, lb.length, hit, p.r2_bp2d)) return;
    const uint32_t len = lb.path_length + s.path_length;
    if (len > p.max_len || len < p.min_len) return;
    const Medium& m = c.sc->media[size_t(lb.med)];
    V att = eval_attenuation(m, idist);
    const float pdf = comp(att, m.min_pos_idx);
    if (p.pbt == UPBP_BEAM_SHORT) att /= pdf;
    float rsf = m.min_pos_att * pdf;
    float rsr = (lb.flags & ORIGIN_IN_MEDIUM) ? rsf : pdf;
    const float ratio = 1.f / m.min_pos_att;
    if (!positive(att)) return;
    rsf *= lb.for_pdf;
    rsr *= lb.rev_pdf;
    VertexWeights wl = lb.sw.w;
    if (lb.path_length > 1u) {
        const float fi = pde_factor(p, wl.prev_in_medium, wl.prev_specular, wl.for_inv, wl.for_ratio, 1.f / rsr,
                                    wl.rev_ratio, lb.sw.last_sin_theta);
        wl.d_pde = lb.sw.d_pde_a * fi + lb.sw.d_pde_b;
        wl.d_bpt = lb.sw.d_bpt_a * fi + lb.sw.d_bpt_b;
    }
    const V cdir = bc.dir_fix_world();
    float sin_theta;
    const float hg = hg_pdf(cdir, -lb.dir, m.mean_cosine, &sin_theta);
    float fpdf = cam_medium_pdf(*c.sc, m.mean_cosine, hit, -cdir, -lb.dir);
    float rpdf = hg;
    const V f = v3(hg);
    if (black_or_neg(f)) return;
    fpdf *= bc.cont_prob;
    rpdf *= m.cont_prob;
    if (!(lb.path_length == 1 && lb.infinite)) wl.d_shared *= idist * idist;
    wl.d_shared /= rsf;
    wl.d_bpt /= rsf;
    wl.d_pde /= rsf;
    wl.for_inv = 1.f / rsf;
    wl.rev_inv = 1.f / rsr;
    wl.for_ratio = ratio;
    wl.rev_ratio = ratio;
    const float kernel = (1.f - ir2 / p.r2_bp2d) / (p.r2_bp2d * kPi * .5f);
    if (!positive(kernel)) return;
    const V unweighted = lb.throughput * att * m.scattering * f * kernel;
    if (black_or_neg(unweighted)) return;
    const VertexWeights& wc = s.sw.w;
    const float fpf = p.pbt == UPBP_BEAM_LONG ? wl.for_inv : wl.for_ratio;
    const float rpf = p.qbt == UPBP_BEAM_LONG ? wc.for_inv : wc.for_ratio;
    const float val = 1.f / (p.mf_bp2d * fpf);
    const float bs1 = float(!wl.prev_specular);
    const float w_light = val * (wl.d_shared * bs1 * p.n_bpt + fpdf * wl.d_pde / wl.rev_inv);
    const float w_local = 1.f + val * (p.mf_bb1d * sin_theta * fpf * rpf + p.mf_pb2d * rpf + p.mf_pp3d);
    const float bt1 = float(!wc.prev_specular);
    const float w_cam = val * (wc.d_shared * bt1 * p.n_bpt + rpdf * wc.d_pde / wc.rev_inv);
    if (g_mis) {
        std::vector<LightVertex> lp;
        if (light_prefix(c, lb.path_idx, lb.path_length - 1, lp)) {
            lp.push_back(beam_point(c, lb, idist, rsf, rsr));
            mis_record(c, MT_BP2D, 1.f / (w_light + w_local + w_cam), c.light_origin[lb.path_idx],
                       c.light_pick[lb.path_idx], lp.data(), int(lp.size()), g_mis->cam.data(), int(g_mis->cam.size()),
                       true, 1.f, 1.f, (lb.path_length == 1 && lb.infinite) ? 1.f : idist, int(lb.path_idx));
        }
    }
    acc += (1.f / (w_light + w_local + w_cam)) * unweighted;
}

// kernel/pde/functs_pb2d.cuh:226-342 and functs_bb1d.cuh:273-397 (segment loops)
template <bool BB1D>
static V eval_segments(const Ctx& c, V ro, V rd, const std::vector<VolSeg>* segs, const std::vector<VolLite>* lite,
                       unsigned flags, uint32_t cam_len, const StateWeights& cw)
{
    const Params& p = c.p;
    const Scene& sc = *c.sc;
    V result = v3(0.f), att = v3(1.f);
    float rsf = 1.f, rsr = 1.f;
    const size_t n = segs ? segs->size() : lite->size();
    for (size_t i = 0; i < n; i++) {
        const float dmin = segs ? (*segs)[i].dmin : (*lite)[i].dmin;
        const float dmax = segs ? (*segs)[i].dmax : (*lite)[i].dmax;
        const int mid = segs ? (*segs)[i].med : (*lite)[i].med;
        if (mid < 0) continue;
        const Medium& m = sc.media[size_t(mid)];
        V sr = v3(0.f);
        if (m.has_scattering) {
            SegQuery q{ro, rd, dmin, dmax, &m, mid, rsf, rsr, END_IN_MEDIUM | (i == 0 ? flags : 0u), cam_len, &cw};
            if (BB1D) {
                c.bvh_beam.query_segment(ro, rd, dmin, dmax, [&](uint32_t bi, float&) {
                    const LightBeam& lb = c.beams[bi];
                    if (lb.path_idx >= p.bb1d_used) return; // BB1D mask, functs_light.cuh:48-51
                    // acceptance test of the intersection program (program_bb1d.cu:31-48) is subsumed
                    eval_bb1d(c, sr, q, lb);
                });
            } else {
                c.bvh_med.query_segment(ro, rd, dmin, dmax, [&](uint32_t k, float&) {
                    const LightVertex& lv = c.verts[c.med_idx[k]];
                    if (!point_beam(ro, rd, dmax, lv.hit, p.r2_pb2d)) return; // program_pb2d.cu:31-46
                    eval_pb2d(c, sr, q, lv);
                });
            }
        }
        result += att * sr;
        if (segs) {
            if (p.qbt == UPBP_BEAM_SHORT) att *= (*segs)[i].att / (*segs)[i].pdf_for;
            else att *= (*segs)[i].att;
            if (!positive(att)) return result;
            rsf *= (*segs)[i].pdf_for;
            rsr *= (*segs)[i].pdf_rev;
        } else {
            att *= eval_attenuation(m, dmax - dmin);
            if (!positive(att)) return result;
            float srr;
            const float sff = ray_sample_pdf(m, dmin, dmax, i == 0 ? flags : 0u, &srr);
            rsf *= sff;
            rsr *= srr;
        }
    }
    return result;
}

// kernel/functs_camera.cuh:41-103
static int generate_camera_sample(const Ctx& c, uint32_t pix, SubpathState& s, float& sx, float& sy, Stream& rng)
{
    const Scene& sc = *c.sc;
    const Camera& cam = sc.cam;
    const int rx = int(cam.res_x);
    const uint32_t x = pix % uint32_t(rx), y = pix / uint32_t(rx);
    const float j0 = rng.next(), j1 = rng.next();
    sx = float(x) + j0;
    sy = float(y) + j1;
    const V wr = transform_point(cam.raster_to_world, {sx, sy, 0.f});
    const V dir = normalize(wr - cam.origin);
    const float cos_at = dot(cam.direction, dir);
    const float ipd = cam.image_plane_dist / cos_at;
    const float cam_pdf_w = (ipd * ipd) / cos_at;
    s.origin = cam.origin;
    s.dir = dir;
    s.throughput = v3(1.f);
    s.path_length = 1;
    s.specular_path = true;
    s.infinite = false;
    s.sw = StateWeights();
    s.sw.w.d_shared = float(c.p.n_light) / cam_pdf_w;
    s.sw.w.d_bpt = 0.f;
    s.sw.w.d_pde = 0.f;
    return init_stack(sc, s.stack, cam.mat_id, cam.med_id);
}

// kernel/trace_camera.cu:49-390
static V trace_camera_path(const Ctx& c, uint32_t pix, float& sx, float& sy)
{
    const Scene& sc = *c.sc;
    const Params& p = c.p;
    SubpathState s;
    Stream jrng(kDomainCamera, pix, p.iter_key, 0, P_JITTER);
    int ret = generate_camera_sample(c, pix, s, sx, sy, jrng);
    MisLogger logger;
    if (c.mis_on && c.mis_slot[pix] >= 0) {
        logger.out = &(*c.mis_cam)[size_t(c.mis_slot[pix])];
        g_mis = &logger;
    }
    V color = v3(0.f);
    bool origin_in_medium = false;
    bool only_spec_surf = (p.techniques & (UPBP_PREVIOUS | UPBP_COMPATIBLE)) != 0;
    std::vector<VolSeg> segs;
    std::vector<VolLite> lite;
    // training samples: (position, scattered direction, weight denominator, luminance collected
    // up to the scattering) of each continuing scattering that trains
    std::vector<std::array<float, 8>> pend;
    for (; ret == OK; s.path_length++) {
        const uint32_t seg = s.path_length - 1;
        Stream mrng(kDomainCamera, pix, p.iter_key, seg, P_MEDIUM);
        Isect is;
        segs.clear();
        lite.clear();
        bool hit;
        const unsigned rf = origin_in_medium ? ORIGIN_IN_MEDIUM : 0u;
        ret = intersect_(c, hit, s.origin, s.dir, is, s.stack, false, rf, segs, lite, &mrng);
        if (ret != OK) break;
        const bool use_full = is.on_surface() || p.qbt == UPBP_BEAM_SHORT || !hit;
        if (p.build_pb2d && !c.med_idx.empty()) {
            const V r = use_full ? eval_segments<false>(c, s.origin, s.dir, &segs, nullptr, rf, s.path_length, s.sw)
                                 : eval_segments<false>(c, s.origin, s.dir, nullptr, &lite, rf, s.path_length, s.sw);
            color += s.throughput * p.norm_pb2d * r;
        }
        if (p.build_bb1d && !c.beams.empty()) {
            const V r = use_full ? eval_segments<true>(c, s.origin, s.dir, &segs, nullptr, rf, s.path_length, s.sw)
                                 : eval_segments<true>(c, s.origin, s.dir, nullptr, &lite, rf, s.path_length, s.sw);
            color += s.throughput * p.norm_bb1d * r;
        }
        if (!hit) {
            if (s.path_length < p.min_len) break;
            if (sc.background_light < 0) break;
            if (sc.media[sc.global_medium].has_attenuation) break;
        }
        float rsf = 1.f, rsr = 1.f;
        if (!segs.empty()) {
            rsf = acc_for(segs);
            rsr = acc_rev(segs);
            s.throughput *= acc_att(segs) / rsf;
        }
        if (black_or_neg(s.throughput)) break;
        VertexWeights& w = s.sw.w;
        if (s.path_length > 1u) {
            const float fi = pde_factor(p, w.prev_in_medium, w.prev_specular, 1.f / rsr, w.rev_ratio, w.for_inv,
                                        w.for_ratio, s.sw.last_sin_theta);
            w.d_bpt = s.sw.d_bpt_a * fi + s.sw.d_bpt_b;
            w.d_pde = s.sw.d_pde_a * fi + s.sw.d_pde_b;
        }
        w.d_shared /= rsf;
        w.d_bpt /= rsf;
        w.d_pde /= rsf;
        if (!hit) {
            color += s.throughput * light_radiance_mis(c, sc.lights[size_t(sc.background_light)], s, v3(0.f), rsf, rsr);
            break;
        }
        BSDF b;
        bsdf_setup(b, s.dir, is, sc, false, relative_ior(sc, is, s.stack));
        if (!b.valid()) break;
        float for_ratio = 0.f, rev_ratio = 0.f;
        if (b.in_medium()) for_ratio = rev_ratio = 1.f / sc.media[size_t(b.med)].min_pos_att;
        const V hitp = s.origin + s.dir * is.dist;
        b.pos = hitp;
        origin_in_medium = is.in_medium();
        {
            const float d = b.d_factor();
            w.d_shared *= is.dist * is.dist;
            w.d_shared /= d;
            w.d_bpt /= d;
            w.d_pde /= d;
            w.for_inv = 1.f / rsf;
            w.rev_inv = 1.f / rsr;
            w.for_ratio = for_ratio;
            w.rev_ratio = rev_ratio;
        }
        if (is.light >= 0) {
            if (s.path_length < p.min_len) break;
            color += s.throughput * light_radiance_mis(c, sc.lights[size_t(is.light)], s, hitp, rsf, rsr);
            break;
        }
        if (g_mis) g_mis->cam.push_back({hitp, b, rsf, rsr});
        if (s.path_length >= p.max_len) break;
        if (p.do_bpt && !b.delta() && s.path_length + 1 >= p.min_len && !sc.lights.empty() &&
            (b.in_medium() || !only_spec_surf)) {
            V ct;
            Stream nrng(kDomainCamera, pix, p.iter_key, seg, P_NEE);
            ret = direct_illumination(c, ct, s, hitp, b, nrng);
            if (ret != OK) break;
            color += s.throughput * ct;
        }
        if (p.do_bpt && !b.delta() && (b.in_medium() || !only_spec_surf)) {
            bool behind = false;
            if (b.in_medium() && s.stack.n > 0) {
                const int m = s.stack.top().mat;
                if (m >= 0 && sc.mats[size_t(m)].real) behind = true;
            }
            Stream brng(kDomainCamera, pix, p.iter_key, seg, P_BPT_PICK);
            const uint32_t lp = uint32_t(brng.next() * float(p.n_light));
            bool err = false;
            for (uint32_t vi = c.path_begin[lp]; vi < c.path_end[lp]; vi++) {
                const LightVertex& lv = c.verts[vi];
                const uint32_t len = lv.path_length + 1 + s.path_length;
                if (len > p.max_len) break;
                const bool small = len < p.min_len;
                const bool diff_media = lv.bsdf.in_medium() && b.in_medium() && lv.bsdf.med != b.med &&
                                        (lv.behind_surf || behind);
                if (!(small || diff_media)) {
                    V ct;
                    ret = connect_vertices(c, ct, lv, b, hitp, s);
                    if (ret != OK) {
                        err = true;
                        break;
                    }
                    color += s.throughput * lv.throughput * ct;
                }
            }
            if (err) break;
        }
        if (p.build_surf && !c.surf_idx.empty() && b.on_surface() && !b.delta() && !only_spec_surf) {
            V acc = v3(0.f);
            c.bvh_surf.query_point(hitp, [&](uint32_t k) {
                const LightVertex& lv = c.verts[c.surf_idx[k]];
                if (!point_point(hitp, lv.hit, p.r2_surf)) return;
                eval_pp(c, acc, s, b, lv);
            });
            color += s.throughput * p.norm_surf * acc;
        }
        if (p.build_pp3d && !c.med_idx.empty() && b.in_medium() && !b.delta()) {
            V acc = v3(0.f);
            c.bvh_med.query_point(hitp, [&](uint32_t k) {
                const LightVertex& lv = c.verts[c.med_idx[k]];
                if (!point_point(hitp, lv.hit, p.r2_pp3d)) return;
                eval_pp(c, acc, s, b, lv);
            });
            color += s.throughput * p.norm_pp3d * acc;
        }
        if (p.build_bp2d && !c.beams.empty() && b.in_medium()) {
            V acc = v3(0.f);
            c.bvh_beam.query_point(hitp, [&](uint32_t k) {
                const LightBeam& lb = c.beams[k];
                if (!point_beam(lb.origin, lb.dir, lb.length, hitp, p.r2_bp2d)) return; // program_bp2d.cu:31-45
                eval_bp2d(c, acc, s, b, hitp, lb);
            });
            color += s.throughput * p.norm_bp2d * acc;
        }
        bool cont;
        Stream srng(kDomainCamera, pix, p.iter_key, seg, P_SCATTER);
        float den = 0.f;
        ret = sample_scattering(c, cont, b, hitp, is, s, srng, c.train ? &den : nullptr);
        if (ret != OK) break;
        if (!cont) break;
        if (b.on_surface()) {
            if (!s.sw.w.prev_specular && (only_spec_surf || (p.techniques & UPBP_SPECULAR_ONLY))) break;
        } else {
            if (p.techniques & UPBP_SPECULAR_ONLY) break;
            if (only_spec_surf) {
                if (p.techniques & UPBP_COMPATIBLE) only_spec_surf = false;
                else break;
            }
        }
        if (c.train && den > 0.f)
            pend.push_back({s.origin.x, s.origin.y, s.origin.z, s.dir.x, s.dir.y, s.dir.z, den, luminance(color)});
    }
    g_mis = nullptr;
    if (ret != OK) const_cast<Ctx&>(c).errors++;
    // a sample's weight: the luminance the path collects after its scattering over the
    // denominator, kept when positive (the GPU's k_guide_finalise)
    if (c.train)
        for (const auto& q : pend) {
            const float w = (luminance(color) - q[7]) / q[6];
            if (w > 0.f && std::isfinite(w)) (*c.train)[pix].push_back({q[0], q[1], q[2], q[3], q[4], q[5], w});
        }
    return color;
}

// renderer_init.cpp:675-929 (flags, counts) and renderer.cpp:56-119 (radius schedule)
static void setup_params(const Scene& sc, const upbp_config& cfg, int iteration, Params& p)
{
    auto rel = [&](float r) { return r < 0.f ? -r * sc.radius : r; };
    const float bb1d_r0 = rel(cfg.bb1d_radius_initial), bp2d_r0 = rel(cfg.bp2d_radius_initial),
                pb2d_r0 = rel(cfg.pb2d_radius_initial), pp3d_r0 = rel(cfg.pp3d_radius_initial),
                surf_r0 = rel(cfg.surf_radius_initial);
    const uint32_t f = cfg.algorithm_flags;
    p.trace_light = p.trace_camera = f != 0;
    p.do_bpt = f & UPBP_BPT;
    p.do_bb1d = f & UPBP_BB1D;
    p.do_bp2d = f & UPBP_BP2D;
    p.do_pb2d = f & UPBP_PB2D;
    p.do_pp3d = f & UPBP_PP3D;
    p.do_surf = f & UPBP_SURF;
    if (f & UPBP_SPECULAR_ONLY) {
        p.trace_light = false;
        p.trace_camera = true;
        p.do_bpt = p.do_bb1d = p.do_bp2d = p.do_pb2d = p.do_pp3d = p.do_surf = false;
    }
    p.connect_from_surf = (f & (UPBP_PREVIOUS | UPBP_COMPATIBLE)) == 0;
    p.n_pixels = cfg.res_x * cfg.res_y;
    p.n_light = uint32_t(cfg.path_count_per_iter);
    float bb1d_used = cfg.bb1d_used_light_subpath_count;
    if (bb1d_used < 0.f) bb1d_used = std::floor(-bb1d_used * float(p.n_light));
    const bool merging = cfg.max_path_length > 1;
    p.build_bb1d = p.trace_light && merging && p.do_bb1d;
    p.build_bp2d = p.trace_light && merging && p.do_bp2d;
    p.build_pb2d = p.trace_light && merging && p.do_pb2d;
    p.build_pp3d = p.trace_light && merging && p.do_pp3d;
    p.build_surf = p.trace_light && merging && p.do_surf;
    p.min_len = cfg.min_path_length;
    p.max_len = cfg.max_path_length;
    p.pbt = cfg.photon_beam_type;
    p.qbt = cfg.query_beam_type;
    p.n_bpt = float(p.do_bpt);
    p.n_bb1d = float(p.do_bb1d) * bb1d_used;
    p.n_bp2d = float(p.do_bp2d) * float(p.n_light);
    p.n_pb2d = float(p.do_pb2d) * float(p.n_light);
    p.n_pp3d = float(p.do_pp3d) * float(p.n_light);
    p.n_surf = float(p.do_surf) * float(p.n_light);
    p.ignore_spec = cfg.ignore_fully_spec_paths != 0;
    p.use_shading_normal = cfg.use_shading_normal != 0;
    p.techniques = f;
    p.bb1d_used = uint32_t(bb1d_used);
    p.iteration = uint32_t(iteration);
    p.iter_key = uint64_t(iteration) + cfg.iteration_offset;

    const float eff = 1.f + float(iteration);
    float r_bb1d = bb1d_r0 * std::pow(1.f + float(iteration) * bb1d_used / float(p.n_light), cfg.bb1d_radius_alpha - 1.f);
    r_bb1d = std::fmax(r_bb1d, 1e-7f);
    float r_bp2d = std::fmax(bp2d_r0 * std::pow(eff, (cfg.bp2d_radius_alpha - 1.f) * .5f), 1e-7f);
    float r_pb2d = std::fmax(pb2d_r0 * std::pow(eff, (cfg.pb2d_radius_alpha - 1.f) * .5f), 1e-7f);
    float r_pp3d = std::fmax(pp3d_r0 * std::pow(eff, (cfg.pp3d_radius_alpha - 1.f) * (1.f / 3.f)), 1e-7f);
    float r_surf = std::fmax(surf_r0 * std::pow(eff, (cfg.surf_radius_alpha - 1.f) * .5f), 1e-7f);
    p.mf_bb1d = .5f * r_bb1d * p.n_bb1d;
    p.mf_bp2d = (kPi * (r_bp2d * r_bp2d)) * p.n_bp2d;
    p.mf_pb2d = (kPi * (r_pb2d * r_pb2d)) * p.n_pb2d;
    p.mf_pp3d = (4.f / 3.f) * (kPi * (r_pp3d * r_pp3d * r_pp3d)) * p.n_pp3d;
    p.mf_surf = (kPi * (r_surf * r_surf)) * p.n_surf;
    p.norm_bb1d = 1.f / bb1d_used;
    p.norm_bp2d = 1.f / float(p.n_light);
    p.norm_pb2d = 1.f / float(p.n_light);
    p.norm_pp3d = 1.f / p.mf_pp3d;
    p.norm_surf = 1.f / p.mf_surf;
    p.r_bb1d = r_bb1d;
    p.r_bp2d = r_bp2d;
    p.r_pb2d = r_pb2d;
    p.r_pp3d = r_pp3d;
    p.r_surf = r_surf;
    p.rad_point_med = 0.f;
    if (p.do_pp3d) p.rad_point_med = std::fmax(p.rad_point_med, r_pp3d);
    if (p.do_pb2d) p.rad_point_med = std::fmax(p.rad_point_med, r_pb2d);
    p.rad_beam = 0.f;
    if (p.do_bb1d) p.rad_beam = std::fmax(p.rad_beam, r_bb1d);
    if (p.do_bp2d) p.rad_beam = std::fmax(p.rad_beam, r_bp2d);
    p.r2_bb1d = r_bb1d * r_bb1d;
    p.r2_bp2d = r_bp2d * r_bp2d;
    p.r2_pb2d = r_pb2d * r_pb2d;
    p.r2_pp3d = r_pp3d * r_pp3d;
    p.r2_surf = r_surf * r_surf;
}

template <typename F>
static void parallel_for(uint32_t n, int threads, F&& fn)
{
    if (threads <= 1 || n < 64) {
        for (uint32_t i = 0; i < n; i++) fn(0, i);
        return;
    }
    std::atomic<uint32_t> next{0};
    std::vector<std::thread> ts;
    for (int t = 0; t < threads; t++)
        ts.emplace_back([&, t] {
            for (;;) {
                const uint32_t b = next.fetch_add(64);
                if (b >= n) break;
                const uint32_t e = std::min(n, b + 64);
                for (uint32_t i = b; i < e; i++) fn(t, i);
            }
        });
    for (auto& t : ts) t.join();
}

static Box point_box(V p, float r)
{
    Box b;
    b.lo = p - v3(r);
    b.hi = p + v3(r);
    return b;
}

// One iteration into fb (W*H*3 floats, zeroed by the caller).
static void render_iteration(Ctx& c, const upbp_config& cfg, int iteration, int threads, std::vector<float>& fb,
                             uint32_t n_camera, uint32_t n_light_limit, oracle_stats* st, double* times = nullptr)
{
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    const Scene& sc = *c.sc;
    setup_params(sc, cfg, iteration, c.p);
    Params& p = c.p;
    const size_t npx = size_t(cfg.res_x) * cfg.res_y;
    c.verts.clear();
    c.beams.clear();
    c.surf_idx.clear();
    c.med_idx.clear();
    c.path_begin.assign(p.n_light, 0);
    c.path_end.assign(p.n_light, 0);
    if (c.mis_on) {
        c.light_origin.assign(p.n_light, v3(0.f));
        c.light_pick.assign(p.n_light, 0);
    }
    if (p.trace_light) {
        const uint32_t nl = std::min(p.n_light, n_light_limit);
        std::vector<LightOut> outs(static_cast<size_t>(nl));
        std::vector<std::vector<float>> tfb(size_t(std::max(threads, 1)), std::vector<float>(3 * npx, 0.f));
        parallel_for(nl, threads, [&](int t, uint32_t i) {
            LightOut& o = outs[i];
            o.fb.swap(tfb[size_t(t)]);
            trace_light_path(c, i, o);
            o.fb.swap(tfb[size_t(t)]);
        });
        for (auto& f : tfb)
            for (size_t k = 0; k < f.size(); k++) fb[k] += f[k];
        for (uint32_t i = 0; i < nl; i++) {
            c.path_begin[i] = uint32_t(c.verts.size());
            for (auto& v : outs[i].verts) c.verts.push_back(v);
            c.path_end[i] = uint32_t(c.verts.size());
            for (auto& b : outs[i].beams) c.beams.push_back(b);
        }
        for (uint32_t i = nl; i < p.n_light; i++) c.path_begin[i] = c.path_end[i] = uint32_t(c.verts.size());
    }
    const auto t1 = clk::now();
    std::vector<Box> sb, mb, bb;
    for (uint32_t i = 0; i < c.verts.size(); i++) {
        if (c.verts[i].in_medium) {
            c.med_idx.push_back(i);
            mb.push_back(point_box(c.verts[i].hit, p.rad_point_med));
        } else {
            c.surf_idx.push_back(i);
            sb.push_back(point_box(c.verts[i].hit, p.r_surf));
        }
    }
    for (const LightBeam& b : c.beams) {
        Box x;
        const V e = b.origin + b.dir * b.length;
        x.grow(b.origin - v3(p.rad_beam));
        x.grow(b.origin + v3(p.rad_beam));
        x.grow(e - v3(p.rad_beam));
        x.grow(e + v3(p.rad_beam));
        bb.push_back(x);
    }
    if (!p.build_surf) c.surf_idx.clear(), sb.clear();
    if (!p.build_pp3d && !p.build_pb2d) c.med_idx.clear(), mb.clear();
    if (!p.build_bb1d && !p.build_bp2d) c.beams.clear(), bb.clear();
    c.bvh_surf.build(sb);
    c.bvh_med.build(mb);
    c.bvh_beam.build(bb);
    const auto t2 = clk::now();
    if (st) {
        st->light_vertices += c.verts.size();
        st->surface_vertices += c.surf_idx.size();
        st->medium_vertices += c.med_idx.size();
        st->beams += c.beams.size();
    }
    if (p.trace_camera) {
        const uint32_t nc = std::min<uint32_t>(p.n_pixels, n_camera);
        std::vector<V> col(nc);
        std::vector<float> sxs(nc), sys(nc);
        // a bounded sample takes pixels spread evenly over the image
        parallel_for(nc, threads, [&](int, uint32_t i) {
            const uint32_t pix = nc == p.n_pixels ? i : uint32_t(uint64_t(i) * p.n_pixels / nc);
            col[i] = trace_camera_path(c, pix, sxs[i], sys[i]);
        });
        for (uint32_t i = 0; i < nc; i++) splat(fb, sc.cam, sxs[i], sys[i], col[i]);
    }
    if (st) st->path_errors += c.errors.exchange(0);
    if (times) {
        times[0] = std::chrono::duration<double>(t1 - t0).count();
        times[1] = std::chrono::duration<double>(t2 - t1).count();
        times[2] = std::chrono::duration<double>(clk::now() - t2).count();
    }
}

} // namespace orc

using namespace orc;

extern "C" int oracle_render(const upbp_scene* scene, const upbp_config* cfg, int first_iteration, int n_iterations,
                             int n_threads, double* out_rgb, float* fb_first, oracle_stats* stats)
{
    if (!scene || !cfg || !out_rgb || n_iterations < 1) return UPBP_E_ARG;
    if (scene->light_count == 0) return UPBP_E_ARG;
    const auto t0 = std::chrono::steady_clock::now();
    Scene sc;
    prepare_scene(scene, cfg->res_x, cfg->res_y, sc);
    Ctx c;
    c.sc = &sc;
    {
        std::vector<Box> rb, ib;
        for (auto& t : sc.real) rb.push_back(tri_box(t));
        for (auto& t : sc.imag) ib.push_back(tri_box(t));
        c.bvh_real.build(rb);
        c.bvh_imag.build(ib);
    }
    const int threads = n_threads > 0 ? n_threads : int(std::max(1u, std::thread::hardware_concurrency()));
    const size_t npx = size_t(cfg->res_x) * cfg->res_y;
    std::vector<double> acc(3 * npx, 0.0);
    if (stats) memset(stats, 0, sizeof(*stats));
    std::vector<float> fb(3 * npx);
    for (int it = 0; it < n_iterations; it++) {
        std::fill(fb.begin(), fb.end(), 0.f);
        render_iteration(c, *cfg, first_iteration + it, threads, fb, UINT32_MAX, UINT32_MAX, stats);
        if (it == 0 && fb_first) memcpy(fb_first, fb.data(), fb.size() * sizeof(float));
        for (size_t k = 0; k < acc.size(); k++) acc[k] += double(fb[k]);
    }
    for (size_t k = 0; k < acc.size(); k++) out_rgb[k] = acc[k] / double(n_iterations);
    if (stats) stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return UPBP_OK;
}

// Renders one iteration with the MIS log on for the given camera paths (pixels) and the
// first n_lt_paths light paths; for every logged contribution writes (technique, s, k,
// constant-time weight, FP64 balance-heuristic weight, the same with the reference's
// next-event weight convention, camera path (pixel; light tracing: light path), light path
// (next-event estimation and light hits: the light)) to out (8 doubles per record).
extern "C" int oracle_mis_check(const upbp_scene* scene, const upbp_config* cfg, int iteration, const uint32_t* pixels,
                                uint32_t n_pixels, uint32_t n_lt_paths, int n_threads, double* out, uint32_t max_records,
                                uint32_t* n_records)
{
    if (!scene || !cfg || (n_pixels && !pixels) || !n_records) return UPBP_E_ARG;
    Scene sc;
    prepare_scene(scene, cfg->res_x, cfg->res_y, sc);
    Ctx c;
    c.sc = &sc;
    {
        std::vector<Box> rb, ib;
        for (auto& t : sc.real) rb.push_back(tri_box(t));
        for (auto& t : sc.imag) ib.push_back(tri_box(t));
        c.bvh_real.build(rb);
        c.bvh_imag.build(ib);
    }
    const size_t npx = size_t(cfg->res_x) * cfg->res_y;
    std::vector<std::vector<MisRec>> cam_recs(n_pixels), lt_recs(n_lt_paths);
    c.mis_on = true;
    c.mis_slot.assign(npx, -1);
    for (uint32_t i = 0; i < n_pixels; i++)
        if (pixels[i] < npx) c.mis_slot[pixels[i]] = int32_t(i);
    c.mis_cam = &cam_recs;
    c.mis_lt = &lt_recs;
    c.mis_lt_paths = n_lt_paths;
    const int threads = n_threads > 0 ? n_threads : int(std::max(1u, std::thread::hardware_concurrency()));
    std::vector<float> fb(3 * npx, 0.f);
    render_iteration(c, *cfg, iteration, threads, fb, UINT32_MAX, UINT32_MAX, nullptr);
    const Params& p = c.p;
    MisParams mp;
    mp.n_bpt = p.n_bpt;
    mp.n_light = double(p.n_light);
    mp.mf_surf = p.mf_surf;
    mp.mf_pp3d = p.mf_pp3d;
    mp.mf_pb2d = p.mf_pb2d;
    mp.mf_bp2d = p.mf_bp2d;
    mp.mf_bb1d = p.mf_bb1d;
    mp.long_photon_beams = p.pbt == UPBP_BEAM_LONG;
    mp.long_query_beams = p.qbt == UPBP_BEAM_LONG;
    mp.connect_from_surf = p.connect_from_surf;
    mp.reference_nee_weight = false;
    mp.reference_beam_distance = false;
    MisParams mq = mp;
    mq.reference_nee_weight = true;
    mq.reference_beam_distance = true;
    uint32_t n = 0;
    for (auto* set : {&cam_recs, &lt_recs})
        for (auto& v : *set)
            for (const MisRec& r : v) {
                if (out && n < max_records) {
                    double* o = out + 8 * size_t(n);
                    o[0] = r.tech;
                    o[1] = r.s;
                    o[2] = r.k;
                    o[3] = r.weight;
                    o[4] = mis_explicit_weight(r, mp);
                    o[5] = mis_explicit_weight(r, mq);
                    o[6] = set == &cam_recs ? double(pixels[&v - cam_recs.data()]) : double(&v - lt_recs.data());
                    o[7] = r.lpath;
                }
                n++;
            }
    *n_records = n;
    return UPBP_OK;
}

// Renders one iteration with the learned proposal's training samples on (scatterings in media
// with train_media, at guidable surfaces with train_surfaces): every kept sample as (pixel,
// world position, scattered direction, weight), 8 floats each, pixel by pixel in the order of
// the path's scatterings; *n_samples receives the total.
extern "C" int oracle_training_samples(const upbp_scene* scene, const upbp_config* cfg, int iteration,
                                       int32_t train_media, int32_t train_surfaces, int n_threads, float* out,
                                       uint64_t max_samples, uint64_t* n_samples)
{
    if (!scene || !cfg || !n_samples) return UPBP_E_ARG;
    Scene sc;
    prepare_scene(scene, cfg->res_x, cfg->res_y, sc);
    Ctx c;
    c.sc = &sc;
    {
        std::vector<Box> rb, ib;
        for (auto& t : sc.real) rb.push_back(tri_box(t));
        for (auto& t : sc.imag) ib.push_back(tri_box(t));
        c.bvh_real.build(rb);
        c.bvh_imag.build(ib);
    }
    const size_t npx = size_t(cfg->res_x) * cfg->res_y;
    std::vector<std::vector<std::array<float, 7>>> recs(npx);
    c.train = &recs;
    c.train_media = train_media != 0;
    c.train_surfaces = train_surfaces != 0;
    const int threads = n_threads > 0 ? n_threads : int(std::max(1u, std::thread::hardware_concurrency()));
    std::vector<float> fb(3 * npx, 0.f);
    render_iteration(c, *cfg, iteration, threads, fb, UINT32_MAX, UINT32_MAX, nullptr);
    uint64_t n = 0;
    for (size_t px = 0; px < npx; px++)
        for (const auto& r : recs[px]) {
            if (out && n < max_samples) {
                float* o = out + 8 * n;
                o[0] = float(px);
                for (int k = 0; k < 7; k++) o[1 + k] = r[size_t(k)];
            }
            n++;
        }
    *n_samples = n;
    return UPBP_OK;
}

extern "C" int oracle_render_sample(const upbp_scene* scene, const upbp_config* cfg, int iteration,
                                    uint32_t n_camera_paths, uint32_t n_light_paths, int n_threads, double* times,
                                    oracle_stats* stats)
{
    if (!scene || !cfg || !times) return UPBP_E_ARG;
    Scene sc;
    prepare_scene(scene, cfg->res_x, cfg->res_y, sc);
    Ctx c;
    c.sc = &sc;
    std::vector<Box> rb, ib;
    for (auto& t : sc.real) rb.push_back(tri_box(t));
    for (auto& t : sc.imag) ib.push_back(tri_box(t));
    c.bvh_real.build(rb);
    c.bvh_imag.build(ib);
    const int threads = n_threads > 0 ? n_threads : int(std::max(1u, std::thread::hardware_concurrency()));
    std::vector<float> fb(3 * size_t(cfg->res_x) * cfg->res_y, 0.f);
    // times[0]: light pass, times[1]: photon-map build, times[2]: camera pass (wall seconds)
    render_iteration(c, *cfg, iteration, threads, fb, n_camera_paths, n_light_paths, stats, times);
    return UPBP_OK;
}

extern "C" int oracle_range_query(int kind, const float* prims, uint32_t n_prims, const float* queries,
                                  uint32_t n_queries, float radius, uint32_t* offsets, uint32_t* indices,
                                  uint64_t max_indices, uint64_t* n_indices)
{
    if (kind < 0 || kind > 3 || !offsets) return UPBP_E_ARG;
    const float r2 = radius * radius;
    uint64_t n = 0;
    for (uint32_t q = 0; q < n_queries; q++) {
        offsets[q] = uint32_t(n);
        for (uint32_t i = 0; i < n_prims; i++) {
            bool acc = false;
            if (kind == 0) {
                acc = point_point(fromp(queries + 3 * q), fromp(prims + 3 * i), r2);
            } else if (kind == 1) {
                const float* Q = queries + 8 * size_t(q);
                float d, h;
                acc = point_beam(fromp(Q), fromp(Q + 3), Q[7], fromp(prims + 3 * i), r2) &&
                      bre(d, h, fromp(Q), fromp(Q + 3), Q[6], Q[7], fromp(prims + 3 * i), r2);
            } else if (kind == 2) {
                const float* B = prims + 7 * size_t(i);
                acc = point_beam(fromp(B), fromp(B + 3), B[6], fromp(queries + 3 * q), r2);
            } else {
                const float* Q = queries + 8 * size_t(q);
                const float* B = prims + 7 * size_t(i);
                float d, s, t1, t2;
                acc = beam_beam(fromp(Q), fromp(Q + 3), Q[6], Q[7], fromp(B), fromp(B + 3), 0.f, B[6], r2, d, s, t1, t2);
            }
            if (acc) {
                if (n < max_indices && indices) indices[n] = i;
                n++;
            }
        }
    }
    offsets[n_queries] = uint32_t(n);
    if (n_indices) *n_indices = n;
    return n <= max_indices ? UPBP_OK : UPBP_E_CAPACITY;
}

extern "C" int oracle_sphere_hits(const float* spheres, uint32_t n_sph, const float* rays, uint32_t n_rays, float* out)
{
    for (uint32_t r = 0; r < n_rays; r++) {
        const float* R = rays + 8 * size_t(r);
        for (uint32_t k = 0; k < n_sph; k++) {
            const Sph s{fromp(spheres + 4 * size_t(k)), spheres[4 * size_t(k) + 3], 0, -1};
            float th[2];
            V nh[2];
            const int n = sphere_hit(s, fromp(R), fromp(R + 3), R[6], R[7], th, nh);
            float* o = out + 3 * (size_t(r) * n_sph + k);
            o[0] = float(n);
            o[1] = n > 0 ? th[0] : INFINITY;
            o[2] = n > 1 ? th[1] : INFINITY;
        }
    }
    return UPBP_OK;
}

extern "C" int oracle_guide_samples(const float* lobes, const float* arrive, float g, int32_t surface, const float* u,
                                    uint32_t n, float* out)
{
    Guide gd;
    const float lo[3] = {0.f, 0.f, 0.f}, hi[3] = {1.f, 1.f, 1.f};
    init_guide(gd, lobes, 1, lo, hi, 1.f);
    const GuideLobe* r = gd.lobes.data();
    const V ar{arrive[0], arrive[1], arrive[2]};
    const PhaseFit h = surface ? cosine_fit(ar) : phase_fit(ar, g);
    for (uint32_t i = 0; i < n; i++) {
        const V d = guide_prod_sample(r, h, u[3 * i], u[3 * i + 1], u[3 * i + 2]);
        out[4 * i] = d.x;
        out[4 * i + 1] = d.y;
        out[4 * i + 2] = d.z;
        out[4 * i + 3] = guide_prod_pdf(r, h, d);
    }
    return UPBP_OK;
}

extern "C" int oracle_set_guide_lobes(const float* lobes, int32_t res, const float* lo, const float* hi, float alpha)
{
    if (!(alpha >= 0.f && alpha <= 1.f)) return UPBP_E_ARG;
    if (alpha > 0.f && (!lobes || !lo || !hi || res < 1 || res > 256)) return UPBP_E_ARG;
    init_guide(g_oracle_guide, lobes, res, lo, hi, alpha);
    return UPBP_OK;
}

extern "C" int oracle_env_samples(const upbp_env_map* env, const float* u, uint32_t n, float* out)
{
    if (!env || !env->rgb || env->height == 0 || env->width < 2 * env->height) return UPBP_E_ARG;
    EnvMap m;
    init_env_map(*env, m);
    for (uint32_t i = 0; i < n; i++) {
        float pdf, lpdf;
        V rad;
        const V d = env_sample(m, u[2 * i], u[2 * i + 1], pdf, rad);
        const V lrad = env_lookup(m, d, &lpdf);
        const float vals[11] = {d.x, d.y, d.z, pdf, rad.x, rad.y, rad.z, lpdf, lrad.x, lrad.y, lrad.z};
        memcpy(out + 11 * size_t(i), vals, sizeof(vals));
    }
    return UPBP_OK;
}

extern "C" int oracle_closest_hit(const float* verts, uint32_t n_tris, const float* rays, uint32_t n_rays, float* hit_t,
                                  int32_t* hit_prim)
{
    std::vector<Tri> tris(n_tris);
    for (uint32_t i = 0; i < n_tris; i++) {
        tris[i].p0 = fromp(verts + 9 * size_t(i));
        tris[i].p1 = fromp(verts + 9 * size_t(i) + 3);
        tris[i].p2 = fromp(verts + 9 * size_t(i) + 6);
    }
    for (uint32_t r = 0; r < n_rays; r++) {
        const float* R = rays + 8 * size_t(r);
        float best = INFINITY;
        int bi = -1;
        for (uint32_t i = 0; i < n_tris; i++) {
            float t, u, v;
            if (tri_hit(tris[i], fromp(R), fromp(R + 3), R[6], R[7], t, u, v) && (t < best || (t == best && int(i) < bi))) {
                best = t;
                bi = int(i);
            }
        }
        hit_t[r] = best;
        hit_prim[r] = bi;
    }
    return UPBP_OK;
}

extern "C" int oracle_camera_ray(const upbp_scene* scene, const upbp_config* cfg, const float* raster_xy, uint32_t n,
                                 float* dirs)
{
    Scene sc;
    prepare_scene(scene, cfg->res_x, cfg->res_y, sc);
    for (uint32_t i = 0; i < n; i++) {
        const V wr = transform_point(sc.cam.raster_to_world, {raster_xy[2 * i], raster_xy[2 * i + 1], 0.f});
        const V d = normalize(wr - sc.cam.origin);
        dirs[3 * i] = d.x;
        dirs[3 * i + 1] = d.y;
        dirs[3 * i + 2] = d.z;
    }
    return UPBP_OK;
}

extern "C" int oracle_world_to_raster(const upbp_scene* scene, const upbp_config* cfg, const float* pts, uint32_t n,
                                      float* raster_xy)
{
    Scene sc;
    prepare_scene(scene, cfg->res_x, cfg->res_y, sc);
    for (uint32_t i = 0; i < n; i++) {
        const V r = transform_point(sc.cam.world_to_raster, fromp(pts + 3 * i));
        raster_xy[2 * i] = r.x;
        raster_xy[2 * i + 1] = r.y;
    }
    return UPBP_OK;
}

extern "C" float oracle_hg_pdf(float g, float cos_theta)
{
    // directions pointing away from the scattering point: fix = -z, gen at angle theta to +z
    const V fix{0, 0, -1};
    const float st = std::sqrt(std::fmax(0.f, 1.f - cos_theta * cos_theta));
    const V gen{st, 0.f, cos_theta};
    return hg_pdf(fix, gen, g, nullptr);
}

extern "C" void oracle_hg_sample(float g, float u0, float u1, float* cos_theta, float* pdf)
{
    const V fix{0, 0, -1};
    Frame f;
    f.set_from_z(-fix);
    float p;
    const V gen = hg_sample(fix, g, u0, u1, f, p, nullptr);
    *cos_theta = gen.z;
    *pdf = p;
}

extern "C" float oracle_fresnel(float cos_inc, float ior) { return fresnel_dielectric(cos_inc, ior); }

// One surface BSDF (kernel/bsdf.cuh:179-660) of material `mat` at a surface with normal +z, set
// up for a ray travelling along ray_dir (3 floats) with relative index of refraction ior, on a
// camera path (from_light 0) or a light path (1). The learned proposal is off.
static bool surface_bsdf(const upbp_material* mat, const float* ray_dir, float ior, int from_light, Scene& sc,
                         BSDF& b)
{
    Material m;
    m.diffuse = {mat->diffuse[0], mat->diffuse[1], mat->diffuse[2]};
    m.phong = {mat->phong[0], mat->phong[1], mat->phong[2]};
    m.mirror = {mat->mirror[0], mat->mirror[1], mat->mirror[2]};
    m.phong_exp = mat->phong_exponent;
    m.ior = mat->ior;
    m.priority = mat->priority;
    m.real = mat->real != 0;
    m.med_idx = mat->med_idx;
    sc.mats.assign(1, m);
    Isect is;
    is.dist = 1.f;
    is.mat = 0;
    is.normal = is.shading = {0.f, 0.f, 1.f};
    is.enter = ray_dir[2] < 0.f;
    bsdf_setup(b, V{ray_dir[0], ray_dir[1], ray_dir[2]}, is, sc, from_light != 0, ior);
    return b.valid();
}

// Evaluates the BSDF toward each of n directions (3 floats each); out: 7 floats per direction
// (f.xyz, cosine at the generated side, forward pdf, reverse pdf, continuation probability).
extern "C" int oracle_bsdf_eval(const upbp_material* mat, const float* ray_dir, float ior, int from_light,
                                const float* dirs, uint32_t n, float* out)
{
    if (!mat || !ray_dir || (n && (!dirs || !out))) return UPBP_E_ARG;
    Scene sc;
    BSDF b;
    if (!surface_bsdf(mat, ray_dir, ior, from_light, sc, b)) return UPBP_E_ARG;
    for (uint32_t i = 0; i < n; i++) {
        float cg = 0.f, fp = 0.f, rp = 0.f, st = 0.f;
        const V f = bsdf_evaluate(b, V{dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]}, sc, cg, &fp, &rp, &st);
        float* o = out + 7 * size_t(i);
        o[0] = f.x;
        o[1] = f.y;
        o[2] = f.z;
        o[3] = cg;
        o[4] = fp;
        o[5] = rp;
        o[6] = b.cont_prob;
    }
    return UPBP_OK;
}

// Samples the BSDF with n variate triples u (3 floats each); out: 9 floats per sample (direction,
// pdf, cosine, f.xyz, event bits of BSDF::Event).
extern "C" int oracle_bsdf_sample(const upbp_material* mat, const float* ray_dir, float ior, int from_light,
                                  const float* u, uint32_t n, float* out)
{
    if (!mat || !ray_dir || (n && (!u || !out))) return UPBP_E_ARG;
    Scene sc;
    BSDF b;
    if (!surface_bsdf(mat, ray_dir, ior, from_light, sc, b)) return UPBP_E_ARG;
    for (uint32_t i = 0; i < n; i++) {
        V d{0, 0, 0};
        float pdf = 0.f, cg = 0.f, st = 0.f;
        unsigned ev = 0;
        const V f = bsdf_sample(b, u[3 * i], u[3 * i + 1], u[3 * i + 2], d, pdf, cg, sc, ev, st);
        float* o = out + 9 * size_t(i);
        o[0] = d.x;
        o[1] = d.y;
        o[2] = d.z;
        o[3] = pdf;
        o[4] = cg;
        o[5] = f.x;
        o[6] = f.y;
        o[7] = f.z;
        o[8] = float(ev);
    }
    return UPBP_OK;
}

extern "C" void oracle_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) { philox(ctr, key, out); }
