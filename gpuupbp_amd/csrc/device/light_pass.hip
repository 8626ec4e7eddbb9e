// Light-subpath wavefront: one generate kernel and one bounce kernel per path segment.
// Each bounce kernel consumes the compacted queue of live light paths, advances every
// path by one segment (intersect, store beams, complete the MIS quantities, store the
// vertex into the path's fixed-stride slot, connect to the camera, sample the next
// direction) and appends the survivors to the next queue with one atomic per wave.
// Semantics: nolmoonen/gpuupbp gpuupbp/kernel/trace_light.cu:42-325 and
// gpuupbp/kernel/functs_light.cuh:86-388.
#include "kernels.h"

namespace upbp {

// kernel/functs_light.cuh:86-156
__device__ __forceinline__ int generate_light_sample(const DScene& sc, const IterParams& p, Sub& s, Rng& rng)
{
    const float lc = float(sc.n_lights);
    const int id = int(rng.next() * lc);
    const float d0 = rng.next(), d1 = rng.next(), p0 = rng.next(), p1 = rng.next();
    const DLight& l = sc.lights[id];
    float epdf, dpdf, cosl;
    s.throughput = light_emit(l, sc, d0, d1, p0, p1, s.origin, s.dir, epdf, &dpdf, &cosl);
    epdf /= lc;
    dpdf /= lc;
    s.throughput /= epdf;
    s.path_length = 1;
    s.infinite = !light_finite(l);
    s.specular_path = false;
    s.sw.w.d_shared = dpdf / epdf;
    s.sw.w.d_bpt = !light_delta(l) ? (light_finite(l) ? cosl : 1.f) / epdf : 0.f;
    s.sw.w.d_pde = s.sw.w.d_bpt * p.n_bpt;
    s.sw.w.for_inv = s.sw.w.rev_inv = s.sw.w.for_ratio = s.sw.w.rev_ratio = 0.f;
    s.sw.w.prev_specular = 0;
    s.sw.w.prev_in_medium = 0;
    s.sw.last_sin = s.sw.d_bpt_a = s.sw.d_bpt_b = s.sw.d_pde_a = s.sw.d_pde_b = 0.f;
    return init_stack(sc, s.stack, l.mat_id, l.med_id);
}

// vb: the block's index within its job (blockIdx.x for a single job)
__device__ __forceinline__ void light_generate_body(const DScene& sc, const Bufs& b, uint32_t vb)
{
    const IterParams& p = *b.params;
    const uint32_t pid = vb * 256u + threadIdx.x;
    if (pid >= p.n_light) return;
    b.nvert[pid] = 0;
    b.nbeam[pid] = 0;
    Rng rng(kDomainLight, pid, p.iter_key, 0, P_EMIT);
    Sub s;
    const int ret = generate_light_sample(sc, p, s, rng);
    const bool alive = ret == S_OK && !(s.infinite && sc.media[sc.global_medium].has_attenuation);
    if (ret != S_OK) stat_add(b.stats, STAT_ERRORS, 1);
    if (alive) {
        PathState st;
        st.flags = 0;
        st.color = mk(0.f);
        st.sx = st.sy = 0.f;
        save_sub(st, s);
        b.lstate[pid] = st;
        store_stack(b.lstack, pid, s.stack);
    }
    queue_append(alive, pid, b.queue[0], &b.qcount[0]);
}

__global__ void __launch_bounds__(256) k_light_generate(DScene sc, Bufs b) { light_generate_body(sc, b, blockIdx.x); }

__global__ void __launch_bounds__(256) k_light_generate_batch(const BatchJob* jobs, const uint2* bmap)
{
    const uint2 m = bmap[blockIdx.x];
    light_generate_body(jobs[m.x].sc, jobs[m.x].b, m.y);
}

// kernel/functs_light.cuh:159-247
__device__ __forceinline__ void add_beams(const DScene& sc, const IterParams& p, const Bufs& b, uint32_t pid,
                                          const Sub& s, unsigned flags, const VolSeg* segs, int nseg,
                                          const LiteSeg* lite, int nlite, uint32_t& nbeam, uint32_t& drops)
{
    f3 t = s.throughput;
    float fpdf = 1.f, rpdf = 1.f;
    auto emit_beam = [&](float dmin, float dmax, const DMedium& m, int mid, int i) {
        if (nbeam >= p.bstride) {
            drops++;
            return;
        }
        LBeam x;
        x.med = mid;
        x.origin = s.origin + s.dir * dmin;
        x.dir = s.dir;
        x.length = fminf(dmax - dmin, m.max_beam_length);
        x.for_pdf = fpdf;
        x.rev_pdf = rpdf;
        x.flags = kEndInMedium | (i == 0 ? flags : 0u);
        x.throughput = t;
        x.path_idx = pid;
        x.sw = s.sw;
        x.path_length = s.path_length;
        x.infinite = s.infinite;
        store_beam(&b.beams[size_t(pid) * p.bstride + nbeam], x);
        nbeam++;
    };
    if (p.pbt == T_BEAM_SHORT) {
        for (int i = 0; i < nseg; i++) {
            const VolSeg& sg = segs[i];
            const DMedium& m = sc.media[sg.med];
            if (m.has_scattering) emit_beam(sg.dmin, sg.dmax, m, sg.med, i);
            t *= sg.att / sg.pf;
            fpdf *= sg.pf;
            rpdf *= sg.pr;
        }
    } else {
        for (int i = 0; i < nlite; i++) {
            const LiteSeg& sg = lite[i];
            if (sg.med < 0) continue;
            const DMedium& m = sc.media[sg.med];
            if (m.has_scattering) emit_beam(sg.dmin, sg.dmax, m, sg.med, i);
            t *= eval_attenuation(m, sg.dmax - sg.dmin);
            float srp;
            const float sfp = ray_sample_pdf(m, sg.dmin, sg.dmax, i == 0 ? flags : 0u, &srp);
            fpdf *= sfp;
            rpdf *= srp;
        }
    }
}

// kernel/functs_light.cuh:251-388
__device__ __forceinline__ int connect_to_camera(const DScene& sc, const IterParams& p, const Bufs& b, const Sub& s,
                                                 f3 hitp, const Bsdf& bs, float* wout = nullptr)
{
    const DCamera& cam = sc.cam;
    f3 to_cam = cam.origin - hitp;
    if (dot(cam.forward, -to_cam) <= 0.f) return S_OK;
    float ix, iy;
    world_to_raster(cam, hitp, ix, iy);
    if (!(ix >= 0 && iy >= 0 && ix < cam.res_x && iy < cam.res_y)) return S_OK;
    const float d2 = dot(to_cam, to_cam);
    const float dist = sqrtf(d2);
    to_cam = to_cam / dist;
    float cos_to_cam, fpdf, rpdf, sin_theta;
    const f3 f = bsdf_evaluate(bs, to_cam, sc, cos_to_cam, &fpdf, &rpdf, &sin_theta);
    if (black_or_neg(f)) return S_OK;
    rpdf *= bs.cont_prob;
    const float cos_at_cam = dot(cam.forward, -to_cam);
    const float ip_dist = cam.ipd / cos_at_cam;
    const float i2sa = ip_dist * ip_dist / cos_at_cam;
    const float i2s = i2sa * fabsf(cos_to_cam) / (dist * dist);
    SegAcc sa;
    bool occ;
    const int ret = occluded(sc, occ, hitp, to_cam, dist, s.stack, bs.in_medium() ? kOriginInMedium : 0u, sa);
    if (ret != S_OK) return ret;
    if (occ) return S_OK;
    const float rsr = sa.pr;
    const f3 att = sa.att;
    if (!positive(att)) return S_OK;
    const Weights& wl = s.sw.w;
    const float f_s = pde_factor(p, bs.in_medium(), false, wl.for_inv, wl.for_ratio, 1.f / rsr, wl.rev_ratio, sin_theta);
    const float bpt_s1 = float(!wl.prev_specular);
    const float w_light = (i2s / float(p.n_light)) * rsr * (f_s + bpt_s1 * wl.d_shared + (rpdf / wl.rev_inv) * wl.d_bpt);
    const float mis = 1.f / (w_light + 1.f);
    if (wout) *wout = mis;
    const float s2i = 1.f / i2s;
    const f3 contrib = mis * s.throughput * f * att / (float(p.n_light) * s2i);
    if (black_or_neg(contrib)) return S_OK;
    const size_t px = size_t(int(ix)) + size_t(int(iy)) * size_t(cam.res_x);
    atomicAdd(&b.fb32[3 * px], contrib.x);
    atomicAdd(&b.fb32[3 * px + 1], contrib.y);
    atomicAdd(&b.fb32[3 * px + 2], contrib.z);
    return S_OK;
}

// One segment of kernel/trace_light.cu:84-321. Returns whether the path continues. LOG adds the
// MIS weight log of light tracing from the first light paths (Bufs::mislog).
template <bool LOG>
__device__ __forceinline__ bool light_step(const DScene& sc, const IterParams& p, const Bufs& b, uint32_t pid, Sub& s,
                                           uint32_t& pflags, int& ret, uint32_t& nvert, uint32_t& nbeam,
                                           uint32_t& bdrops)
{
    const bool store_beams = (p.do_bb1d && pid < p.bb1d_used) || p.do_bp2d;
    const bool origin_in_medium = (pflags & PF_ORIGIN_IN_MEDIUM) != 0;
    const uint32_t seg = s.path_length - 1;
    Rng mrng(kDomainLight, pid, p.iter_key, seg, P_MEDIUM);
    Hit is;
    hit_reset(is, kInf);
    VolSeg segs[kMaxSegments];
    LiteSeg lite[kMaxSegments];
    int nseg, nlite;
    bool hit;
    const unsigned rf = origin_in_medium ? kOriginInMedium : 0u;
    ret = intersect(sc, hit, s.origin, s.dir, is, s.stack, rf, segs, nseg, lite, nlite, &mrng);
    if (ret != S_OK) return false;
    if (store_beams) add_beams(sc, p, b, pid, s, rf, segs, nseg, lite, nlite, nbeam, bdrops);
    if (!hit) return false;
    float rsf = 1.f, rsr = 1.f;
    if (nseg) {
        rsf = acc_for(segs, nseg);
        rsr = acc_rev(segs, nseg);
        s.throughput *= acc_att(segs, nseg) / rsf;
    }
    if (black_or_neg(s.throughput)) return false;
    Weights& w = s.sw.w;
    if (s.path_length > 1u) {
        const float fi = pde_factor(p, w.prev_in_medium, w.prev_specular, w.for_inv, w.for_ratio, 1.f / rsr,
                                    w.rev_ratio, s.sw.last_sin);
        w.d_bpt = s.sw.d_bpt_a * fi + s.sw.d_bpt_b;
        w.d_pde = s.sw.d_pde_a * fi + s.sw.d_pde_b;
    }
    w.d_shared /= rsf;
    w.d_bpt /= rsf;
    w.d_pde /= rsf;
    Bsdf bs;
    bsdf_setup(bs, s.dir, is, sc, true, relative_ior(sc, is, s.stack));
    if (!bs.valid()) return false;
    float for_ratio = 0.f, rev_ratio = 0.f;
    if (bs.in_medium()) for_ratio = rev_ratio = 1.f / sc.media[bs.id].min_pos_att;
    const f3 hitp = s.origin + s.dir * is.dist;
    bs.pos = hitp;
    const bool now_in_medium = is.in_medium();
    pflags = now_in_medium ? (pflags | PF_ORIGIN_IN_MEDIUM) : (pflags & ~PF_ORIGIN_IN_MEDIUM);
    {
        const float d = bs.d_factor();
        if (!(s.path_length == 1 && s.infinite)) w.d_shared *= is.dist * is.dist;
        w.d_shared /= d;
        w.d_bpt /= d;
        w.d_pde /= d;
        w.for_inv = 1.f / rsf;
        w.rev_inv = 1.f / rsr;
        w.for_ratio = for_ratio;
        w.rev_ratio = rev_ratio;
    }
    if (!bs.delta() && nvert < p.max_len - 1) {
        LVertex lv;
        lv.hit = hitp;
        lv.throughput = s.throughput;
        lv.path_length = s.path_length;
        lv.bsdf = bs;
        lv.in_medium = now_in_medium;
        lv.behind_surf = false;
        if (now_in_medium && s.stack.n > 0) {
            const int m = s.stack.mat[s.stack.n - 1];
            if (m >= 0 && sc.mats[m].real) lv.behind_surf = true;
        }
        lv.w = w;
        store_vertex(&b.verts[size_t(pid) * p.vstride + nvert], lv, bs.in_medium() ? s.dir : is.shading);
        nvert++;
    }
    if (p.do_bpt && !bs.delta() && (bs.in_medium() || p.connect_from_surf)) {
        if (s.path_length + 1 >= p.min_len) {
            // logged when every vertex of the path so far was stored
            const bool logged = LOG && pid < b.mis_lt && nvert == s.path_length;
            float wl = -1.f;
            ret = connect_to_camera(sc, p, b, s, hitp, bs, logged ? &wl : nullptr);
            if (ret != S_OK) return false;
            if (logged && wl >= 0.f) mis_log(b, MT_LT, s.path_length + 1, s.path_length + 1, pid, pid, wl);
        }
    }
    if (s.path_length + 2 > p.max_len) return false;
    bool cont;
    Rng srng(kDomainLight, pid, p.iter_key, seg, P_SCATTER);
    ret = sample_scattering(sc, p, cont, bs, hitp, is, s, srng);
    if (ret != S_OK) return false;
    return cont;
}

template <bool LOG>
__device__ __forceinline__ void light_bounce_body(const DScene& sc, const Bufs& b, int bounce, uint32_t vb)
{
    const IterParams& p = *b.params;
    const uint32_t i = vb * blockDim.x + threadIdx.x;
    const uint32_t n = b.qcount[bounce];
    if (i >= n) return;
    const uint32_t pid = b.queue[bounce & 1][i];
    PathState st = b.lstate[pid];
    Sub s;
    load_sub(st, s);
    load_stack(b.lstack, pid, s.stack);
    uint32_t nvert = b.nvert[pid], nbeam = b.nbeam[pid];
    const uint32_t nvert0 = nvert, nbeam0 = nbeam;
    uint32_t pflags = st.flags, bdrops = 0;
    int ret = S_OK;
    bool cont = light_step<LOG>(sc, p, b, pid, s, pflags, ret, nvert, nbeam, bdrops);
    if (ret != S_OK) {
        cont = false;
        stat_add(b.stats, STAT_ERRORS, 1);
    }
    if (bdrops) stat_add(b.stats, STAT_BDROP, bdrops);
    if (nvert != nvert0) b.nvert[pid] = nvert;
    if (nbeam != nbeam0) b.nbeam[pid] = nbeam;
    if (cont) {
        s.path_length++;
        st.flags = pflags;
        save_sub(st, s);
        b.lstate[pid] = st;
        store_stack(b.lstack, pid, s.stack);
    }
    queue_append(cont, pid, b.queue[(bounce + 1) & 1], &b.qcount[bounce + 1]);
}

template <bool LOG>
__global__ void __launch_bounds__(256) k_light_bounce(DScene sc, Bufs b, int bounce)
{
    light_bounce_body<LOG>(sc, b, bounce, blockIdx.x);
}

__global__ void __launch_bounds__(256) k_light_bounce_batch(const BatchJob* jobs, const uint2* bmap, int bounce)
{
    const uint2 m = bmap[blockIdx.x];
    light_bounce_body<false>(jobs[m.x].sc, jobs[m.x].b, bounce, m.y);
}

void launch_light_generate(const DScene& sc, const IterParams& p, const Bufs& b, hipStream_t s)
{
    hipLaunchKernelGGL(k_light_generate, dim3((p.n_light + 255) / 256), dim3(256), 0, s, sc, b);
}

void launch_light_generate_batch(const BatchJob* jobs, const uint2* bmap, uint32_t blocks, hipStream_t s)
{
    if (blocks) hipLaunchKernelGGL(k_light_generate_batch, dim3(blocks), dim3(256), 0, s, jobs, bmap);
}

void launch_light_bounce_batch(const BatchJob* jobs, const uint2* bmap, uint32_t blocks, int bounce, hipStream_t s)
{
    if (blocks) hipLaunchKernelGGL(k_light_bounce_batch, dim3(blocks), dim3(256), 0, s, jobs, bmap, bounce);
}

void launch_light_bounce(const DScene& sc, const IterParams& p, const Bufs& b, int bounce, hipStream_t s)
{
    if (b.mislog)
        hipLaunchKernelGGL(k_light_bounce<true>, dim3((p.n_light + p.block - 1) / p.block), dim3(p.block), 0, s, sc, b,
                           bounce);
    else
        hipLaunchKernelGGL(k_light_bounce<false>, dim3((p.n_light + p.block - 1) / p.block), dim3(p.block), 0, s, sc, b,
                           bounce);
}

} // namespace upbp
