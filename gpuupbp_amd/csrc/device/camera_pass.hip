// Camera-subpath wavefront with the five photon-density estimators and BPT fused into
// one bounce kernel per path segment: P-B2D and B-B1D along the segment's media
// (segment-vs-point and segment-vs-capsule LBVH queries), light hits, next-event
// estimation, vertex connections over the chosen light path's vertex range, surface /
// P-P3D / B-P2D merges (point queries), then scattering and wave-aggregated queue
// compaction. Every contribution carries its constant-time MIS weight computed in place.
// Semantics: nolmoonen/gpuupbp gpuupbp/kernel/trace_camera.cu:49-390,
// gpuupbp/kernel/functs_camera.cuh:41-527 and gpuupbp/kernel/pde/functs_*.cuh.
#include "kernels.h"

namespace upbp {

// kernel/functs_camera.cuh:107-165
__device__ __forceinline__ f3 light_radiance_mis(const DScene& sc, const IterParams& p, const DLight& l, const Sub& s,
                                                 float rsr)
{
    if (s.specular_path && p.ignore_spec) return mk(0.f);
    const float pick = 1.f / float(sc.n_lights);
    float dpdf_a = 0.f, epdf_w = 0.f;
    const f3 rad = light_radiance(l, sc, s.dir, &dpdf_a, &epdf_w);
    if (black_or_neg(rad)) return mk(0.f);
    if (s.path_length == 1) return rad;
    if (p.techniques && !(p.techniques & T_BPT)) return s.specular_path ? rad : mk(0.f);
    dpdf_a *= pick;
    epdf_w *= pick;
    const Weights& w = s.sw.w;
    const float bpt_t1 = float(!w.prev_specular);
    const float w_cam = dpdf_a * bpt_t1 * w.d_shared + epdf_w * rsr * w.d_bpt;
    return rad * (1.f / (1.f + w_cam));
}

// kernel/functs_camera.cuh:170-344
__device__ __forceinline__ int direct_illumination(const DScene& sc, const IterParams& p, f3& contrib, const Sub& s,
                                                   f3 hitp, const Bsdf& bs, Rng& rng)
{
    const float pick = 1.f / float(sc.n_lights);
    const int id = int(rng.next() * float(sc.n_lights));
    const float u0 = rng.next(), u1 = rng.next();
    const DLight& l = sc.lights[id];
    contrib = mk(0.f);
    if (!light_finite(l) && sc.media[sc.global_medium].has_attenuation) return S_OK;
    f3 dir;
    float dist, dpdf_w, epdf_w, cos_at;
    const f3 rad = light_illuminate(l, sc, hitp, u0, u1, dir, dist, dpdf_w, &epdf_w, &cos_at);
    if (black_or_neg(rad)) return S_OK;
    float fpdf, rpdf, cos_to, sin_theta;
    const f3 f = bsdf_evaluate(bs, dir, sc, cos_to, &fpdf, &rpdf, &sin_theta);
    if (black_or_neg(f)) return S_OK;
    const float cp = bs.cont_prob;
    fpdf *= light_delta(l) ? 0.f : cp;
    rpdf *= cp;
    VolSeg segs[kMaxSegments];
    int nseg;
    bool occ;
    const int ret = occluded(sc, occ, hitp, dir, dist, s.stack, bs.in_medium() ? kOriginInMedium : 0u, segs, nseg);
    if (ret != S_OK) return ret;
    if (occ) return S_OK;
    float nrr = 1.f;
    f3 natt = mk(1.f);
    if (nseg) {
        nrr = acc_rev(segs, nseg);
        natt = acc_att(segs, nseg);
        if (!positive(natt)) return S_OK;
    }
    const float w_light = fpdf / (pick * dpdf_w);
    const Weights& w = s.sw.w;
    const float ft = pde_factor(p, bs.in_medium(), false, 1.f / nrr, w.rev_ratio, w.for_inv, w.for_ratio, sin_theta);
    const float bt1 = float(!w.prev_specular);
    const float w_cam =
        (epdf_w * cos_to / (dpdf_w * cos_at)) * nrr * (ft + bt1 * w.d_shared + (rpdf / w.rev_inv) * w.d_bpt);
    const float mis = 1.f / (w_light + 1.f + w_cam);
    contrib = ((cos_to / (pick * dpdf_w)) * (rad * natt * f)) * mis;
    if (black_or_neg(contrib)) contrib = mk(0.f);
    return S_OK;
}

// kernel/functs_camera.cuh:349-527
__device__ __forceinline__ int connect_vertices(const DScene& sc, const IterParams& p, f3& contrib, const LVertex& lv,
                                                const Bsdf& bc, f3 chit, const Sub& s)
{
    contrib = mk(0.f);
    f3 dir = lv.hit - chit;
    const float d2 = dot(dir, dir);
    const float dist = sqrtf(d2);
    dir = dir / dist;
    float cos_c, cfp, crp, sin_c;
    const f3 fc = bsdf_evaluate(bc, dir, sc, cos_c, &cfp, &crp, &sin_c);
    if (black_or_neg(fc)) return S_OK;
    cfp *= bc.cont_prob;
    crp *= bc.cont_prob;
    float cos_l, lfp, lrp, sin_l;
    const f3 fl = bsdf_evaluate(lv.bsdf, -dir, sc, cos_l, &lfp, &lrp, &sin_l);
    if (black_or_neg(fl)) return S_OK;
    lfp *= lv.bsdf.cont_prob;
    lrp *= lv.bsdf.cont_prob;
    const float g = cos_l * cos_c / d2;
    if (g < 0.f) return S_OK;
    const float c_pdf_a = pdf_w_to_a(cfp, dist, cos_l);
    const float l_pdf_a = pdf_w_to_a(lfp, dist, cos_c);
    unsigned rf = 0;
    if (bc.in_medium()) rf |= kOriginInMedium;
    if (lv.in_medium) rf |= kEndInMedium;
    VolSeg segs[kMaxSegments];
    int nseg;
    bool occ;
    const int ret = occluded(sc, occ, chit, dir, dist, s.stack, rf, segs, nseg);
    if (ret != S_OK) return ret;
    if (occ) return S_OK;
    float rsf = 1.f, rsr = 1.f;
    f3 att = mk(1.f);
    if (nseg) {
        rsf = acc_for(segs, nseg);
        rsr = acc_rev(segs, nseg);
        att = acc_att(segs, nseg);
        if (!positive(att)) return S_OK;
    }
    const Weights& wl = lv.w;
    const Weights& wc = s.sw.w;
    const float fs = pde_factor(p, lv.bsdf.in_medium(), false, wl.for_inv, wl.for_ratio, 1.f / rsf, wl.rev_ratio, sin_l);
    const float bs1 = float(!wl.prev_specular);
    const float w_light = c_pdf_a * rsf * (fs + bs1 * wl.d_shared + (lrp / wl.rev_inv) * wl.d_bpt);
    const float ft = pde_factor(p, bc.in_medium(), false, 1.f / rsr, wc.rev_ratio, wc.for_inv, wc.for_ratio, sin_c);
    const float bt1 = float(!wc.prev_specular);
    const float w_cam = l_pdf_a * rsr * (ft + bt1 * wc.d_shared + (crp / wc.rev_inv) * wc.d_bpt);
    const float mis = 1.f / (w_light + 1.f + w_cam);
    contrib = (g * fc * fl * att) * mis;
    if (black_or_neg(contrib)) contrib = mk(0.f);
    return S_OK;
}

// kernel/pde/functs_pp.cuh:44-136 (SURF and P-P3D merge of one photon)
__device__ __forceinline__ void eval_pp(const DScene& sc, const IterParams& p, f3& acc, const Sub& s, const Bsdf& bc,
                                        const LVertex& lv)
{
    const bool medium = lv.in_medium;
    const uint32_t len = lv.path_length + s.path_length;
    if (len > p.max_len || len < p.min_len) return;
    const f3 pdir = lv.bsdf.world_dir_fix();
    float cos_c, fpdf, rpdf, sin_theta;
    const f3 f = bsdf_evaluate(bc, pdir, sc, cos_c, &fpdf, &rpdf, &sin_theta);
    if (black_or_neg(f)) return;
    fpdf *= bc.cont_prob;
    rpdf *= lv.bsdf.cont_prob;
    const Weights& wl = lv.w;
    const Weights& wc = s.sw.w;
    const float fpf = p.pbt == T_BEAM_LONG ? wl.for_inv : wl.for_ratio;
    const float rpf = p.qbt == T_BEAM_LONG ? wc.for_inv : wc.for_ratio;
    const float val = medium ? 1.f / p.mf_pp3d : 1.f / p.mf_surf;
    const float bs1 = float(!wl.prev_specular);
    const float w_light = val * (wl.d_shared * bs1 * p.n_bpt + fpdf * wl.d_pde / wl.rev_inv);
    const float w_local =
        1.f + float(medium) * val * (p.mf_bb1d * sin_theta * fpf * rpf + p.mf_bp2d * fpf + p.mf_pb2d * rpf);
    const float bt1 = float(!wc.prev_specular);
    const float w_cam = val * (wc.d_shared * bt1 * p.n_bpt + rpdf * wc.d_pde / wc.rev_inv);
    const float mis = 1.f / (w_light + w_local + w_cam);
    acc += mis * f * lv.throughput;
}

struct SegQ {
    f3 ro, rd;
    float dmin, dmax;
    int med;
    float rsf, rsr;
    unsigned flags;
    uint32_t cam_len;
};

// kernel/pde/functs_pb2d.cuh:57-201
__device__ __forceinline__ void eval_pb2d(const DScene& sc, const IterParams& p, f3& acc, const SegQ& q,
                                          const StateWeights& cw, const LVertex& lv, float idist, float ir2)
{
    const uint32_t len = lv.path_length + q.cam_len;
    if (len > p.max_len || len < p.min_len) return;
    const DMedium& m = sc.media[q.med];
    f3 att = eval_attenuation(m, idist - q.dmin);
    const float pdf = comp(att, m.min_pos_idx);
    if (p.qbt == T_BEAM_SHORT) att /= pdf;
    float rsf = m.min_pos_att * pdf;
    float rsr = (q.flags & kOriginInMedium) ? rsf : pdf;
    const float ratio = 1.f / m.min_pos_att;
    if (!positive(att)) return;
    rsf *= q.rsf;
    rsr *= q.rsr;
    Weights wc = cw.w;
    if (q.cam_len > 1u) {
        const float fi = pde_factor(p, wc.prev_in_medium, wc.prev_specular, 1.f / rsr, wc.rev_ratio, wc.for_inv,
                                    wc.for_ratio, cw.last_sin);
        wc.d_pde = cw.d_pde_a * fi + cw.d_pde_b;
        wc.d_bpt = cw.d_bpt_a * fi + cw.d_bpt_b;
    }
    const f3 pdir = lv.bsdf.world_dir_fix();
    float sin_theta;
    float fpdf = hg_pdf(-q.rd, pdir, m.mean_cosine, &sin_theta);
    float rpdf = fpdf;
    const f3 f = mk(fpdf);
    if (black_or_neg(f)) return;
    fpdf *= m.cont_prob;
    rpdf *= lv.bsdf.cont_prob;
    wc.d_shared *= idist * idist;
    wc.d_shared /= rsf;
    wc.d_bpt /= rsf;
    wc.d_pde /= rsf;
    wc.for_inv = 1.f / rsf;
    wc.rev_inv = 1.f / rsr;
    wc.for_ratio = ratio;
    wc.rev_ratio = ratio;
    const float kernel = (1.f - ir2 / p.r2_pb2d) / (p.r2_pb2d * kPi * .5f);
    if (!positive(kernel)) return;
    const f3 unweighted = lv.throughput * att * m.scattering * f * kernel;
    if (black_or_neg(unweighted)) return;
    const Weights& wl = lv.w;
    const float fpf = p.pbt == T_BEAM_LONG ? wl.for_inv : wl.for_ratio;
    const float rpf = p.qbt == T_BEAM_LONG ? wc.for_inv : wc.for_ratio;
    const float val = 1.f / (p.mf_pb2d * rpf);
    const float bs1 = float(!wl.prev_specular);
    const float w_light = val * (wl.d_shared * bs1 * p.n_bpt + fpdf * wl.d_pde / wl.rev_inv);
    const float w_local = 1.f + val * (p.mf_bb1d * sin_theta * fpf * rpf + p.mf_bp2d * fpf + p.mf_pp3d);
    const float bt1 = float(!wc.prev_specular);
    const float w_cam = val * (wc.d_shared * bt1 * p.n_bpt + rpdf * wc.d_pde / wc.rev_inv);
    acc += (1.f / (w_light + w_local + w_cam)) * unweighted;
}

// kernel/pde/functs_bb1d.cuh:57-248 (acceptance already decided by the caller)
__device__ __forceinline__ void eval_bb1d(const DScene& sc, const IterParams& p, f3& acc, const SegQ& q,
                                          const StateWeights& cw, const LBeam& lb, float bbd, float sin_theta,
                                          float tc, float tp)
{
    const uint32_t len = lb.path_length + q.cam_len;
    if (len > p.max_len || len < p.min_len) return;
    const DMedium& m = sc.media[q.med];
    f3 attc = eval_attenuation(m, tc - q.dmin);
    const float pdfc = comp(attc, m.min_pos_idx);
    if (p.qbt == T_BEAM_SHORT) attc /= pdfc;
    float rsfc = m.min_pos_att * pdfc;
    float rsrc = (q.flags & kOriginInMedium) ? rsfc : pdfc;
    const float ratio = 1.f / m.min_pos_att;
    f3 attp = eval_attenuation(m, tp);
    const float pdfp = comp(attp, m.min_pos_idx);
    if (p.pbt == T_BEAM_SHORT) attp /= pdfp;
    float rsfp = m.min_pos_att * pdfp;
    float rsrp = (lb.flags & kOriginInMedium) ? rsfp : pdfp;
    const f3 att = attc * attp;
    if (!positive(att)) return;
    rsfc *= q.rsf;
    rsrc *= q.rsr;
    Weights wc = cw.w;
    if (q.cam_len > 1u) {
        const float fi = pde_factor(p, wc.prev_in_medium, wc.prev_specular, 1.f / rsrc, wc.rev_ratio, wc.for_inv,
                                    wc.for_ratio, cw.last_sin);
        wc.d_pde = cw.d_pde_a * fi + cw.d_pde_b;
        wc.d_bpt = cw.d_bpt_a * fi + cw.d_bpt_b;
    }
    rsfp *= lb.for_pdf;
    rsrp *= lb.rev_pdf;
    Weights wl = lb.sw.w;
    if (lb.path_length > 1u) {
        const float fi = pde_factor(p, wl.prev_in_medium, wl.prev_specular, wl.for_inv, wl.for_ratio, 1.f / rsrp,
                                    wl.rev_ratio, lb.sw.last_sin);
        wl.d_pde = lb.sw.d_pde_a * fi + lb.sw.d_pde_b;
        wl.d_bpt = lb.sw.d_bpt_a * fi + lb.sw.d_bpt_b;
    }
    float fpdf = hg_pdf(-q.rd, -lb.dir, m.mean_cosine, nullptr);
    float rpdf = fpdf;
    const f3 f = mk(fpdf);
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
    // light-side condition as written in functs_bb1d.cuh:191
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
    const f3 unweighted = lb.throughput * m.scattering * att * f * kernel;
    if (black_or_neg(unweighted)) return;
    const float fpf = p.pbt == T_BEAM_LONG ? wl.for_inv : wl.for_ratio;
    const float rpf = p.qbt == T_BEAM_LONG ? wc.for_inv : wc.for_ratio;
    const float val = 1.f / (p.mf_bb1d * sin_theta * fpf * rpf);
    const float bs1 = float(!wl.prev_specular);
    const float w_light = val * (wl.d_shared * bs1 * p.n_bpt + fpdf * wl.d_pde / wl.rev_inv);
    const float w_local = 1.f + val * (p.mf_bp2d * fpf + p.mf_pb2d * rpf + p.mf_pp3d);
    const float bt1 = float(!wc.prev_specular);
    const float w_cam = val * (wc.d_shared * bt1 * p.n_bpt + rpdf * wc.d_pde / wc.rev_inv);
    acc += (1.f / (w_light + w_local + w_cam)) * unweighted;
}

// kernel/pde/functs_bp2d.cuh:51-198
__device__ __forceinline__ void eval_bp2d(const DScene& sc, const IterParams& p, f3& acc, const Sub& s, const Bsdf& bc,
                                          const LBeam& lb, float idist, float ir2)
{
    const uint32_t len = lb.path_length + s.path_length;
    if (len > p.max_len || len < p.min_len) return;
    const DMedium& m = sc.media[lb.med];
    f3 att = eval_attenuation(m, idist);
    const float pdf = comp(att, m.min_pos_idx);
    if (p.pbt == T_BEAM_SHORT) att /= pdf;
    float rsf = m.min_pos_att * pdf;
    float rsr = (lb.flags & kOriginInMedium) ? rsf : pdf;
    const float ratio = 1.f / m.min_pos_att;
    if (!positive(att)) return;
    rsf *= lb.for_pdf;
    rsr *= lb.rev_pdf;
    Weights wl = lb.sw.w;
    if (lb.path_length > 1u) {
        const float fi = pde_factor(p, wl.prev_in_medium, wl.prev_specular, wl.for_inv, wl.for_ratio, 1.f / rsr,
                                    wl.rev_ratio, lb.sw.last_sin);
        wl.d_pde = lb.sw.d_pde_a * fi + lb.sw.d_pde_b;
        wl.d_bpt = lb.sw.d_bpt_a * fi + lb.sw.d_bpt_b;
    }
    const f3 cdir = bc.world_dir_fix();
    float sin_theta;
    float fpdf = hg_pdf(cdir, -lb.dir, m.mean_cosine, &sin_theta);
    float rpdf = fpdf;
    const f3 f = mk(fpdf);
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
    const f3 unweighted = lb.throughput * att * m.scattering * f * kernel;
    if (black_or_neg(unweighted)) return;
    const Weights& wc = s.sw.w;
    const float fpf = p.pbt == T_BEAM_LONG ? wl.for_inv : wl.for_ratio;
    const float rpf = p.qbt == T_BEAM_LONG ? wc.for_inv : wc.for_ratio;
    const float val = 1.f / (p.mf_bp2d * fpf);
    const float bs1 = float(!wl.prev_specular);
    const float w_light = val * (wl.d_shared * bs1 * p.n_bpt + fpdf * wl.d_pde / wl.rev_inv);
    const float w_local = 1.f + val * (p.mf_bb1d * sin_theta * fpf * rpf + p.mf_pb2d * rpf + p.mf_pp3d);
    const float bt1 = float(!wc.prev_specular);
    const float w_cam = val * (wc.d_shared * bt1 * p.n_bpt + rpdf * wc.d_pde / wc.rev_inv);
    acc += (1.f / (w_light + w_local + w_cam)) * unweighted;
}

// kernel/pde/functs_pb2d.cuh:226-342 and functs_bb1d.cuh:273-397: walk the segment's media,
// query the medium-point (P-B2D) or beam (B-B1D) LBVH per scattering medium
template <bool BB1D>
__device__ __noinline__ f3 eval_segments(const DScene& sc, const IterParams& p, const Bufs& b, f3 ro, f3 rd, const VolSeg* segs,
                            const LiteSeg* lite, int n, unsigned flags, uint32_t cam_len, const StateWeights& cw,
                            uint32_t& nq, uint32_t& nacc, bool& stack_ok)
{
    f3 result = mk(0.f), att = mk(1.f);
    float rsf = 1.f, rsr = 1.f;
    for (int i = 0; i < n; i++) {
        const float dmin = segs ? segs[i].dmin : lite[i].dmin;
        const float dmax = segs ? segs[i].dmax : lite[i].dmax;
        const int mid = segs ? segs[i].med : lite[i].med;
        if (mid < 0) continue;
        const DMedium& m = sc.media[mid];
        f3 sr = mk(0.f);
        if (m.has_scattering) {
            const SegQ q{ro, rd, dmin, dmax, mid, rsf, rsr, kEndInMedium | (i == 0 ? flags : 0u), cam_len};
            nq++;
            if (BB1D) {
                const PhotonMaps& M = b.maps;
                stack_ok &= bvh_ray_query(M.beam, ro, rd, dmin, dmax, [&](uint32_t k, float&) {
                    const float4 g0 = M.beam_geo[2 * k], g1 = M.beam_geo[2 * k + 1];
                    const uint32_t slot = uint32_t(__float_as_int(g1.w));
                    float bbd, st, tc, tp;
                    if (!pred_beam_beam(ro, rd, dmin, dmax, mk(g0.x, g0.y, g0.z), mk(g1.x, g1.y, g1.z), 0.f, g0.w,
                                        p.r2_bb1d, bbd, st, tc, tp))
                        return;
                    LBeam lb;
                    load_beam(&b.beams[slot], lb);
                    if (lb.path_idx >= p.bb1d_used || lb.med != mid) return; // BB1D mask, same-medium test
                    nacc++;
                    eval_bb1d(sc, p, sr, q, cw, lb, bbd, st, tc, tp);
                });
            } else {
                const PhotonMaps& M = b.maps;
                stack_ok &= bvh_ray_query(M.med, ro, rd, dmin, dmax, [&](uint32_t k, float&) {
                    const float4 pt = M.med_pts[k];
                    const f3 ce = mk(pt.x, pt.y, pt.z);
                    float idist, ir2;
                    if (!pred_point_beam(ro, rd, dmax, ce, p.r2_pb2d)) return;
                    if (!pred_bre(idist, ir2, ro, rd, dmin, dmax, ce, p.r2_pb2d)) return;
                    LVertex lv;
                    load_vertex(&b.verts[__float_as_int(pt.w)], lv);
                    nacc++;
                    eval_pb2d(sc, p, sr, q, cw, lv, idist, ir2);
                });
            }
        }
        result += att * sr;
        if (segs) {
            if (p.qbt == T_BEAM_SHORT) att *= segs[i].att / segs[i].pf;
            else att *= segs[i].att;
            if (!positive(att)) return result;
            rsf *= segs[i].pf;
            rsr *= segs[i].pr;
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

__global__ void __launch_bounds__(256) k_camera_generate(DScene sc, IterParams p, Bufs b)
{
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= p.n_pixels) return;
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

struct CamCounters {
    uint32_t conn = 0, pq = 0, sq = 0, acc = 0;
};

// One segment of kernel/trace_camera.cu:79-382. Returns whether the path continues.
__device__ bool camera_step(const DScene& sc, const IterParams& p, const Bufs& b, uint32_t pix, Sub& s, uint32_t& pflags,
                            f3& color, int& ret, CamCounters& cc)
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
    ret = intersect(sc, hit, s.origin, s.dir, is, s.stack, false, rf, segs, nseg, lite, nlite, &mrng);
    if (ret != S_OK) return false;
    const bool use_full = is.on_surface() || p.qbt == T_BEAM_SHORT || !hit;
    bool ok = true;
    if (p.build_pb2d && b.maps.med.n > 0) {
        const f3 r = use_full ? eval_segments<false>(sc, p, b, s.origin, s.dir, segs, nullptr, nseg, rf, s.path_length,
                                                     s.sw, cc.sq, cc.acc, ok)
                              : eval_segments<false>(sc, p, b, s.origin, s.dir, nullptr, lite, nlite, rf,
                                                     s.path_length, s.sw, cc.sq, cc.acc, ok);
        color += s.throughput * p.norm_pb2d * r;
    }
    if (p.build_bb1d && b.maps.beam.n > 0) {
        const f3 r = use_full ? eval_segments<true>(sc, p, b, s.origin, s.dir, segs, nullptr, nseg, rf, s.path_length,
                                                    s.sw, cc.sq, cc.acc, ok)
                              : eval_segments<true>(sc, p, b, s.origin, s.dir, nullptr, lite, nlite, rf, s.path_length,
                                                    s.sw, cc.sq, cc.acc, ok);
        color += s.throughput * p.norm_bb1d * r;
    }
    if (!ok) {
        ret = S_ERR_STACK;
        return false;
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
    if (is.light >= 0) {
        if (s.path_length < p.min_len) return false;
        color += s.throughput * light_radiance_mis(sc, p, sc.lights[is.light], s, rsr);
        return false;
    }
    if (s.path_length >= p.max_len) return false;
    if (p.do_bpt && !bs.delta() && s.path_length + 1 >= p.min_len && sc.n_lights > 0 &&
        (bs.in_medium() || !only_spec_surf)) {
        f3 ct;
        Rng nrng(kDomainCamera, pix, p.iter_key, seg, P_NEE);
        ret = direct_illumination(sc, p, ct, s, hitp, bs, nrng);
        if (ret != S_OK) return false;
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
                ret = connect_vertices(sc, p, ct, lv, bs, hitp, s);
                if (ret != S_OK) return false;
                color += s.throughput * lv.throughput * ct;
            }
        }
    }
    if (p.build_surf && b.maps.surf.n > 0 && bs.on_surface() && !bs.delta() && !only_spec_surf) {
        f3 acc = mk(0.f);
        cc.pq++;
        ok = bvh_point_query(b.maps.surf, hitp, [&](uint32_t k) {
            const float4 pt = b.maps.surf_pts[k];
            if (!pred_point_point(hitp, mk(pt.x, pt.y, pt.z), p.r2_surf)) return;
            LVertex lv;
            load_vertex(&b.verts[__float_as_int(pt.w)], lv);
            cc.acc++;
            eval_pp(sc, p, acc, s, bs, lv);
        });
        color += s.throughput * p.norm_surf * acc;
    }
    if (p.build_pp3d && b.maps.med.n > 0 && bs.in_medium() && !bs.delta()) {
        f3 acc = mk(0.f);
        cc.pq++;
        ok &= bvh_point_query(b.maps.med, hitp, [&](uint32_t k) {
            const float4 pt = b.maps.med_pts[k];
            if (!pred_point_point(hitp, mk(pt.x, pt.y, pt.z), p.r2_pp3d)) return;
            LVertex lv;
            load_vertex(&b.verts[__float_as_int(pt.w)], lv);
            cc.acc++;
            eval_pp(sc, p, acc, s, bs, lv);
        });
        color += s.throughput * p.norm_pp3d * acc;
    }
    if (p.build_bp2d && b.maps.beam.n > 0 && bs.in_medium()) {
        f3 acc = mk(0.f);
        cc.pq++;
        ok &= bvh_point_query(b.maps.beam, hitp, [&](uint32_t k) {
            const float4 g0 = b.maps.beam_geo[2 * k], g1 = b.maps.beam_geo[2 * k + 1];
            const f3 bo = mk(g0.x, g0.y, g0.z), bd = mk(g1.x, g1.y, g1.z);
            if (!pred_point_beam(bo, bd, g0.w, hitp, p.r2_bp2d)) return; // program_bp2d.cu:31-45
            float idist, ir2;
            if (!pred_bre(idist, ir2, bo, bd, 0.f, g0.w, hitp, p.r2_bp2d)) return;
            LBeam lb;
            load_beam(&b.beams[__float_as_int(g1.w)], lb);
            cc.acc++;
            eval_bp2d(sc, p, acc, s, bs, lb, idist, ir2);
        });
        color += s.throughput * p.norm_bp2d * acc;
    }
    if (!ok) {
        ret = S_ERR_STACK;
        return false;
    }
    bool cont;
    Rng srng(kDomainCamera, pix, p.iter_key, seg, P_SCATTER);
    ret = sample_scattering(sc, p, cont, bs, hitp, is, s, srng);
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

__global__ void __launch_bounds__(256) k_camera_bounce(DScene sc, IterParams p, Bufs b, int bounce)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
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
    bool cont = camera_step(sc, p, b, pix, s, pflags, color, ret, cc);
    if (ret != S_OK) {
        cont = false;
        stat_add(b.stats, STAT_ERRORS, 1);
    }
    stat_add(b.stats, STAT_BPT_CONN, cc.conn);
    stat_add(b.stats, STAT_POINT_Q, cc.pq);
    stat_add(b.stats, STAT_SEG_Q, cc.sq);
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

// adds every camera path's colour at its jittered raster position (frame_buffer.cuh:35-44)
__global__ void __launch_bounds__(256) k_camera_splat(IterParams p, Bufs b, float res_x, float res_y)
{
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= p.n_pixels) return;
    const PathState& st = b.cstate[pix];
    if (!(st.sx >= 0.f && st.sx < res_x && st.sy >= 0.f && st.sy < res_y)) return;
    const size_t tgt = size_t(int(st.sx)) + size_t(int(st.sy)) * size_t(res_x);
    atomicAdd(&b.fb32[3 * tgt], st.color.x);
    atomicAdd(&b.fb32[3 * tgt + 1], st.color.y);
    atomicAdd(&b.fb32[3 * tgt + 2], st.color.z);
}

// folds the FP32 iteration buffer into the FP64 cross-iteration accumulator
__global__ void __launch_bounds__(256) k_accumulate(const float* fb32, double* accum, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) accum[i] += double(fb32[i]);
}

void launch_camera_generate(const DScene& sc, const IterParams& p, const Bufs& b, hipStream_t s)
{
    hipLaunchKernelGGL(k_camera_generate, dim3((p.n_pixels + 255) / 256), dim3(256), 0, s, sc, p, b);
}

void launch_camera_bounce(const DScene& sc, const IterParams& p, const Bufs& b, int bounce, hipStream_t s)
{
    hipLaunchKernelGGL(k_camera_bounce, dim3((p.n_pixels + 255) / 256), dim3(256), 0, s, sc, p, b, bounce);
}

void launch_camera_finish(const IterParams& p, const Bufs& b, float res_x, float res_y, hipStream_t s)
{
    hipLaunchKernelGGL(k_camera_splat, dim3((p.n_pixels + 255) / 256), dim3(256), 0, s, p, b, res_x, res_y);
    const uint32_t n = 3 * p.n_pixels;
    hipLaunchKernelGGL(k_accumulate, dim3((n + 255) / 256), dim3(256), 0, s, b.fb32, b.accum, n);
}

} // namespace upbp
