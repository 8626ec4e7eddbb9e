// Connection and merge estimators with their constant-time MIS weights, shared by the
// camera bounce kernel (light hits, next-event estimation, vertex connections, SURF /
// P-P3D / B-P2D point merges) and the segment merge kernels (P-B2D, B-B1D).
// Semantics: nolmoonen/gpuupbp gpuupbp/kernel/functs_camera.cuh:107-527 and
// gpuupbp/kernel/pde/functs_{pp,pb2d,bb1d,bp2d}.cuh.
#pragma once

#include "passes.h"

namespace upbp {

// kernel/functs_camera.cuh:107-165
// Every estimator below takes an optional wout: when given, it receives the MIS weight of the
// contribution if one is weighted, and is left unchanged otherwise.
__device__ __forceinline__ f3 light_radiance_mis(const DScene& sc, const IterParams& p, const DLight& l, const Sub& s,
                                                 float rsr, float* wout = nullptr)
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
    if (wout) *wout = 1.f / (1.f + w_cam);
    return rad * (1.f / (1.f + w_cam));
}

// kernel/functs_camera.cuh:170-344
__device__ __forceinline__ int direct_illumination(const DScene& sc, const IterParams& p, f3& contrib, const Sub& s,
                                                   f3 hitp, const Bsdf& bs, Rng& rng, float* wout = nullptr,
                                                   int* light_out = nullptr)
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
    SegAcc sa;
    bool occ;
    const int ret = occluded(sc, occ, hitp, dir, dist, s.stack, bs.in_medium() ? kOriginInMedium : 0u, sa);
    if (ret != S_OK) return ret;
    if (occ) return S_OK;
    const float nrr = sa.pr;
    const f3 natt = sa.att;
    if (!positive(natt)) return S_OK;
    const float w_light = fpdf / (pick * dpdf_w);
    const Weights& w = s.sw.w;
    const float ft = pde_factor(p, bs.in_medium(), false, 1.f / nrr, w.rev_ratio, w.for_inv, w.for_ratio, sin_theta);
    const float bt1 = float(!w.prev_specular);
    const float w_cam =
        (epdf_w * cos_to / (dpdf_w * cos_at)) * nrr * (ft + bt1 * w.d_shared + (rpdf / w.rev_inv) * w.d_bpt);
    const float mis = 1.f / (w_light + 1.f + w_cam);
    if (wout) {
        *wout = mis;
        *light_out = id;
    }
    contrib = ((cos_to / (pick * dpdf_w)) * (rad * natt * f)) * mis;
    if (black_or_neg(contrib)) contrib = mk(0.f);
    return S_OK;
}

// kernel/functs_camera.cuh:349-527
__device__ __forceinline__ int connect_vertices(const DScene& sc, const IterParams& p, f3& contrib, const LVertex& lv,
                                                const Bsdf& bc, f3 chit, const Sub& s, float* wout = nullptr)
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
    SegAcc sa;
    bool occ;
    const int ret = occluded(sc, occ, chit, dir, dist, s.stack, rf, sa);
    if (ret != S_OK) return ret;
    if (occ) return S_OK;
    const float rsf = sa.pf, rsr = sa.pr;
    const f3 att = sa.att;
    if (!positive(att)) return S_OK;
    const Weights& wl = lv.w;
    const Weights& wc = s.sw.w;
    const float fs = pde_factor(p, lv.bsdf.in_medium(), false, wl.for_inv, wl.for_ratio, 1.f / rsf, wl.rev_ratio, sin_l);
    const float bs1 = float(!wl.prev_specular);
    const float w_light = c_pdf_a * rsf * (fs + bs1 * wl.d_shared + (lrp / wl.rev_inv) * wl.d_bpt);
    const float ft = pde_factor(p, bc.in_medium(), false, 1.f / rsr, wc.rev_ratio, wc.for_inv, wc.for_ratio, sin_c);
    const float bt1 = float(!wc.prev_specular);
    const float w_cam = l_pdf_a * rsr * (ft + bt1 * wc.d_shared + (crp / wc.rev_inv) * wc.d_bpt);
    const float mis = 1.f / (w_light + 1.f + w_cam);
    if (wout) *wout = mis;
    contrib = (g * fc * fl * att) * mis;
    if (black_or_neg(contrib)) contrib = mk(0.f);
    return S_OK;
}

// kernel/pde/functs_pp.cuh:44-136 (SURF and P-P3D merge of one photon)
__device__ __forceinline__ void eval_pp(const DScene& sc, const IterParams& p, f3& acc, const Sub& s, const Bsdf& bc,
                                        const LVertex& lv, float* wout = nullptr)
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
    if (wout) *wout = mis;
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
                                          const StateWeights& cw, const LVertex& lv, float idist, float ir2,
                                          float* wout = nullptr)
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
    const float hg = hg_pdf(-q.rd, pdir, m.mean_cosine, &sin_theta);
    // camera technique (learned mixture when enabled) toward the light side; light technique
    // (the phase function) toward the camera side
    float fpdf = cam_medium_pdf(sc, m, q.ro + q.rd * idist, q.rd, pdir, nullptr);
    float rpdf = hg;
    const f3 f = mk(hg);
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
    if (wout) *wout = 1.f / (w_light + w_local + w_cam);
    acc += (1.f / (w_light + w_local + w_cam)) * unweighted;
}

// kernel/pde/functs_bb1d.cuh:57-248 (acceptance already decided by the caller)
__device__ __forceinline__ void eval_bb1d(const DScene& sc, const IterParams& p, f3& acc, const SegQ& q,
                                          const StateWeights& cw, const LBeam& lb, float bbd, float sin_theta,
                                          float tc, float tp, float* wout = nullptr)
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
    const float hg = hg_pdf(-q.rd, -lb.dir, m.mean_cosine, nullptr);
    float fpdf = cam_medium_pdf(sc, m, q.ro + q.rd * tc, q.rd, -lb.dir, nullptr);
    float rpdf = hg;
    const f3 f = mk(hg);
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
    if (wout) *wout = 1.f / (w_light + w_local + w_cam);
    acc += (1.f / (w_light + w_local + w_cam)) * unweighted;
}

// kernel/pde/functs_bp2d.cuh:51-198
__device__ __forceinline__ void eval_bp2d(const DScene& sc, const IterParams& p, f3& acc, const Sub& s, const Bsdf& bc,
                                          const LBeam& lb, float idist, float ir2, float* wout = nullptr)
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
    const float hg = hg_pdf(cdir, -lb.dir, m.mean_cosine, &sin_theta);
    float fpdf = cam_medium_pdf(sc, m, bc.pos, -cdir, -lb.dir, nullptr);
    float rpdf = hg;
    const f3 f = mk(hg);
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
    if (wout) *wout = 1.f / (w_light + w_local + w_cam);
    acc += (1.f / (w_light + w_local + w_cam)) * unweighted;
}

// SegRec <-> query and camera state
__device__ __forceinline__ void load_segrec(const SegRec* r, SegQ& q, StateWeights& cw, f3& tatt, uint32_t& pix)
{
    const float4 a = r->a, b = r->b, c = r->c, d = r->d, e = r->e, f = r->f, g = r->g;
    const uint32_t bits = uint32_t(__float_as_int(d.w));
    q.ro = mk(a.x, a.y, a.z);
    q.dmin = a.w;
    q.rd = mk(b.x, b.y, b.z);
    q.dmax = b.w;
    tatt = mk(c.x, c.y, c.z);
    pix = uint32_t(__float_as_int(c.w));
    q.rsf = d.x;
    q.rsr = d.y;
    q.med = __float_as_int(d.z);
    q.flags = bits & 3u;
    q.cam_len = bits >> 8;
    cw.w.prev_in_medium = (bits >> 2) & 1u;
    cw.w.prev_specular = (bits >> 3) & 1u;
    cw.w.d_shared = e.x;
    cw.w.d_bpt = e.y;
    cw.w.d_pde = e.z;
    cw.w.for_inv = e.w;
    cw.w.rev_inv = f.x;
    cw.w.for_ratio = f.y;
    cw.w.rev_ratio = f.z;
    cw.last_sin = f.w;
    cw.d_bpt_a = g.x;
    cw.d_bpt_b = g.y;
    cw.d_pde_a = g.z;
    cw.d_pde_b = g.w;
}

__device__ __forceinline__ void store_segrec(SegRec* r, const SegQ& q, const StateWeights& cw, f3 tatt, uint32_t pix)
{
    const uint32_t bits = (q.flags & 3u) | (cw.w.prev_in_medium ? 4u : 0u) | (cw.w.prev_specular ? 8u : 0u) |
                          (q.cam_len << 8);
    r->a = make_float4(q.ro.x, q.ro.y, q.ro.z, q.dmin);
    r->b = make_float4(q.rd.x, q.rd.y, q.rd.z, q.dmax);
    r->c = make_float4(tatt.x, tatt.y, tatt.z, __int_as_float(int(pix)));
    r->d = make_float4(q.rsf, q.rsr, __int_as_float(q.med), __int_as_float(int(bits)));
    r->e = make_float4(cw.w.d_shared, cw.w.d_bpt, cw.w.d_pde, cw.w.for_inv);
    r->f = make_float4(cw.w.rev_inv, cw.w.for_ratio, cw.w.rev_ratio, cw.last_sin);
    r->g = make_float4(cw.d_bpt_a, cw.d_bpt_b, cw.d_pde_a, cw.d_pde_b);
}

} // namespace upbp
