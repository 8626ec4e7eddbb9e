// BSDF / phase-function records and light sampling on the device.
//
// Semantics of nolmoonen/gpuupbp gpuupbp/kernel/bsdf.cuh:64-660 and
// gpuupbp/kernel/light.cuh:41-543 (constant-colour background only).
// The BSDF record is a compact value type: it keeps the vector the shading
// frame is derived from, so a light vertex stores 52 bytes of BSDF state and
// the consumer rebuilds the identical frame.
#pragma once

#include "env_map.h"
#include "guide.h"
#include "upbp_device.h"

namespace upbp {

struct Bsdf {
    enum Event : unsigned { DIFFUSE = 1, PHONG = 2, REFLECT = 4, REFRACT = 8, SCATTER = 16, SPECULAR = 12 };
    enum Flag : int { IN_MEDIUM = 1, FROM_LIGHT = 2, VALID = 4, DELTA = 8 };
    int flags;
    int id; // material id on a surface, medium id in a medium
    Frame frame;
    float cont_prob;
    // surface
    f3 local_dir_fix;
    f3 geom_normal; // local
    float cos_fix_geom;
    float diff_prob, phong_prob, refl_prob, refr_prob, reflect_coeff, ior;
    // medium: incoming ray direction (world_dir_fix = -wdir, frame built from wdir) and the
    // scattering position (for the learned proposal's record lookup)
    f3 wdir, pos;

    UPBP_HD bool valid() const { return flags & VALID; }
    UPBP_HD bool delta() const { return flags & DELTA; }
    UPBP_HD bool in_medium() const { return flags & IN_MEDIUM; }
    UPBP_HD bool on_surface() const { return !in_medium(); }
    UPBP_HD bool from_light() const { return flags & FROM_LIGHT; }
    UPBP_HD float d_factor() const { return on_surface() ? fabsf(local_dir_fix.z) : 1.f; }
    UPBP_HD f3 world_dir_fix() const { return on_surface() ? frame.to_world(local_dir_fix) : -wdir; }
};

struct Hit {
    float dist;
    int mat, med, light;
    f3 normal, shading;
    bool enter;
    UPBP_HD bool in_medium() const { return mat < 0; }
    UPBP_HD bool on_surface() const { return mat >= 0; }
};

UPBP_HD void hit_reset(Hit& h, float d)
{
    h.dist = d;
    h.mat = h.med = h.light = -1;
    h.normal = mk(0.f);
    h.shading = mk(0.f);
    h.enter = false;
}

// kernel/bsdf.cuh:143-177
UPBP_HD void component_probabilities(Bsdf& b, const DMaterial& m)
{
    b.reflect_coeff = fresnel_dielectric(b.local_dir_fix.z, b.ior);
    const float ad = luminance(m.diffuse);
    const float ap = luminance(m.phong);
    const float ar = b.reflect_coeff * luminance(m.mirror);
    const float at = (1.f - b.reflect_coeff) * (m.ior > 0.f ? 1.f : 0.f);
    const float total = ad + ap + ar + at;
    if (total < 1e-9f) {
        b.diff_prob = b.phong_prob = b.refl_prob = b.refr_prob = 0.f;
        b.cont_prob = 0.f;
        return;
    }
    b.diff_prob = ad / total;
    b.phong_prob = ap / total;
    b.refl_prob = ar / total;
    b.refr_prob = at / total;
    const f3 tmp = m.diffuse + m.phong + b.reflect_coeff * m.mirror;
    b.cont_prob = maxc(tmp) + (1.f - b.reflect_coeff);
    b.cont_prob = fminf(1.f, fmaxf(0.f, b.cont_prob));
}

// kernel/bsdf.cuh:179-223
UPBP_HD void bsdf_setup(Bsdf& b, f3 ray_dir, const Hit& h, const DScene& sc, bool from_light, float ior)
{
    b.flags = from_light ? Bsdf::FROM_LIGHT : 0;
    b.diff_prob = b.phong_prob = b.refl_prob = b.refr_prob = b.reflect_coeff = 0.f;
    b.cos_fix_geom = 0.f;
    b.ior = -1.f;
    b.local_dir_fix = b.geom_normal = b.wdir = b.pos = mk(0.f);
    if (h.on_surface()) {
        const DMaterial& m = sc.mats[h.mat];
        b.id = h.mat;
        b.frame.set_from_z(h.shading);
        b.local_dir_fix = b.frame.to_local(-ray_dir);
        b.geom_normal = b.frame.to_local(h.normal);
        b.cos_fix_geom = dot(b.local_dir_fix, b.geom_normal);
        if (fabsf(b.local_dir_fix.z) < kEpsCosine || fabsf(b.cos_fix_geom) < kEpsCosine) {
            b.flags = 0;
            return;
        }
        b.ior = ior;
        component_probabilities(b, m);
        if (b.diff_prob == 0.f && b.phong_prob == 0.f) b.flags |= Bsdf::DELTA;
    } else {
        b.flags |= Bsdf::IN_MEDIUM;
        const int idx = int(fminf(float(h.med), float(sc.n_media - 1)));
        const DMedium& m = sc.media[idx];
        b.id = idx;
        b.wdir = ray_dir;
        b.frame.set_from_z(ray_dir); // from -world_dir_fix
        b.cont_prob = m.cont_prob;
    }
    b.flags |= Bsdf::VALID;
}

// Rebuilds a surface BSDF from the fields a light vertex stores (same values as at setup).
UPBP_HD void bsdf_restore_surface(Bsdf& b, f3 frame_src, f3 local_dir_fix, f3 geom_local, int mat, float diff_prob,
                                  float phong_prob, float cont_prob, int flags)
{
    b.flags = flags;
    b.id = mat;
    b.frame.set_from_z(frame_src);
    b.local_dir_fix = local_dir_fix;
    b.geom_normal = geom_local;
    b.cos_fix_geom = dot(local_dir_fix, geom_local);
    b.diff_prob = diff_prob;
    b.phong_prob = phong_prob;
    b.refl_prob = b.refr_prob = b.reflect_coeff = 0.f;
    b.ior = -1.f;
    b.cont_prob = cont_prob;
    b.wdir = mk(0.f);
}

UPBP_HD void bsdf_restore_medium(Bsdf& b, f3 ray_dir, int med, float cont_prob, int flags)
{
    b.flags = flags;
    b.id = med;
    b.wdir = ray_dir;
    b.frame.set_from_z(ray_dir);
    b.local_dir_fix = b.geom_normal = mk(0.f);
    b.cos_fix_geom = 0.f;
    b.diff_prob = b.phong_prob = b.refl_prob = b.refr_prob = b.reflect_coeff = 0.f;
    b.ior = -1.f;
    b.cont_prob = cont_prob;
    b.pos = mk(0.f);
}

// kernel/bsdf.cuh:227-284
UPBP_HD f3 eval_diffuse(const Bsdf& b, const DMaterial& m, f3 lg, float* fp, float* rp)
{
    if (b.diff_prob == 0) return mk(0.f);
    if (b.local_dir_fix.z < kEpsCosine || lg.z < kEpsCosine) return mk(0.f);
    if (fp) *fp += b.diff_prob * fmaxf(0.f, lg.z * kInvPi);
    if (rp) *rp += b.diff_prob * fmaxf(0.f, b.local_dir_fix.z * kInvPi);
    return m.diffuse * kInvPi;
}
UPBP_HD f3 eval_phong(const Bsdf& b, const DMaterial& m, f3 lg, float* fp, float* rp)
{
    if (b.phong_prob == 0) return mk(0.f);
    if (b.local_dir_fix.z < kEpsCosine || lg.z < kEpsCosine) return mk(0.f);
    const f3 r = reflect_local(b.local_dir_fix);
    const float dr = dot(r, lg);
    if (dr <= kEpsPhong) return mk(0.f);
    const float pw = powf(dr, m.phong_exp);
    if (!positive(pw)) return mk(0.f);
    if (fp || rp) {
        const float pdf = b.phong_prob * power_cos_hemisphere_pdf_w(r, lg, m.phong_exp);
        if (fp) *fp += pdf;
        if (rp) *rp += pdf;
    }
    const f3 rho = m.phong * (m.phong_exp + 2.f) * .5f * kInvPi;
    return rho * pw;
}
// kernel/bsdf.cuh:290-326
UPBP_HD void pdf_diffuse(const Bsdf& b, f3 lg, float* fp, float* rp)
{
    if (b.diff_prob == 0) return;
    if (fp) *fp += b.diff_prob * fmaxf(0.f, lg.z * kInvPi);
    if (rp) *rp += b.diff_prob * fmaxf(0.f, b.local_dir_fix.z * kInvPi);
}
UPBP_HD void pdf_phong(const Bsdf& b, const DMaterial& m, f3 lg, float* fp, float* rp)
{
    if (b.phong_prob == 0) return;
    const f3 r = reflect_local(b.local_dir_fix);
    const float dr = dot(r, lg);
    if (dr <= kEpsPhong) return;
    const float pdf = power_cos_hemisphere_pdf_w(r, lg, m.phong_exp) * b.phong_prob;
    if (fp) *fp += pdf;
    if (rp) *rp += pdf;
}

// world shading normal on the side of local cosine c (the side a camera path leaves to)
UPBP_HD f3 surface_side(const Bsdf& b, float c) { return c < 0.f ? -b.frame.z : b.frame.z; }

// kernel/bsdf.cuh:334-396
UPBP_HD f3 bsdf_evaluate(const Bsdf& b, f3 dir_gen, const DScene& sc, float& cos_gen, float* for_pdf, float* rev_pdf,
                         float* sin_theta)
{
    if (b.on_surface()) {
        f3 res = mk(0.f);
        if (for_pdf) *for_pdf = 0.f;
        if (rev_pdf) *rev_pdf = 0.f;
        if (sin_theta) *sin_theta = 0.f;
        const f3 lg = b.frame.to_local(dir_gen);
        const float cg = lg.z;
        const float cgg = dot(lg, b.geom_normal);
        cos_gen = fabsf(cg);
        if (cos_gen < kEpsCosine || fabsf(cgg) < kEpsCosine) return res;
        if (cg * b.local_dir_fix.z < 0.f || cgg * b.cos_fix_geom < 0.f) return res;
        const DMaterial& m = sc.mats[b.id];
        res += eval_diffuse(b, m, lg, for_pdf, rev_pdf);
        res += eval_phong(b, m, lg, for_pdf, rev_pdf);
        // the camera technique samples the learned product at guidable surfaces: its pdf is the
        // forward one at a camera vertex and the reverse one at a light vertex
        const float a = sc.guide.alpha_s;
        if (a > 0.f && surface_guidable(m)) {
            if (!b.from_light() && for_pdf)
                *for_pdf = a * guide_surface_pdf(sc.guide, b.pos, surface_side(b, b.local_dir_fix.z), dir_gen) +
                           (1.f - a) * *for_pdf;
            else if (b.from_light() && rev_pdf)
                *rev_pdf = a * guide_surface_pdf(sc.guide, b.pos, surface_side(b, cg), b.world_dir_fix()) +
                           (1.f - a) * *rev_pdf;
        }
        return res;
    }
    cos_gen = 1.f;
    const DMedium& m = sc.media[b.id];
    if (black_or_neg(m.scattering)) {
        if (for_pdf) *for_pdf = 0.f;
        if (rev_pdf) *rev_pdf = 0.f;
        if (sin_theta) *sin_theta = 0.f;
        return mk(0.f);
    }
    const float p = hg_pdf(-b.wdir, dir_gen, m.mean_cosine, sin_theta);
    // the camera technique samples the learned mixture: its pdf is the forward one at a camera
    // vertex and the reverse one at a light vertex (arriving along -dir_gen, leaving along -wdir)
    float fp = p, rp = p;
    if (sc.guide.alpha > 0.f) {
        const float a = sc.guide.alpha;
        if (!b.from_light()) fp = a * guide_pdf(sc.guide, m.mean_cosine, b.pos, b.wdir, dir_gen) + (1.f - a) * p;
        else rp = a * guide_pdf(sc.guide, m.mean_cosine, b.pos, -dir_gen, -b.wdir) + (1.f - a) * p;
    }
    if (for_pdf) *for_pdf = fp;
    if (rev_pdf) *rev_pdf = rp;
    return m.scattering * p;
}

// kernel/bsdf.cuh:402-447
UPBP_HD float bsdf_pdf_rev(const Bsdf& b, f3 dir_gen, const DScene& sc)
{
    if (b.on_surface()) {
        const f3 lg = b.frame.to_local(dir_gen);
        const float cg = lg.z;
        const float cgg = dot(lg, b.geom_normal);
        if (fabsf(cg) < kEpsCosine || fabsf(cgg) < kEpsCosine) return 0.f;
        if (cg * b.local_dir_fix.z < 0.f || cgg * b.cos_fix_geom < 0.f) return 0.f;
        float fp = 0.f, rp = 0.f;
        const DMaterial& m = sc.mats[b.id];
        pdf_diffuse(b, lg, &fp, &rp);
        pdf_phong(b, m, lg, &fp, &rp);
        const float a = sc.guide.alpha_s;
        if (a > 0.f && b.from_light() && surface_guidable(m)) // the camera technique's side
            rp = a * guide_surface_pdf(sc.guide, b.pos, surface_side(b, cg), b.world_dir_fix()) + (1.f - a) * rp;
        return rp;
    }
    const float g = sc.media[b.id].mean_cosine;
    const float p = hg_pdf(-b.wdir, dir_gen, g, nullptr);
    if (!(sc.guide.alpha > 0.f) || !b.from_light()) return p; // light technique: the phase function
    const float a = sc.guide.alpha;
    return a * guide_pdf(sc.guide, g, b.pos, -dir_gen, -b.wdir) + (1.f - a) * p;
}

// kernel/bsdf.cuh:567-660 with the component samplers of :455-559, surface part
UPBP_HD f3 bsdf_sample_surface(const Bsdf& b, float u0, float u1, float u2, f3& dir_gen, float& pdf, float& cos_gen,
                               const DScene& sc, unsigned& event, float& sin_theta)
{
    {
        unsigned ev;
        if (u2 < b.diff_prob) ev = Bsdf::DIFFUSE;
        else if (u2 < b.diff_prob + b.phong_prob) ev = Bsdf::PHONG;
        else if (u2 < b.diff_prob + b.phong_prob + b.refl_prob) ev = Bsdf::REFLECT;
        else ev = Bsdf::REFRACT;
        event = ev;
        sin_theta = 0.f;
        pdf = 0.f;
        f3 res = mk(0.f);
        f3 lg = mk(0.f);
        const DMaterial& m = sc.mats[b.id];
        if (ev == Bsdf::DIFFUSE) {
            if (b.local_dir_fix.z < kEpsCosine) return mk(0.f);
            float up;
            lg = sample_cos_hemisphere_w(u0, u1, &up);
            pdf += up * b.diff_prob;
            res += m.diffuse * kInvPi;
            if (black_or_neg(res)) return mk(0.f);
            res += eval_phong(b, m, lg, &pdf, nullptr);
        } else if (ev == Bsdf::PHONG) {
            lg = sample_power_cos_hemisphere_w(u0, u1, m.phong_exp);
            const f3 r = reflect_local(b.local_dir_fix);
            {
                Frame f;
                f.set_from_z(r);
                lg = f.to_world(lg);
            }
            const float dr = dot(r, lg);
            if (dr <= kEpsPhong) return mk(0.f);
            pdf_phong(b, m, lg, &pdf, nullptr);
            const f3 rho = m.phong * (m.phong_exp + 2.f) * .5f * kInvPi;
            res += rho * powf(dr, m.phong_exp);
            if (black_or_neg(res)) return mk(0.f);
            res += eval_diffuse(b, m, lg, &pdf, nullptr);
        } else if (ev == Bsdf::REFLECT) {
            lg = reflect_local(b.local_dir_fix);
            pdf += b.refl_prob;
            res += b.reflect_coeff * m.mirror / fabsf(lg.z);
            if (black_or_neg(res)) return mk(0.f);
        } else {
            if (b.ior < 0) return mk(0.f);
            float cos_i = b.local_dir_fix.z;
            float cos_t;
            const float eta = b.ior;
            if (cos_i < 0.f) {
                cos_i = -cos_i;
                cos_t = 1.f;
            } else {
                cos_t = -1.f;
            }
            const float sin_t2 = eta * eta * (1.f - cos_i * cos_i);
            if (!(sin_t2 < 1.f)) return mk(0.f);
            cos_t *= sqrtf(fmaxf(0.f, 1.f - sin_t2));
            lg = {-eta * b.local_dir_fix.x, -eta * b.local_dir_fix.y, cos_t};
            pdf += b.refr_prob;
            const float rc = 1.f - b.reflect_coeff;
            res += b.from_light() ? mk(rc / fabsf(cos_t)) : mk(rc * eta * eta / fabsf(cos_t));
            if (black_or_neg(res)) return mk(0.f);
        }
        dir_gen = b.frame.to_world(lg);
        lg = b.frame.to_local(dir_gen);
        const float cg = lg.z;
        const float cgg = dot(lg, b.geom_normal);
        cos_gen = fabsf(cg);
        if (cos_gen < kEpsCosine || fabsf(cgg) < kEpsCosine) return mk(0.f);
        const bool refr = ev == Bsdf::REFRACT;
        if ((refr && cg * b.local_dir_fix.z > 0.f) || (!refr && cg * b.local_dir_fix.z < 0.f) ||
            (refr && cgg * b.cos_fix_geom > 0.f) || (!refr && cgg * b.cos_fix_geom < 0.f))
            return mk(0.f);
        return res;
    }
}

// kernel/bsdf.cuh:567-660
UPBP_HD f3 bsdf_sample(const Bsdf& b, float u0, float u1, float u2, f3& dir_gen, float& pdf, float& cos_gen,
                       const DScene& sc, unsigned& event, float& sin_theta)
{
    if (b.on_surface()) {
        const float a = sc.guide.alpha_s;
        if (!(a > 0.f) || b.from_light() || !surface_guidable(sc.mats[b.id]))
            return bsdf_sample_surface(b, u0, u1, u2, dir_gen, pdf, cos_gen, sc, event, sin_theta);
        // camera path at a guidable surface: the learned product with the cosine lobe
        // (probability alpha_s, chosen by the component variate u2, which is then rescaled for
        // the BSDF's own component choice), else the BSDF; the pdf is the mixture's either way
        const f3 n = surface_side(b, b.local_dir_fix.z);
        if (u2 < a) {
            dir_gen = vmf_prod_sample(guide_record(sc.guide, b.pos), cosine_lobes(n), u2 / a, u0, u1);
            event = Bsdf::DIFFUSE;
            sin_theta = 0.f;
            pdf = 0.f;
            return bsdf_evaluate(b, dir_gen, sc, cos_gen, &pdf, nullptr, nullptr);
        }
        const f3 f = bsdf_sample_surface(b, u0, u1, (u2 - a) / (1.f - a), dir_gen, pdf, cos_gen, sc, event, sin_theta);
        if (black_or_neg(f)) return f;
        pdf = a * guide_surface_pdf(sc.guide, b.pos, n, dir_gen) + (1.f - a) * pdf;
        return f;
    }
    cos_gen = 1.f;
    event = Bsdf::SCATTER;
    const DMedium& m = sc.media[b.id];
    if (black_or_neg(m.scattering)) {
        dir_gen = mk(0.f);
        pdf = 0.f;
        sin_theta = 0.f;
        return mk(0.f);
    }
    const DGuide& g = sc.guide;
    if (!(g.alpha > 0.f) || b.from_light()) {
        dir_gen = hg_sample(-b.wdir, m.mean_cosine, u0, u1, b.frame, pdf, &sin_theta);
        return m.scattering * pdf;
    }
    // camera path: one-sample mixture of the learned product (probability alpha, chosen by the
    // otherwise unused component variate u2) and the phase function
    const float4* rec = guide_record(g, b.pos);
    const PhaseLobes h = phase_lobes(b.wdir, m.mean_cosine);
    float hg;
    if (u2 < g.alpha) {
        dir_gen = vmf_prod_sample(rec, h, u2 / g.alpha, u0, u1);
        hg = hg_pdf(-b.wdir, dir_gen, m.mean_cosine, &sin_theta);
    } else {
        dir_gen = hg_sample(-b.wdir, m.mean_cosine, u0, u1, b.frame, hg, &sin_theta);
    }
    pdf = g.alpha * vmf_prod_pdf(rec, h, dir_gen) + (1.f - g.alpha) * hg;
    return m.scattering * hg;
}

// ---- lights: kernel/light.cuh ----------------------------------------------------
UPBP_HD bool light_finite(const DLight& l) { return l.type == L_AREA || l.type == L_POINT; }
UPBP_HD bool light_delta(const DLight& l) { return l.type == L_DIRECTIONAL || l.type == L_POINT; }

UPBP_HD f3 light_illuminate(const DLight& l, const DScene& sc, f3 recv, float u0, float u1, f3& dir, float& dist,
                            float& for_pdf_w, float* emission_pdf_w, float* cos_at_light)
{
    if (l.type == L_AREA) { // :43-72
        const float t = sqrtf(u0);
        const float b0 = 1.f - t, b1 = u1 * t;
        const f3 p = l.p0 + l.e1 * b0 + l.e2 * b1;
        dir = p - recv;
        const float d2 = dot(dir, dir);
        dist = sqrtf(d2);
        dir = dir / dist;
        const float cn = dot(l.frame.z, -dir);
        if (cn < kEpsCosine) return mk(0.f);
        for_pdf_w = l.inv_area * d2 / cn;
        if (cos_at_light) *cos_at_light = cn;
        if (emission_pdf_w) *emission_pdf_w = l.inv_area * cn * kInvPi;
        return l.intensity;
    }
    if (l.type == L_BACKGROUND) { // :128-151
        f3 rad;
        if (l.uses_env && sc.env.img) {
            dir = env_sample(sc.env, u0, u1, for_pdf_w, rad);
            rad = rad * l.scale;
        } else {
            dir = sample_uniform_sphere_w(u0, u1, &for_pdf_w);
            rad = l.intensity * l.scale;
        }
        dist = kInf;
        if (emission_pdf_w) *emission_pdf_w = for_pdf_w * concentric_disc_pdf_a() * sc.inv_radius_sqr;
        if (cos_at_light) *cos_at_light = 1.f;
        return rad;
    }
    if (l.type == L_DIRECTIONAL) { // :230-248
        dir = -l.frame.z;
        dist = kInf;
        for_pdf_w = 1.f;
        if (cos_at_light) *cos_at_light = 1.f;
        if (emission_pdf_w) *emission_pdf_w = concentric_disc_pdf_a() * sc.inv_radius_sqr;
        return l.intensity;
    }
    dir = l.position - recv; // point, :301-317
    const float d2 = dot(dir, dir);
    for_pdf_w = d2;
    dist = sqrtf(d2);
    dir = dir / dist;
    if (cos_at_light) *cos_at_light = 1.f;
    if (emission_pdf_w) *emission_pdf_w = uniform_sphere_pdf_w();
    return l.intensity;
}

UPBP_HD f3 light_emit(const DLight& l, const DScene& sc, float d0, float d1, float p0, float p1, f3& pos, f3& dir,
                      float& emission_pdf_w, float* for_pdf_a, float* cos_light)
{
    if (l.type == L_AREA) { // :74-99
        const float t = sqrtf(p0);
        const float b0 = 1.f - t, b1 = p1 * t;
        pos = l.p0 + l.e1 * b0 + l.e2 * b1;
        f3 lo = sample_cos_hemisphere_w(d0, d1, &emission_pdf_w);
        emission_pdf_w *= l.inv_area;
        lo.z = fmaxf(lo.z, kEpsCosine);
        dir = l.frame.to_world(lo);
        if (for_pdf_a) *for_pdf_a = l.inv_area;
        if (cos_light) *cos_light = lo.z;
        return l.intensity * lo.z;
    }
    if (l.type == L_BACKGROUND) { // :153-190
        float fp;
        f3 rad;
        if (l.uses_env && sc.env.img) {
            dir = -env_sample(sc.env, d0, d1, fp, rad);
            rad = rad * l.scale;
        } else {
            dir = sample_uniform_sphere_w(d0, d1, &fp);
            rad = l.intensity * l.scale;
        }
        float x, y;
        sample_concentric_disc(p0, p1, x, y);
        Frame f;
        f.set_from_z(dir);
        pos = sc.center + sc.radius * (-dir + f.x * x + f.y * y);
        emission_pdf_w = fp * concentric_disc_pdf_a() * sc.inv_radius_sqr;
        if (for_pdf_a) *for_pdf_a = fp;
        if (cos_light) *cos_light = 1.f;
        return rad;
    }
    if (l.type == L_DIRECTIONAL) { // :250-272
        float x, y;
        sample_concentric_disc(p0, p1, x, y);
        pos = sc.center + sc.radius * (-l.frame.z + l.frame.x * x + l.frame.y * y);
        dir = l.frame.z;
        emission_pdf_w = concentric_disc_pdf_a() * sc.inv_radius_sqr;
        if (for_pdf_a) *for_pdf_a = 1.f;
        if (cos_light) *cos_light = 1.f;
        return l.intensity;
    }
    pos = l.position; // point, :319-337
    dir = sample_uniform_sphere_w(d0, d1, &emission_pdf_w);
    if (for_pdf_a) *for_pdf_a = 1.f;
    if (cos_light) *cos_light = 1.f;
    return l.intensity;
}

UPBP_HD f3 light_radiance(const DLight& l, const DScene& sc, f3 ray_dir, float* for_pdf_a, float* emission_pdf_w)
{
    if (l.type == L_AREA) { // :101-122
        const float c = fmaxf(0.f, dot(l.frame.z, -ray_dir));
        if (c == 0.f) return mk(0.f);
        if (for_pdf_a) *for_pdf_a = l.inv_area;
        if (emission_pdf_w) *emission_pdf_w = cos_hemisphere_pdf_w(l.frame.z, -ray_dir) * l.inv_area;
        return l.intensity;
    }
    if (l.type == L_BACKGROUND) { // :192-217
        float fp;
        f3 rad;
        if (l.uses_env && sc.env.img) {
            rad = env_lookup(sc.env, ray_dir, &fp) * l.scale;
        } else {
            fp = uniform_sphere_pdf_w();
            rad = l.intensity * l.scale;
        }
        if (for_pdf_a) *for_pdf_a = fp;
        if (emission_pdf_w) *emission_pdf_w = fp * concentric_disc_pdf_a() * sc.inv_radius_sqr;
        return rad;
    }
    return mk(0.f);
}

// raster position of a world point (kernel/camera.cuh:42-46 semantics)
UPBP_HD void world_to_raster(const DCamera& c, f3 p, float& x, float& y)
{
    const f3 d = p - c.origin;
    const float depth = dot(c.forward, d);
    x = c.res_x * .5f + c.ipd * dot(c.right, d) / depth;
    y = c.res_y * .5f - c.ipd * dot(c.up, d) / depth;
}

UPBP_HD f3 raster_dir(const DCamera& c, float x, float y)
{
    return normalize(c.forward * c.ipd + c.right * (x - c.res_x * .5f) - c.up * (y - c.res_y * .5f));
}

} // namespace upbp
