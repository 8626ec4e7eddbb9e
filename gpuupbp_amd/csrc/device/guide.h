// Learned scattering-direction proposal for camera paths in media: the product of a learned
// incident-radiance mixture of eight von Mises-Fisher lobes with three vMF lobes standing in
// for the Henyey-Greenstein phase function, one-sample mixed with the phase function itself. No
// counterpart in the reference (gpuupbp samples the phase function only,
// gpuupbp/kernel/phase_function.cuh:34-137).
//
// The radiance mixture is piecewise constant over a uniform grid on the scene bounds: each cell
// holds the lobes the network (guide_mlp.hip) predicts at its centre. The phase lobes depend on
// the arrival direction at the vertex and are multiplied in at the vertex, in closed form (the
// product of vMF lobes is a vMF lobe), so that the proposal follows a sharp phase function
// (g = 0.9: most of its mass within a few degrees) for every arrival direction, which no record
// binned by direction can. This is the product sampling of Herholz et al., "Volume Path Guiding
// Based on Zero-Variance Random Walk Theory" (2019); the phase function's three lobes are the
// KL-closest fit tabulated over g (phase_lobes.h, made by fit_phase_lobes.py): one lobe of
// matching mean cosine leaves E_HG[HG / lobe] near 1000 at g = 0.9, three leave it at 1.01. Any
// lobes give an exact pdf; the fit only decides how closely the proposal follows the phase
// function. The pdf of the camera technique at any point is one record lookup and the product,
// so the forward and reverse MIS recursions, the connections and every merge evaluate the same
// exact pdf the camera paths sample from.
//
// Record layout (kGuideRec float4, 160 bytes): lobes i = 0..7 as (mu.xyz, kappa), then the lobe
// constants c_i = w_i kappa_i / (2 pi (1 - exp(-2 kappa_i))) (two float4) with the normalised
// lobe weights w_i, so that the radiance mixture is q(w) = sum_i c_i exp(kappa_i (mu_i . w - 1)).
//
// Phase lobes j = 0..2 about mu_h: h(w) = sum_j a_j exp(k_j (mu_h . w - 1)). Lobe (i, j) of the
// product has v_ij = kappa_i mu_i + k_j mu_h, and the normalised product has the pdf
//   p(w) = q(w) h(w) / Z,
//   Z = sum_ij c_i a_j 2 pi (exp(|v_ij| - kappa_i - k_j) - exp(-|v_ij| - kappa_i - k_j)) / |v_ij|,
// its lobe (i, j) chosen with probability (term ij of Z) / Z. The second exponential of a term
// is below 1e-7 of the first once |v_ij| > 8 and is then left out.
#pragma once

#include "phase_lobes.h"
#include "upbp_device.h"

namespace upbp {

UPBP_HD int guide_cell(const DGuide& g, f3 p)
{
    const float fx = (p.x - g.lo.x) * g.inv_cell.x, fy = (p.y - g.lo.y) * g.inv_cell.y,
                fz = (p.z - g.lo.z) * g.inv_cell.z;
    const int r = g.res;
    const int ix = fx < 0.f ? 0 : (fx >= float(r) ? r - 1 : int(fx));
    const int iy = fy < 0.f ? 0 : (fy >= float(r) ? r - 1 : int(fy));
    const int iz = fz < 0.f ? 0 : (fz >= float(r) ? r - 1 : int(fz));
    return (iz * r + iy) * r + ix;
}

UPBP_HD const float4* guide_record(const DGuide& g, f3 p) { return g.recs + size_t(guide_cell(g, p)) * kGuideRec; }

UPBP_HD float lane4(const float4& v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w)); }

// 2 pi (1 - exp(-2 k)) / k, the integral of exp(k (mu . w - 1)) over the sphere
UPBP_HD float vmf_mass(float k) { return k > 1e-6f ? 2.f * kPi * -expm1f(-2.f * k) / k : 4.f * kPi; }

// exp(kp - s) vmf_mass(kp): the sphere integral of exp(kp (mu . w - 1) + kp - s)
UPBP_HD float prod_mass(float kp, float s)
{
    if (kp > 8.f) return 2.f * kPi * expf(kp - s) / kp;
    return expf(kp - s) * vmf_mass(kp);
}

// the phase function of mean cosine g, for a path travelling along `arrive`, as three vMF lobes
// about mu (along the arrival for g >= 0, against it for g < 0); the table's rows are
// interpolated linearly in |g| (clamped to 0.99)
struct PhaseLobes {
    f3 mu;
    int n; // lobes in use
    float k[kPhaseLobes], a[kPhaseLobes];
};

UPBP_HD PhaseLobes phase_lobes(f3 arrive, float g)
{
    PhaseLobes h;
    h.mu = g < 0.f ? -arrive : arrive;
    h.n = kPhaseLobes;
    const float t = fminf(fabsf(g), 0.99f) / kPhaseFitStep;
    const int i = t >= float(kPhaseFitN - 1) ? kPhaseFitN - 2 : int(t);
    const float f = t - float(i);
    for (int j = 0; j < kPhaseLobes; j++) {
        h.k[j] = (1.f - f) * kPhaseFit[i][j] + f * kPhaseFit[i + 1][j];
        const float w = (1.f - f) * kPhaseFit[i][3 + j] + f * kPhaseFit[i + 1][3 + j];
        h.a[j] = w / vmf_mass(h.k[j]);
    }
    return h;
}

// the clamped cosine about the surface normal n as its one fitted vMF lobe (phase_lobes.h)
UPBP_HD PhaseLobes cosine_lobes(f3 n)
{
    PhaseLobes h;
    h.mu = n;
    h.n = 1;
    h.k[0] = kCosineKappa;
    h.a[0] = 1.f / vmf_mass(kCosineKappa);
    for (int j = 1; j < kPhaseLobes; j++) h.k[j] = h.a[j] = 0.f;
    return h;
}

// out of line: keeps the product's record loads and exponentials out of the register
// allocation of the bounce and merge kernels, which call them only when alpha > 0
inline __host__ __device__ __noinline__ float vmf_prod_pdf(const float4* rec, PhaseLobes h, f3 w)
{
    float q = 0.f, z = 0.f, hv = 0.f;
    const float hw = dot(h.mu, w);
    for (int j = 0; j < h.n; j++) hv += h.a[j] * expf(h.k[j] * (hw - 1.f));
#pragma unroll
    for (int i = 0; i < kGuideLobes; i++) {
        const float4 l = rec[i];
        const float c = lane4(rec[kGuideLobes + i / 4], i & 3);
        const float lw = l.x * w.x + l.y * w.y + l.z * w.z, lh = l.x * h.mu.x + l.y * h.mu.y + l.z * h.mu.z;
        q += c * expf(l.w * (lw - 1.f));
        float zi = 0.f;
        for (int j = 0; j < h.n; j++) {
            // |v|^2 = kappa^2 + k^2 + 2 kappa k (mu_i . mu_h)
            const float kp = sqrtf(fmaxf(0.f, l.w * l.w + h.k[j] * h.k[j] + 2.f * l.w * h.k[j] * lh));
            zi += h.a[j] * prod_mass(kp, l.w + h.k[j]);
        }
        z += c * zi;
    }
    return z > 0.f ? q * hv / z : 0.f;
}

// direction of the vMF lobe (mu, kappa) by the inversion of Jakob, "Numerically stable sampling
// of the von Mises Fisher distribution on S2" (2012), in the form
// cos = 1 + log1p((1 - u0) expm1(-2 kappa)) / kappa that stays exact as kappa -> 0
UPBP_HD f3 vmf_sample_dir(f3 mu, float kappa, float u0, float u1)
{
    const float c = fmaxf(-1.f, 1.f + log1pf((1.f - u0) * expm1f(-2.f * kappa)) / kappa);
    const float s = sqrtf(fmaxf(0.f, 1.f - c * c));
    const float phi = 2.f * kPi * u1;
    Frame f;
    f.set_from_z(mu);
    return f.to_world(mk(cosf(phi) * s, sinf(phi) * s, c));
}

// lobe (i, j) of the product chosen with us in [0, 1), its direction with u0, u1
inline __host__ __device__ __noinline__ f3 vmf_prod_sample(const float4* rec, PhaseLobes h, float us, float u0,
                                                           float u1)
{
    float tot = 0.f;
#pragma unroll
    for (int i = 0; i < kGuideLobes; i++) {
        const float4 l = rec[i];
        const float c = lane4(rec[kGuideLobes + i / 4], i & 3);
        const float lh = l.x * h.mu.x + l.y * h.mu.y + l.z * h.mu.z;
        for (int j = 0; j < h.n; j++) {
            const float kp = sqrtf(fmaxf(0.f, l.w * l.w + h.k[j] * h.k[j] + 2.f * l.w * h.k[j] * lh));
            tot += c * h.a[j] * prod_mass(kp, l.w + h.k[j]);
        }
    }
    const float t = us * tot;
    int pick = -1;
    float acc = 0.f;
    for (int i = 0; i < kGuideLobes; i++) {
        const float4 l = rec[i];
        const float c = lane4(rec[kGuideLobes + i / 4], i & 3);
        const float lh = l.x * h.mu.x + l.y * h.mu.y + l.z * h.mu.z;
        for (int j = 0; j < h.n; j++) {
            const float kp = sqrtf(fmaxf(0.f, l.w * l.w + h.k[j] * h.k[j] + 2.f * l.w * h.k[j] * lh));
            acc += c * h.a[j] * prod_mass(kp, l.w + h.k[j]);
            if (t < acc && pick < 0) pick = i * kPhaseLobes + j;
        }
    }
    if (pick < 0) pick = (kGuideLobes - 1) * kPhaseLobes + h.n - 1;
    const int pi = pick / kPhaseLobes, pj = pick % kPhaseLobes;
    const float4 l = rec[pi];
    const f3 v = mk(l.x * l.w + h.mu.x * h.k[pj], l.y * l.w + h.mu.y * h.k[pj], l.z * l.w + h.mu.z * h.k[pj]);
    const float kp = sqrtf(dot(v, v));
    if (!(kp > 1e-6f)) return vmf_sample_dir(mk(0.f, 0.f, 1.f), 1e-6f, u0, u1);
    return vmf_sample_dir(v / kp, kp, u0, u1);
}

// pdf of the learned part of the proposal at p, for a path arriving along `arrive` in a medium
// of mean cosine g and leaving along gen
UPBP_HD float guide_pdf(const DGuide& gd, float g, f3 p, f3 arrive, f3 gen)
{
    return vmf_prod_pdf(guide_record(gd, p), phase_lobes(arrive, g), gen);
}

// Surfaces that take the learned proposal: no mirror and no refraction component, so that the
// camera path's BSDF sample is always a non-specular reflection (a material property: the
// same at camera vertices and at stored light vertices)
UPBP_HD bool surface_guidable(const DMaterial& m) { return luminance(m.mirror) == 0.f && !(m.ior > 0.f); }

// pdf of the learned part of the proposal at surface point p with the normal n on the side the
// camera path leaves to, along gen
UPBP_HD float guide_surface_pdf(const DGuide& gd, f3 p, f3 n, f3 gen)
{
    return vmf_prod_pdf(guide_record(gd, p), cosine_lobes(n), gen);
}

// Camera technique's solid-angle pdf of leaving p along gen after arriving along arrive, in
// medium m (continuation probability excluded).
UPBP_HD float cam_medium_pdf(const DScene& sc, const DMedium& m, f3 p, f3 arrive, f3 gen, float* sin_theta)
{
    const float hg = hg_pdf(-arrive, gen, m.mean_cosine, sin_theta);
    const DGuide& g = sc.guide;
    if (!(g.alpha > 0.f)) return hg;
    return g.alpha * guide_pdf(g, m.mean_cosine, p, arrive, gen) + (1.f - g.alpha) * hg;
}

// (mu, kappa, w) x 8 -> the device record: lobes, normalised weights, lobe constants
UPBP_HD void pack_guide_record(const float* in, float4* out)
{
    float wsum = 0.f;
    for (int i = 0; i < kGuideLobes; i++) wsum += fmaxf(in[5 * i + 4], 0.f);
    float w[kGuideLobes], c[kGuideLobes];
    for (int i = 0; i < kGuideLobes; i++) {
        const float* l = in + 5 * i;
        const float n = sqrtf(l[0] * l[0] + l[1] * l[1] + l[2] * l[2]);
        const float kappa = fminf(fmaxf(l[3], kGuideKappaMin), kGuideKappaMax);
        out[i] = n > 0.f ? make_float4(l[0] / n, l[1] / n, l[2] / n, kappa) : make_float4(0.f, 0.f, 1.f, kappa);
        w[i] = wsum > 0.f ? fmaxf(l[4], 0.f) / wsum : 1.f / float(kGuideLobes);
        c[i] = w[i] * kappa / (2.f * kPi * -expm1f(-2.f * kappa));
    }
    out[8] = make_float4(c[0], c[1], c[2], c[3]);
    out[9] = make_float4(c[4], c[5], c[6], c[7]);
}

} // namespace upbp
