// Learned scattering-direction proposal for camera paths in media: a mixture of eight
// von Mises-Fisher lobes, one-sample mixed with the Henyey-Greenstein phase function. No
// counterpart in the reference (gpuupbp samples the phase function only,
// gpuupbp/kernel/phase_function.cuh:34-137).
//
// The proposal is piecewise constant: a uniform grid over the scene bounds, and per cell six
// bins of the camera path's arrival direction (the dominant axis and its sign). Each
// (cell, bin) record holds the lobes the network (guide_mlp.hip) predicts at the cell centre
// for the bin's axis direction. The pdf of the camera technique at any point is therefore
// one record lookup, so the forward and reverse MIS recursions, the connections and every
// merge evaluate the same exact mixture pdf the camera paths sample from.
//
// Record layout (kGuideRec float4): lobes i = 0..7 as (mu.xyz, kappa); the lobe weights w_i
// (two float4); the lobe constants c_i = w_i kappa_i / (2 pi (1 - exp(-2 kappa_i))) (two
// float4), so that the mixture pdf is sum_i c_i exp(kappa_i (mu_i . w - 1)).
#pragma once

#include "upbp_device.h"

namespace upbp {

UPBP_HD int guide_bin(f3 d)
{
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    if (ax >= ay && ax >= az) return d.x < 0.f ? 1 : 0;
    if (ay >= az) return d.y < 0.f ? 3 : 2;
    return d.z < 0.f ? 5 : 4;
}

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

UPBP_HD const float4* guide_record(const DGuide& g, f3 p, f3 arrive)
{
    return g.recs + size_t(guide_cell(g, p) * kGuideBins + guide_bin(arrive)) * kGuideRec;
}

UPBP_HD float lane4(const float4& v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w)); }

// sum_i c_i exp(kappa_i (mu_i . w - 1))
// out of line: keeps the mixture's twelve record loads and eight exponentials out of the
// register allocation of the bounce and merge kernels, which call it only when alpha > 0
inline __host__ __device__ __noinline__ float vmf_mix_pdf(const float4* rec, f3 w)
{
    float pdf = 0.f;
#pragma unroll
    for (int i = 0; i < kGuideLobes; i++) {
        const float4 l = rec[i];
        const float c = lane4(rec[kGuideLobes + 2 + i / 4], i & 3);
        pdf += c * expf(l.w * (l.x * w.x + l.y * w.y + l.z * w.z - 1.f));
    }
    return pdf;
}

// lobe chosen by weight with us in [0, 1), direction by the inversion of Jakob ("Numerically
// stable sampling of the von Mises Fisher distribution on S2", 2012) with u0, u1
inline __host__ __device__ __noinline__ f3 vmf_mix_sample(const float4* rec, float us, float u0, float u1)
{
    int pick = -1;
    float acc = 0.f;
#pragma unroll
    for (int i = 0; i < kGuideLobes; i++) {
        acc += lane4(rec[kGuideLobes + i / 4], i & 3);
        if (pick < 0 && us < acc) pick = i;
    }
    const float4 l = rec[pick < 0 ? kGuideLobes - 1 : pick];
    const float kappa = l.w;
    const float c = 1.f + logf(u0 + (1.f - u0) * expf(-2.f * kappa)) / kappa;
    const float s = sqrtf(fmaxf(0.f, 1.f - c * c));
    const float phi = 2.f * kPi * u1;
    Frame f;
    f.set_from_z(mk(l.x, l.y, l.z));
    return f.to_world(mk(cosf(phi) * s, sinf(phi) * s, c));
}

// Camera technique's solid-angle pdf of leaving p along gen after arriving along arrive, in
// medium m (continuation probability excluded).
UPBP_HD float cam_medium_pdf(const DScene& sc, const DMedium& m, f3 p, f3 arrive, f3 gen, float* sin_theta)
{
    const float hg = hg_pdf(-arrive, gen, m.mean_cosine, sin_theta);
    const DGuide& g = sc.guide;
    if (!(g.alpha > 0.f)) return hg;
    return g.alpha * vmf_mix_pdf(guide_record(g, p, arrive), gen) + (1.f - g.alpha) * hg;
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
    out[8] = make_float4(w[0], w[1], w[2], w[3]);
    out[9] = make_float4(w[4], w[5], w[6], w[7]);
    out[10] = make_float4(c[0], c[1], c[2], c[3]);
    out[11] = make_float4(c[4], c[5], c[6], c[7]);
}

} // namespace upbp
