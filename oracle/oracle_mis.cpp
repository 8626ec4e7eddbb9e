// FP64 full-path MIS oracle (test infrastructure only): explicit balance-heuristic
// weights from per-vertex sampling densities. See oracle_mis.h for the definitions.
//
// Path pdfs in area / volume measure, x_0 on the light, x_k the camera:
//   light direction   P>_0 = position pdf of x_0,
//                     P>_1 = emission pdf of the direction * rsf_0 * cin_1 / d_0^2,
//                     P>_j = fpdf_{j-1} * rsf_{j-1} * cin_j / d_{j-1}^2,
//   camera direction  P<_k = 1,
//                     P<_{k-1} = camera pdf * rsr_{k-1} * cout_{k-1} / d_{k-1}^2,
//                     P<_j = rpdf_{j+1} * rsr_j * cout_j / d_j^2,
//                     P<_0 = rpdf_1 * rsr_0 * cos_light / d_0^2.
// Techniques and their pdfs:
//   BPT with s light vertices (0 <= s <= k-1): n_bpt * prod_{i<s} P>_i * prod_{i>=s} P<_i,
//     valid when x_{s-1} and x_s can be connected (s = 0: the camera path hits the light);
//   light tracing (s = k): n_bpt * N_L * prod_{i<k} P>_i, since every light path reaches
//     the image;
//   merging at x_m (1 <= m <= k-1): prod_{i<=m} P>_i * prod_{i>=m} P<_i * F_m with
//     F_m = mf_surf on surfaces and mf_pp3d + mf_pb2d rp + mf_bp2d fp + mf_bb1d sin fp rp in
//     media; fp = 1/rsf_{m-1} (long photon beams) or the extinction ratio, rp = 1/rsr_m
//     (long query beams) or the ratio, the merge factors mf including N_L.
#include "oracle_mis.h"

#include <cmath>

namespace orc {

double mis_explicit_weight(const MisRec& r, const MisParams& p)
{
    const int k = r.k;
    std::vector<double> pf(size_t(k) + 1, 0.0), pr(size_t(k) + 1, 0.0);
    pf[0] = r.dpdf;
    if (k >= 1) pf[1] = r.eps * r.rsf[0] * r.cin[1] / (r.dist[0] * r.dist[0]);
    for (int j = 2; j <= k - 1; j++)
        pf[size_t(j)] = r.fpdf[size_t(j - 1)] * r.rsf[size_t(j - 1)] * r.cin[size_t(j)] /
                        (r.dist[size_t(j - 1)] * r.dist[size_t(j - 1)]);
    pr[size_t(k)] = 1.0;
    pr[size_t(k - 1)] = r.cam_pdf * r.rsr[size_t(k - 1)] * r.cout[size_t(k - 1)] /
                        (r.dist[size_t(k - 1)] * r.dist[size_t(k - 1)]);
    for (int j = k - 2; j >= 1; j--)
        pr[size_t(j)] = r.rpdf[size_t(j + 1)] * r.rsr[size_t(j)] * r.cout[size_t(j)] / (r.dist[size_t(j)] * r.dist[size_t(j)]);
    if (k >= 2) pr[0] = r.rpdf[1] * r.rsr[0] * r.cos_light / (r.dist[0] * r.dist[0]);
    if (p.reference_nee_weight && r.tech == MT_NEE && k >= 2) pr[0] /= r.rsr[0];

    auto prod_f = [&](int a, int b) { // prod_{a <= i < b} P>_i
        double v = 1.0;
        for (int i = a; i < b; i++) v *= pf[size_t(i)];
        return v;
    };
    auto prod_r = [&](int a, int b) { // prod_{a <= i < b} P<_i
        double v = 1.0;
        for (int i = a; i < b; i++) v *= pr[size_t(i)];
        return v;
    };
    // merge factors at x_m: total, and the share of r's own technique
    auto merge_factor = [&](int m, double* own_share) {
        if (r.type[size_t(m)] == MV_SURFACE) {
            *own_share = r.tech == MT_SURF ? p.mf_surf : 0.0;
            return p.mf_surf;
        }
        const double fp = p.long_photon_beams ? 1.0 / r.rsf[size_t(m - 1)] : r.ratio[size_t(m)];
        const double rp = p.long_query_beams ? 1.0 / r.rsr[size_t(m)] : r.ratio[size_t(m)];
        const double pp = p.mf_pp3d, pb = p.mf_pb2d * rp, bp = p.mf_bp2d * fp,
                     bb = p.mf_bb1d * r.sin_t[size_t(m)] * fp * rp;
        *own_share = r.tech == MT_PP3D ? pp : r.tech == MT_PB2D ? pb : r.tech == MT_BP2D ? bp : r.tech == MT_BB1D ? bb : 0;
        return pp + pb + bp + bb;
    };

    const bool beam_merge = r.tech == MT_BP2D || r.tech == MT_BB1D;
    double sum = 0.0, own = 0.0;
    // BPT, s = 0 .. k-1 (every vertex here is non-delta and every pair can be connected)
    for (int s = 0; s <= k - 1; s++) {
        if (s == 0 && k < 2) continue;
        double v = p.n_bpt * prod_f(0, s) * prod_r(s, k + 1);
        if (p.reference_beam_distance && beam_merge && r.ref_light_dist > 0 && s == r.s) { // 1/P>_m as the reference has it
            const double d = r.dist[size_t(s - 1)];
            v *= (r.ref_light_dist * r.ref_light_dist) / (d * d);
        }
        sum += v;
        if ((r.tech == MT_BPT || r.tech == MT_NEE || r.tech == MT_LIGHT_HIT) && r.s == s) own = v;
    }
    // light tracing
    if (r.type[size_t(k - 1)] == MV_MEDIUM || p.connect_from_surf) {
        const double v = p.n_bpt * p.n_light * prod_f(0, k);
        sum += v;
        if (r.tech == MT_LT) own = v;
    }
    // merges
    for (int m = 1; m <= k - 1; m++) {
        double share;
        const double f = merge_factor(m, &share);
        const double base = prod_f(0, m + 1) * prod_r(m, k + 1);
        sum += base * f;
        if (r.s == m && r.tech >= MT_SURF) own = base * share;
    }
    return sum > 0.0 ? own / sum : 0.0;
}

} // namespace orc
