/*
 * oracle_mis.h -- FP64 full-path MIS oracle (test infrastructure only).
 *
 * A full path x_0 (on a light) .. x_k (the camera) is described by the primitive
 * sampling densities the renderer's estimators consume: per interior vertex the
 * direction pdfs of continuing toward either neighbour (continuation probability
 * included), the cosines that convert them to area measure, per edge the distance
 * sampling pdfs in light and camera direction, and the light and camera pdfs at the
 * ends. From these, mis_explicit_weight() forms the pdf of every technique that can
 * produce the path -- BPT connections with s light vertices, light tracing, and the
 * point/beam merges at every non-delta interior vertex -- and returns the balance
 * heuristic weight p_T / sum_T' p_T' in FP64, as defined by Krivanek et al. 2014
 * ("Unifying points, beams, and paths in volumetric light transport simulation",
 * Sec. 5) and by the VCM weights it extends (Georgiev et al. 2012). The renderer's
 * constant-time recursion (gpuupbp/kernel/path_weight.cuh:35-51,
 * functs_shared.cuh:81-141 and the estimators citing them) must reproduce that
 * weight; tests/test_mis_oracle.py checks it within 1e-4.
 */
#ifndef UPBP_ORACLE_MIS_H
#define UPBP_ORACLE_MIS_H

#include <cstdint>
#include <vector>

namespace orc {

enum MisTech : int { MT_LIGHT_HIT = 0, MT_NEE = 1, MT_BPT = 2, MT_LT = 3, MT_SURF = 4, MT_PP3D = 5, MT_PB2D = 6,
                     MT_BP2D = 7, MT_BB1D = 8 };
enum MisVType : int { MV_LIGHT = 0, MV_SURFACE = 1, MV_MEDIUM = 2, MV_CAMERA = 3 };

struct MisRec {
    int tech = 0;
    int s = 0;     // BPT family: number of light-side vertices (x_0 .. x_{s-1}); merges: merge vertex m
    int k = 0;     // x_k is the camera
    double weight = 0.0; // FP32 constant-time weight the estimator computed
    // per vertex j = 0..k
    std::vector<int> type;
    std::vector<double> fpdf; // interior: pdf of sampling x_{j+1} from x_j given x_{j-1} (solid angle) * continuation
    std::vector<double> rpdf; // interior: pdf of sampling x_{j-1} from x_j given x_{j+1} (solid angle) * continuation
    std::vector<double> cin;  // |cos| at x_j against the edge to x_{j-1} (1 in media)
    std::vector<double> cout; // |cos| at x_j against the edge to x_{j+1} (1 in media)
    std::vector<double> sin_t; // sin of the angle between the two edges at x_j (medium vertices)
    std::vector<double> ratio; // 1 / smallest positive extinction at medium vertices
    // per edge j = 0..k-1 (x_j -> x_{j+1})
    std::vector<double> dist, rsf, rsr; // length, distance pdfs toward x_{j+1} (light) and toward x_j (camera)
    double dpdf = 0, eps = 0, cos_light = 0; // light: position pdf (area, pick included), emission pdf (solid angle)
    double cam_pdf = 0;                      // camera: solid-angle pdf of the direction to x_{k-1}
    // beam merges on the light side: the length the reference squares into the light-side
    // d_shared at the merge point (the distance along the beam from its segment start, or 1
    // when it skips the factor); 0 when the true edge length is used
    double ref_light_dist = 0;
    // key of the contribution: the light path of its light-side vertices (light hits and
    // next-event estimation: the light)
    int64_t lpath = -1;
};

struct MisParams {
    double n_bpt, n_light, mf_surf, mf_pp3d, mf_pb2d, mf_bp2d, mf_bb1d;
    bool long_photon_beams, long_query_beams, connect_from_surf;
    // Model the reference's next-event estimation weight, whose light-hit term omits the
    // camera-direction distance pdf of the shadow edge (gpuupbp/kernel/functs_camera.cuh:290,
    // w_light = bsdf pdf / (pick * direct pdf)); the light-hit estimator (:146-162) includes it.
    bool reference_nee_weight;
    // Model the reference's light-side beam distance in B-P2D / B-B1D (measured from the
    // beam's medium-segment start, functs_bp2d.cuh:138, and skipped for beams leaving a
    // finite light, functs_bb1d.cuh:191) instead of the light vertex distance.
    bool reference_beam_distance;
};

// Balance-heuristic weight of r's own technique among all techniques that sample its path.
double mis_explicit_weight(const MisRec& r, const MisParams& p);

} // namespace orc

#endif
