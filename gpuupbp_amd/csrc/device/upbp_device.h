// Device-side data model and leaf functions of the UPBP hot path (HIP, gfx950).
//
// Restates the semantics of nolmoonen/gpuupbp gpuupbp/kernel/{bsdf,phase_function,
// sample,medium,light,camera,intersection}.cuh with a data layout of our own:
// POD scene tables in HBM, compact light-vertex and beam records, and a
// counter-based Philox stream per (path, segment, purpose) instead of a
// per-thread cuRAND state. All functions are __host__ __device__ so that the
// host side can size buffers with the same structs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define UPBP_HD __host__ __device__ __forceinline__

namespace upbp {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.318309886183790671538f;
constexpr float kInf = 1e36f;         // shared/shared_defs.h:40
constexpr float kEpsRay = 1e-3f;      // kernel/defs.cuh:114
constexpr float kEpsCosine = 1e-6f;   // kernel/defs.cuh:120
constexpr float kEpsPhong = 1e-3f;    // kernel/defs.cuh:123
constexpr int kThinWall = -1;         // shared/shared_defs.h:30
constexpr int kGlobalPriority = -2;   // shared/shared_defs.h:31
constexpr int kMaxSegments = 10;      // kernel/defs.cuh:53
constexpr int kMaxRealHits = 10;      // kernel/defs.cuh:57
constexpr int kMaxStack = 20;         // kernel/defs.cuh:61
constexpr int kMaxImaginary = 8;      // kernel/defs.cuh:65
constexpr unsigned kOriginInMedium = 1u, kEndInMedium = 2u;

enum Status : int { S_OK = 0, S_ERR_STACK = 1, S_ERR_REAL_HITS = 2, S_ERR_SEGMENTS = 3, S_ERR_IMAGINARY = 4 };

// technique bits (shared/shared_enums.h:30-56)
constexpr uint32_t T_BEAM_SHORT = 1u << 0, T_BEAM_LONG = 1u << 1, T_BPT = 1u << 2, T_SURF = 1u << 3,
                   T_PP3D = 1u << 4, T_PB2D = 1u << 5, T_BB1D = 1u << 6, T_BP2D = 1u << 7, T_PREVIOUS = 1u << 9,
                   T_COMPATIBLE = 1u << 10, T_SPECULAR_ONLY = 1u << 12;

// ---- float3 math ---------------------------------------------------------------
struct f3 {
    float x, y, z;
};
UPBP_HD f3 mk(float a) { return {a, a, a}; }
UPBP_HD f3 mk(float a, float b, float c) { return {a, b, c}; }
UPBP_HD f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
UPBP_HD f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
UPBP_HD f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
UPBP_HD f3 operator*(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
UPBP_HD f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
UPBP_HD f3 operator*(float s, f3 a) { return {a.x * s, a.y * s, a.z * s}; }
UPBP_HD f3 operator/(f3 a, f3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
UPBP_HD f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
UPBP_HD f3& operator+=(f3& a, f3 b)
{
    a = a + b;
    return a;
}
UPBP_HD f3& operator*=(f3& a, f3 b)
{
    a = a * b;
    return a;
}
UPBP_HD f3& operator*=(f3& a, float s)
{
    a = a * s;
    return a;
}
UPBP_HD f3& operator/=(f3& a, float s)
{
    a = a / s;
    return a;
}
UPBP_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
UPBP_HD f3 cross(f3 a, f3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
UPBP_HD f3 normalize(f3 a) { return a * (1.f / sqrtf(dot(a, a))); }
UPBP_HD float maxc(f3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
UPBP_HD float comp(f3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
UPBP_HD bool black_or_neg(f3 a) { return a.x <= 0.f && a.y <= 0.f && a.z <= 0.f; } // kernel/defs.cuh:126
UPBP_HD bool positive(float a) { return a > 1e-20f; }                                // kernel/defs.cuh:128
UPBP_HD bool positive(f3 a) { return positive(a.x) && positive(a.y) && positive(a.z); }

// orthonormal frame from a z direction (shared/frame.h:45-51)
struct Frame {
    f3 x, y, z;
    UPBP_HD void set_from_z(f3 pz)
    {
        f3 tz = z = normalize(pz);
        f3 tx = (fabsf(tz.x) > .99f) ? mk(0, 1, 0) : mk(1, 0, 0);
        y = normalize(cross(tz, tx));
        x = cross(y, tz);
    }
    UPBP_HD f3 to_world(f3 a) const { return x * a.x + y * a.y + z * a.z; }
    UPBP_HD f3 to_local(f3 a) const { return {dot(a, x), dot(a, y), dot(a, z)}; }
};

// ---- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) -------------------------
UPBP_HD void philox(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = uint64_t(0xD2511F53u) * c[0];
        const uint64_t p1 = uint64_t(0xCD9E8D57u) * c[2];
        const uint32_t n0 = uint32_t(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n2 = uint32_t(p0 >> 32) ^ c[3] ^ k1;
        c[1] = uint32_t(p1);
        c[3] = uint32_t(p0);
        c[0] = n0;
        c[2] = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

enum Purpose : uint32_t { P_EMIT = 1, P_MEDIUM = 2, P_SCATTER = 3, P_JITTER = 4, P_NEE = 5, P_BPT_PICK = 6 };
constexpr uint32_t kDomainLight = 0x6C696768u, kDomainCamera = 0x63616D65u;

// Stream of uniforms in [0,1), keyed by (domain, path, iteration, segment, purpose).
struct Rng {
    uint32_t path, seg_purpose, iter_lo, key0, key1, idx;
    uint32_t buf[4];
    UPBP_HD Rng(uint32_t domain, uint32_t p, uint64_t iter, uint32_t seg, uint32_t purpose)
        : path(p), seg_purpose((seg << 8) | purpose), iter_lo(uint32_t(iter)), key0(domain),
          key1(uint32_t(iter >> 32)), idx(0)
    {
    }
    UPBP_HD float next()
    {
        if ((idx & 3u) == 0) {
            buf[0] = idx >> 2;
            buf[1] = seg_purpose;
            buf[2] = path;
            buf[3] = iter_lo;
            philox(buf, key0, key1);
        }
        const uint32_t u = buf[idx & 3u];
        idx++;
        return float(u >> 8) * (1.0f / 16777216.0f);
    }
};

// ---- scene tables ------------------------------------------------------------
struct DMaterial {
    f3 diffuse, phong, mirror;
    float phong_exp, ior;
    int priority, real, med_idx;
};

struct DMedium { // shared/medium.h, derived as in host/medium.hpp:39-85
    f3 absorption, emission, scattering, attenuation;
    float min_pos_att;
    int min_pos_idx;
    float max_beam_length;
    int has_scattering, has_attenuation;
    float cont_prob, mean_cosine;
};

enum LightType : int { L_AREA = 0, L_BACKGROUND = 1, L_DIRECTIONAL = 2, L_POINT = 3 };
struct DLight {
    int type, mat_id, med_id;
    f3 p0, e1, e2;
    Frame frame;
    f3 intensity;
    float inv_area;
    f3 position;
    float scale;
    int uses_env; // background light reading sc.env
};

// Pinhole camera. Rays are formed directly from the image-plane basis; this is the
// same mapping as the raster<->world matrices of host/camera.hpp:33-76.
struct DCamera {
    f3 origin, forward, right, up;
    float res_x, res_y, ipd; // ipd: distance of the unit-pixel image plane
    int mat_id, med_id;
};

// LBVH node: both child boxes inline; child >= 0 is an inner node, child < 0 is
// leaf ~child (a position in the sorted primitive order).
struct __align__(16) BNode {
    float lx0, ly0, lz0, ux0;
    float uy0, uz0, lx1, ly1;
    float lz1, ux1, uy1, uz1;
    int c0, c1, pad0, pad1;
};

struct Bvh {
    const BNode* nodes;
    const uint32_t* order; // sorted position -> primitive index
    uint32_t n;            // primitives
};

// 8-wide node collapsed from the binary LBVH: 128 bytes, one cache line per node visit. Child
// boxes are quantised to 8 bits per plane on a power-of-two grid anchored at the node box's low
// corner; decoding is origin + q * step in FP32 and the builder picks q so that the decoded box
// contains the exact child box.
constexpr int kWide = 8;
struct __align__(128) WNode {
    float ox, oy, oz;
    uint32_t meta;      // bits 0-7 / 8-15 / 16-23: biased exponents of the x / y / z steps; 24-27: child count
    int child[kWide];   // >= 0: wide node, < 0: ~sorted leaf position
    uint32_t q[12];     // plane p (low x, y, z, high x, y, z) of child k: byte k & 3 of q[2 p + (k >> 2)]
    uint32_t pad[8];
};

struct WBvh {
    const WNode* nodes;
    uint32_t n; // primitives
};

UPBP_HD float wnode_step(uint32_t meta, int axis) { return __builtin_bit_cast(float, ((meta >> (8 * axis)) & 255u) << 23); }
UPBP_HD float wnode_decode(float origin, uint32_t word, int k, float step)
{
#pragma clang fp contract(off)
    return origin + float((word >> (8 * k)) & 255u) * step;
}
UPBP_HD int wnode_count(uint32_t meta) { return int((meta >> 24) & 15u); }
// child k's box, decoded (steps sx, sy, sz from wnode_step)
UPBP_HD void wnode_box(const WNode& nd, int k, float sx, float sy, float sz, float& lx, float& ly, float& lz, float& ux,
                       float& uy, float& uz)
{
    const int w = k >> 2, b = k & 3;
    lx = wnode_decode(nd.ox, nd.q[0 + w], b, sx);
    ly = wnode_decode(nd.oy, nd.q[2 + w], b, sy);
    lz = wnode_decode(nd.oz, nd.q[4 + w], b, sz);
    ux = wnode_decode(nd.ox, nd.q[6 + w], b, sx);
    uy = wnode_decode(nd.oy, nd.q[8 + w], b, sy);
    uz = wnode_decode(nd.oz, nd.q[10 + w], b, sz);
}

struct TriSet {
    const float4* tri;  // 3 float4 per sorted triangle: p0, e1 = p1 - p0, e2 = p2 - p0
    const float4* nrm;  // 4 float4 per sorted triangle: n0, n1, n2, face normal
    const int2* rec;    // (material, light) per sorted triangle
    Bvh bvh;
};

// analytic spheres: centre and radius, (material, light) per sphere
struct SphSet {
    const float4* sph;
    const int2* rec;
    uint32_t n;
};

// environment map of the background light and its luminance distribution (env_map.h)
struct DEnvMap {
    const float* img;       // width * height * 3, rows from the top
    const float* cond_func; // nv * nu
    const float* cond_cdf;  // nv * (nu + 1)
    const float* cond_int;  // nv
    const float* marg_func; // nv
    const float* marg_cdf;  // nv + 1
    float marg_int;
    int width, height, nu, nv; // nu = 2 * height columns, nv = height rows
    float norm;                // 1 / (2 pi^2): (u, v) density to solid angle, before the 1 / sin(theta)
};

// learned proposal of camera paths in media (guide.h): per (grid cell, arrival bin) record
constexpr int kGuideLobes = 8, kGuideRec = 10;
constexpr int kGuidePend = 12, kGuideRaw = 12; // floats per pending / raw training record
constexpr float kGuideKappaMin = 1e-2f, kGuideKappaMax = 1e4f; // lobe concentration range
struct DGuide {
    const float4* recs; // res^3 records of kGuideRec float4, cell (z * res + y) * res + x
    f3 lo, inv_cell;    // grid origin, cells per unit length per axis
    int res;            // cells per axis
    float alpha;        // probability of sampling the product at camera medium vertices (0: off)
    float alpha_s;      // the same at non-specular camera surface vertices (0: off)
    int train_media, train_surfaces; // which scatterings feed the training samples
};

struct DScene {
    const DMaterial* mats;
    const DMedium* media;
    const DLight* lights;
    uint32_t n_mats, n_media, n_lights, global_medium;
    int background_light;
    DCamera cam;
    f3 center;
    float radius, inv_radius_sqr;
    TriSet real, imag;
    SphSet sph_real, sph_imag;
    DEnvMap env; // img == nullptr: no environment map
    DGuide guide;
    int use_shading_normal;
};

// ---- sampling helpers: kernel/sample.cuh ----------------------------------------
UPBP_HD float power_cos_hemisphere_pdf_w(f3 n, f3 d, float power)
{
    const float c = fmaxf(0.f, dot(n, d));
    return (power + 1.f) * powf(c, power) * (kInvPi * .5f);
}
UPBP_HD f3 sample_power_cos_hemisphere_w(float u0, float u1, float power)
{
    const float t1 = 2.f * kPi * u0;
    const float t2 = powf(u1, 1.f / (power + 1.f));
    const float t3 = sqrtf(1.f - t2 * t2);
    return {cosf(t1) * t3, sinf(t1) * t3, t2};
}
UPBP_HD void sample_concentric_disc(float u0, float u1, float& ox, float& oy)
{
    float phi, r;
    const float a = 2 * u0 - 1, b = 2 * u1 - 1;
    if (a > -b) {
        if (a > b) {
            r = a;
            phi = (kPi / 4.f) * (b / a);
        } else {
            r = b;
            phi = (kPi / 4.f) * (2.f - (a / b));
        }
    } else {
        if (a < b) {
            r = -a;
            phi = (kPi / 4.f) * (4.f + (b / a));
        } else {
            r = -b;
            phi = b != 0 ? (kPi / 4.f) * (6.f - (a / b)) : 0.f;
        }
    }
    ox = r * cosf(phi);
    oy = r * sinf(phi);
}
UPBP_HD f3 sample_cos_hemisphere_w(float u0, float u1, float* pdf)
{
    const float t1 = 2.f * kPi * u0, t2 = sqrtf(u1);
    f3 r{cosf(t1) * t2, sinf(t1) * t2, sqrtf(1.f - u1)};
    if (pdf) *pdf = r.z * kInvPi;
    return r;
}
UPBP_HD float cos_hemisphere_pdf_w(f3 n, f3 d) { return fmaxf(0.f, dot(n, d)) * kInvPi; }
UPBP_HD f3 sample_uniform_sphere_w(float u0, float u1, float* pdf)
{
    const float t1 = 2.f * kPi * u0, t2 = 2.f * sqrtf(u1 - u1 * u1);
    if (pdf) *pdf = kInvPi * .25f;
    return {cosf(t1) * t2, sinf(t1) * t2, 1.f - 2.f * u1};
}
UPBP_HD float uniform_sphere_pdf_w() { return kInvPi * .25f; }
UPBP_HD float concentric_disc_pdf_a() { return kInvPi; }
UPBP_HD float pdf_w_to_a(float pdf_w, float dist, float cos_there) { return pdf_w * fabsf(cos_there) / (dist * dist); }
UPBP_HD float luminance(f3 c) { return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z; }
UPBP_HD float fresnel_dielectric(float cos_inc, float ior) // kernel/sample.cuh:184-201
{
    if (ior < 0) return 1.f;
    if (cos_inc < 0.f) cos_inc = -cos_inc;
    const float sin_t2 = ior * ior * (1.f - cos_inc * cos_inc);
    const float cos_t = sqrtf(fmaxf(0.f, 1.f - sin_t2));
    const float t1 = ior * cos_t;
    const float rpar = (cos_inc - t1) / (cos_inc + t1);
    const float t2 = ior * cos_inc;
    const float rper = (t2 - cos_t) / (t2 + cos_t);
    return 0.5f * (rpar * rpar + rper * rper);
}
UPBP_HD f3 reflect_local(f3 d) { return {-d.x, -d.y, d.z}; }

// ---- Henyey-Greenstein: kernel/phase_function.cuh:34-137 ------------------------
UPBP_HD bool isotropic(float g) { return fabsf(g) < 1e-3f; }
UPBP_HD float hg_pdf(f3 dir_fix, f3 dir_gen, float g, float* sin_theta)
{
    const float c = -dot(dir_gen, dir_fix);
    if (sin_theta) *sin_theta = sqrtf(fmaxf(0.f, 1.f - c * c));
    if (isotropic(g)) return uniform_sphere_pdf_w();
    const float g2 = g * g;
    const float d = 1.f + g2 - (g + g) * c;
    return d > 0.f ? uniform_sphere_pdf_w() * (1.f - g2) / (d * sqrtf(d)) : 0.f;
}
UPBP_HD f3 hg_sample(f3 dir_fix, float g, float u0, float u1, const Frame& frame, float& pdf, float* sin_theta)
{
    f3 gen;
    if (isotropic(g)) {
        gen = sample_uniform_sphere_w(u0, u1, &pdf);
        if (sin_theta) {
            const float c = -dot(gen, dir_fix);
            *sin_theta = sqrtf(fmaxf(0.f, 1.f - c * c));
        }
    } else {
        const float g2 = g * g, two_g = g + g;
        const float s = (1.f - g2) / (1.f - g + two_g * u0);
        const float ct = (1.f + g2 - s * s) / two_g;
        const float st = sqrtf(fmaxf(0.f, 1.f - ct * ct));
        const float phi = 2.f * kPi * u1;
        const float d = 1.f + g2 - two_g * ct;
        gen = frame.to_world({cosf(phi) * st, sinf(phi) * st, ct});
        pdf = d > 0.f ? uniform_sphere_pdf_w() * (1.f - g2) / (d * sqrtf(d)) : 0.f;
        if (sin_theta) *sin_theta = st;
    }
    return gen;
}

// ---- homogeneous media: kernel/medium.cuh:33-171 --------------------------------
UPBP_HD f3 eval_attenuation(const DMedium& m, float dist)
{
    return {expf(-m.attenuation.x * dist), expf(-m.attenuation.y * dist), expf(-m.attenuation.z * dist)};
}
UPBP_HD float sample_ray(const DMedium& m, float dist_to_boundary, float rnd, float* for_pdf, unsigned flags,
                         float* rev_pdf)
{
    if (m.min_pos_att != 0.f && m.has_scattering) {
        const float s = -logf(rnd) / m.min_pos_att;
        if (s < dist_to_boundary) {
            const float att = expf(-m.min_pos_att * s);
            *for_pdf = m.min_pos_att * att;
            *rev_pdf = (flags & kOriginInMedium) ? *for_pdf : att;
            return s;
        }
        const float att = expf(-m.min_pos_att * dist_to_boundary);
        *for_pdf = att;
        *rev_pdf = (flags & kOriginInMedium) ? m.min_pos_att * att : att;
        return dist_to_boundary;
    }
    *for_pdf = 1.f;
    *rev_pdf = 1.f;
    return dist_to_boundary;
}
UPBP_HD float ray_sample_pdf(const DMedium& m, float dmin, float dmax, unsigned flags, float* rev_pdf)
{
    float for_pdf = 1.f;
    if (rev_pdf) *rev_pdf = 1.f;
    if (m.min_pos_att != 0.f && m.has_scattering) {
        const float att = fmaxf(expf(-m.min_pos_att * (dmax - dmin)), 1e-35f);
        const float minatt = fmaxf(m.min_pos_att * att, 1e-35f);
        for_pdf = (flags & kEndInMedium) ? minatt : att;
        if (rev_pdf) *rev_pdf = (flags & kOriginInMedium) ? minatt : att;
    }
    return for_pdf;
}

// ---- range-query predicates: kernel/intersection.cuh:33-173 ---------------------
// No contraction inside these bodies: the accepted index sets must not depend on
// FMA formation, whatever the translation unit's default is.
UPBP_HD bool pred_point_point(f3 p0, f3 p1, float ra2)
{
#pragma clang fp contract(off)
    const f3 d = p0 - p1;
    return dot(d, d) <= ra2;
}
UPBP_HD bool pred_point_beam(f3 ro, f3 rd, float rt, f3 ce, float ra2)
{
#pragma clang fp contract(off)
    const f3 eo = ce - ro;
    const float t = dot(eo, rd);
    if (t <= 0.f || t >= rt) return false;
    const f3 tmp = ro + rd * t - ce;
    return dot(tmp, tmp) <= ra2;
}
UPBP_HD bool pred_bre(float& isect_dist, float& isect_rad2, f3 ro, f3 rd, float min_t, float max_t, f3 p, float r2)
{
#pragma clang fp contract(off)
    const f3 op = p - ro;
    const float dist = dot(op, rd);
    if (dist > min_t && dist < max_t) {
        const f3 ir = ro + rd * dist - p;
        const float ir2 = dot(ir, ir);
        if (ir2 <= r2) {
            isect_dist = dist;
            isect_rad2 = ir2;
            return true;
        }
    }
    return false;
}
UPBP_HD bool pred_beam_beam(f3 o1, f3 d1, float min1, float max1, f3 o2, f3 d2, float min2, float max2, float maxd2,
                            float& distance, float& sin_theta, float& t1, float& t2)
{
#pragma clang fp contract(off)
    const f3 c = cross(d1, d2);
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
    if (t2 <= min2 || t2 >= max2 || isnan(t2)) return false;
    sin_theta = sqrtf(s2);
    distance = fabsf(ad) / sin_theta;
    return true;
}

} // namespace upbp
