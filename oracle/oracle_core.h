// Oracle internals: math, scene preparation, RNG streams, BSDF / phase / light /
// medium leaf functions. TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Every function restates a reference function (nolmoonen/gpuupbp, paths relative
// to the reference root); the citation is on each definition.
#ifndef UPBP_ORACLE_CORE_H
#define UPBP_ORACLE_CORE_H

#include "oracle.h"

#include "../gpuupbp_amd/csrc/device/phase_lobes.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace orc {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.318309886183790671538f;
constexpr float kInf = 1e36f;          // shared/shared_defs.h:40
constexpr float kEpsRay = 1e-3f;       // kernel/defs.cuh:114
constexpr float kEpsLaunch = 1e-5f;    // kernel/defs.cuh:117
constexpr float kEpsCosine = 1e-6f;    // kernel/defs.cuh:120
constexpr float kEpsPhong = 1e-3f;     // kernel/defs.cuh:123
constexpr int kThinWall = -1;          // shared/shared_defs.h:30
constexpr int kGlobalPriority = -2;    // shared/shared_defs.h:31
constexpr int kMaxSegments = 10;       // kernel/defs.cuh:53
constexpr int kMaxIntersections = 10;  // kernel/defs.cuh:57
constexpr int kMaxStack = 20;          // kernel/defs.cuh:61
constexpr int kMaxImaginary = 8;       // kernel/defs.cuh:65

enum Status { OK = 0, ERR_BOUNDARY_STACK = 1, ERR_INTERSECTION_ARRAY = 2, ERR_VOL_SEGMENT_ARRAY = 3,
              ERR_INTERSECTION_STACK = 4 };
constexpr unsigned ORIGIN_IN_MEDIUM = 1u, END_IN_MEDIUM = 2u;

struct V {
    float x, y, z;
};
inline V v3(float a) { return {a, a, a}; }
inline V v3(float a, float b, float c) { return {a, b, c}; }
inline V operator+(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V operator-(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V operator-(V a) { return {-a.x, -a.y, -a.z}; }
inline V operator*(V a, V b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline V operator*(V a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V operator*(float s, V a) { return {a.x * s, a.y * s, a.z * s}; }
inline V operator/(V a, V b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
inline V operator/(V a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline V& operator+=(V& a, V b) { a = a + b; return a; }
inline V& operator*=(V& a, V b) { a = a * b; return a; }
inline V& operator*=(V& a, float s) { a = a * s; return a; }
inline V& operator/=(V& a, float s) { a = a / s; return a; }
inline float dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V cross(V a, V b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float length(V a) { return std::sqrt(dot(a, a)); }
inline V normalize(V a) { return a * (1.f / std::sqrt(dot(a, a))); }
inline float maxc(V a) { return std::fmax(a.x, std::fmax(a.y, a.z)); }
inline V fromp(const float* p) { return {p[0], p[1], p[2]}; }

// kernel/defs.cuh:126-136
inline bool black_or_neg(V a) { return a.x <= 0.f && a.y <= 0.f && a.z <= 0.f; }
inline bool positive(float a) { return a > 1e-20f; }
inline bool positive(V a) { return positive(a.x) && positive(a.y) && positive(a.z); }

// shared/frame.h:34-65
struct Frame {
    V x{1, 0, 0}, y{0, 1, 0}, z{0, 0, 1};
    void set_from_z(V pz)
    {
        V tz = z = normalize(pz);
        V tx = (std::fabs(tz.x) > .99f) ? v3(0, 1, 0) : v3(1, 0, 0);
        y = normalize(cross(tz, tx));
        x = cross(y, tz);
    }
    V to_world(V a) const { return x * a.x + y * a.y + z * a.z; }
    V to_local(V a) const { return {dot(a, x), dot(a, y), dot(a, z)}; }
};

// ---- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) ----------------------
inline void philox(const uint32_t in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = uint64_t(0xD2511F53u) * c0;
        const uint64_t p1 = uint64_t(0xCD9E8D57u) * c2;
        const uint32_t n0 = uint32_t(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = uint32_t(p0 >> 32) ^ c3 ^ k1;
        c0 = n0;
        c1 = uint32_t(p1);
        c2 = n2;
        c3 = uint32_t(p0);
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// Random stream keyed by (domain, path, iteration, segment, purpose); values in [0,1).
// Replaces the cuRAND Philox state of kernel/rng.cuh:36-65 (not its seeding/conversion).
enum Purpose : uint32_t { P_EMIT = 1, P_MEDIUM = 2, P_SCATTER = 3, P_JITTER = 4, P_NEE = 5, P_BPT_PICK = 6 };
constexpr uint32_t kDomainLight = 0x6C696768u;
constexpr uint32_t kDomainCamera = 0x63616D65u;

struct Stream {
    uint32_t ctr[4];
    uint32_t key[2];
    uint32_t idx = 0;
    uint32_t buf[4];
    Stream(uint32_t domain, uint32_t path, uint64_t iter, uint32_t seg, uint32_t purpose)
    {
        ctr[0] = 0;
        ctr[1] = (seg << 8) | purpose;
        ctr[2] = path;
        ctr[3] = uint32_t(iter);
        key[0] = domain;
        key[1] = uint32_t(iter >> 32);
    }
    float next()
    {
        if ((idx & 3u) == 0) {
            ctr[0] = idx >> 2;
            philox(ctr, key, buf);
        }
        const uint32_t u = buf[idx & 3u];
        idx++;
        return float(u >> 8) * (1.0f / 16777216.0f);
    }
};

// ---- scene -------------------------------------------------------------------
struct Material {
    V diffuse, phong, mirror;
    float phong_exp, ior;
    int priority;
    bool real;
    int med_idx;
};

// shared/medium.h + host/medium.hpp:39-85
struct Medium {
    V absorption, emission, scattering, attenuation;
    float min_pos_att;
    int min_pos_idx;
    float max_beam_length;
    bool has_scattering, has_attenuation;
    float cont_prob, mean_cosine;
};

enum LightType { L_AREA = 0, L_BACKGROUND = 1, L_DIRECTIONAL = 2, L_POINT = 3 };
struct Light {
    int type, mat_id, med_id;
    V p0, e1, e2;
    Frame frame;
    V intensity;
    float inv_area;
    V position;
    V bg_color;
    float scale;
    bool uses_env = false;
};

// piecewise-constant distribution over [0, 1) (host/env_map.hpp:42-68)
struct Dist1D {
    std::vector<float> func, cdf; // n, n + 1
    float integral = 0.f;
};

// environment map: lat-long image and its luminance distribution (shared/env_map.h:34-71)
struct EnvMap {
    int width = 0, height = 0;
    std::vector<V> img;          // rows from the top
    std::vector<Dist1D> rows;    // conditional over u per image row
    Dist1D marginal;             // over rows
    float norm = 0.f;
};

struct Mat4 {
    float m[16];
};

// shared/camera.h:32-50
struct Camera {
    V origin, direction;
    float res_x, res_y;
    Mat4 raster_to_world, world_to_raster;
    float image_plane_dist;
    int mat_id, med_id;
};

struct Tri {
    V p0, p1, p2;
    V n0, n1, n2, face;
    int mat, light;
};

// analytic sphere (predefined scenes)
struct Sph {
    V c;
    float r;
    int mat, light;
};

// Learned proposal of camera paths in media: res^3 cells, each a mixture of 8 vMF lobes of
// incident radiance, sampled as its product with the vMF lobe fitted to the phase function
// (no counterpart in the reference). alpha 0: off.
struct GuideLobe {
    V mu;
    float kappa, w, c; // concentration, normalised weight, w kappa / (2 pi (1 - exp(-2 kappa)))
};
struct Guide {
    float alpha = 0.f, alpha_s = 0.f; // media, non-specular surfaces
    int res = 0;
    V lo{0, 0, 0}, inv_cell{0, 0, 0};
    std::vector<GuideLobe> lobes; // 8 per record, record (z * res + y) * res + x
};

struct Scene {
    std::vector<Tri> real, imag;
    Guide guide;
    std::vector<Sph> sreal, simag;
    EnvMap env; // width 0: none
    std::vector<Material> mats;
    std::vector<Medium> media;
    std::vector<Light> lights;
    uint32_t global_medium;
    int background_light;
    Camera cam;
    V center;
    float radius, inv_radius_sqr;
};

void prepare_scene(const upbp_scene* in, uint32_t res_x, uint32_t res_y, Scene& out);
// learned proposal: records from (mu, kappa, weight) x 8 per record; the product of a record
// with the phase function's three fitted vMF lobes (the table of
// gpuupbp_amd/csrc/device/phase_lobes.h, shared data); the camera technique's pdf at a medium
// point (arriving along `arrive`, leaving along `gen`)
void init_guide(Guide& g, const float* lobes, int res, const float* lo, const float* hi, float alpha);
extern Guide g_oracle_guide;
const GuideLobe* guide_record(const Guide& g, V p);
struct PhaseFit { // h(w) = sum_{j < n} a_j exp(k_j (mu . w - 1))
    V mu;
    int n;
    float k[3], a[3];
};
PhaseFit phase_fit(V arrive, float mean_cosine);
PhaseFit cosine_fit(V normal);
bool surface_guidable(const Material& m);
float guide_prod_pdf(const GuideLobe* rec, const PhaseFit& h, V w);
V guide_prod_sample(const GuideLobe* rec, const PhaseFit& h, float us, float u0, float u1);
float guide_pdf(const Guide& g, float mean_cosine, V p, V arrive, V gen);
float cam_medium_pdf(const Scene& sc, float mean_cosine, V p, V arrive, V gen);
// environment map (host/env_map.hpp:42-144, kernel/env_map.cuh:36-137)
void init_env_map(const upbp_env_map& in, EnvMap& m);
V env_sample(const EnvMap& m, float u0, float u1, float& pdf_w, V& radiance);
V env_lookup(const EnvMap& m, V direction, float* pdf_w);
float env_pdf_uv(const EnvMap& m, float u, float v);
// Ray against an analytic sphere (gpuupbp/kernel/intersector.cu:35-95): roots in (tmin, tmax),
// near first, with outward normals; returns their number (0..2).
int sphere_hit(const Sph& s, V ro, V rd, float tmin, float tmax, float* th, V* nh);
V transform_point(const Mat4& m, V p);

// ---- sampling helpers: kernel/sample.cuh -----------------------------------
// kernel/sample.cuh:38-45
inline float power_cos_hemisphere_pdf_w(V n, V d, float power)
{
    const float c = std::fmax(0.f, dot(n, d));
    return (power + 1.f) * std::pow(c, power) * (kInvPi * .5f);
}
// kernel/sample.cuh:49-58
inline V sample_power_cos_hemisphere_w(float u0, float u1, float power)
{
    const float t1 = 2.f * kPi * u0;
    const float t2 = std::pow(u1, 1.f / (power + 1.f));
    const float t3 = std::sqrt(1.f - t2 * t2);
    return {std::cos(t1) * t3, std::sin(t1) * t3, t2};
}
// kernel/sample.cuh:63-103
inline void sample_concentric_disc(float u0, float u1, float& ox, float& oy)
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
    ox = r * std::cos(phi);
    oy = r * std::sin(phi);
}
// kernel/sample.cuh:111-121
inline V sample_cos_hemisphere_w(float u0, float u1, float* pdf)
{
    const float t1 = 2.f * kPi * u0, t2 = std::sqrt(u1);
    V r{std::cos(t1) * t2, std::sin(t1) * t2, std::sqrt(1.f - u1)};
    if (pdf) *pdf = r.z * kInvPi;
    return r;
}
// kernel/sample.cuh:124-127
inline float cos_hemisphere_pdf_w(V n, V d) { return std::fmax(0.f, dot(n, d)) * kInvPi; }
// kernel/sample.cuh:131-135
inline void sample_uniform_triangle(float u0, float u1, float& b0, float& b1)
{
    const float t = std::sqrt(u0);
    b0 = 1.f - t;
    b1 = u1 * t;
}
// kernel/sample.cuh:142-152
inline V sample_uniform_sphere_w(float u0, float u1, float* pdf)
{
    const float t1 = 2.f * kPi * u0, t2 = 2.f * std::sqrt(u1 - u1 * u1);
    if (pdf) *pdf = kInvPi * .25f;
    return {std::cos(t1) * t2, std::sin(t1) * t2, 1.f - 2.f * u1};
}
inline float uniform_sphere_pdf_w() { return kInvPi * .25f; }
inline float concentric_disc_pdf_a() { return kInvPi; }
// kernel/sample.cuh:164-174
inline float pdf_w_to_a(float pdf_w, float dist, float cos_there) { return pdf_w * std::fabs(cos_there) / (dist * dist); }
// kernel/sample.cuh:176-179
inline float luminance(V c) { return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z; }
// kernel/sample.cuh:184-201
inline float fresnel_dielectric(float cos_inc, float ior)
{
    if (ior < 0) return 1.f;
    if (cos_inc < 0.f) cos_inc = -cos_inc;
    const float sin_t2 = ior * ior * (1.f - cos_inc * cos_inc);
    const float cos_t = std::sqrt(std::fmax(0.f, 1.f - sin_t2));
    const float t1 = ior * cos_t;
    const float rpar = (cos_inc - t1) / (cos_inc + t1);
    const float t2 = ior * cos_inc;
    const float rper = (t2 - cos_t) / (t2 + cos_t);
    return 0.5f * (rpar * rpar + rper * rper);
}
inline V reflect_local(V d) { return {-d.x, -d.y, d.z}; }

// ---- phase function: kernel/phase_function.cuh:34-137 ----------------------
inline bool isotropic(float g) { return std::fabs(g) < 1e-3f; }
inline float hg_pdf(V dir_fix, V dir_gen, float g, float* sin_theta)
{
    const float c = -dot(dir_gen, dir_fix);
    if (sin_theta) *sin_theta = std::sqrt(std::fmax(0.f, 1.f - c * c));
    if (isotropic(g)) return uniform_sphere_pdf_w();
    const float g2 = g * g;
    const float d = 1.f + g2 - (g + g) * c;
    return d > 0.f ? uniform_sphere_pdf_w() * (1.f - g2) / (d * std::sqrt(d)) : 0.f;
}
inline V hg_sample(V dir_fix, float g, float u0, float u1, const Frame& frame, float& pdf, float* sin_theta)
{
    V gen;
    if (isotropic(g)) {
        gen = sample_uniform_sphere_w(u0, u1, &pdf);
        if (sin_theta) {
            const float c = -dot(gen, dir_fix);
            *sin_theta = std::sqrt(std::fmax(0.f, 1.f - c * c));
        }
    } else {
        const float g2 = g * g, two_g = g + g;
        const float s = (1.f - g2) / (1.f - g + two_g * u0);
        const float ct = (1.f + g2 - s * s) / two_g;
        const float st = std::sqrt(std::fmax(0.f, 1.f - ct * ct));
        const float phi = 2.f * kPi * u1;
        const float d = 1.f + g2 - two_g * ct;
        gen = frame.to_world({std::cos(phi) * st, std::sin(phi) * st, ct});
        pdf = d > 0.f ? uniform_sphere_pdf_w() * (1.f - g2) / (d * std::sqrt(d)) : 0.f;
        if (sin_theta) *sin_theta = st;
    }
    return gen;
}

// ---- medium: kernel/medium.cuh:33-171 --------------------------------------
inline V eval_attenuation(const Medium& m, float dist)
{
    return {std::exp(-m.attenuation.x * dist), std::exp(-m.attenuation.y * dist), std::exp(-m.attenuation.z * dist)};
}
inline float comp(V v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
float sample_ray(const Medium& m, float dist_to_boundary, float rnd, float* for_pdf, unsigned flags, float* rev_pdf);
float ray_sample_pdf(const Medium& m, float dmin, float dmax, unsigned flags, float* rev_pdf);

// ---- BSDF: kernel/bsdf.cuh -------------------------------------------------
struct BSDF {
    enum Event { DIFFUSE = 1, PHONG = 2, REFLECT = 4, REFRACT = 8, SCATTER = 16, SPECULAR = 12 };
    enum Flag { IN_MEDIUM = 1, FROM_LIGHT = 2, VALID = 4, DELTA = 8 };
    int flags = 0;
    Frame frame;
    float cont_prob = 0.f;
    // surface
    int mat = -1;
    float reflect_coeff = 0.f;
    V local_dir_fix{0, 0, 0};
    float diff_prob = 0, phong_prob = 0, refl_prob = 0, refr_prob = 0;
    V geom_normal{0, 0, 0};
    float cos_fix_geom = 0.f;
    float ior = -1.f;
    // medium
    int med = -1;
    float mean_cosine = 0.f;
    V world_dir_fix{0, 0, 0};
    V scatter_coef{0, 0, 0};
    bool scatter_pos = false;
    V pos{0, 0, 0}; // scattering position (learned proposal lookup)

    bool valid() const { return flags & VALID; }
    bool delta() const { return flags & DELTA; }
    bool in_medium() const { return flags & IN_MEDIUM; }
    bool on_surface() const { return !in_medium(); }
    bool from_light() const { return flags & FROM_LIGHT; }
    float d_factor() const { return on_surface() ? std::fabs(local_dir_fix.z) : 1.f; }
    V dir_fix_world() const { return on_surface() ? frame.to_world(local_dir_fix) : world_dir_fix; }
};

struct Isect {
    float dist = kInf;
    int mat = -1, med = -1, light = -1;
    V normal{0, 0, 0};
    bool enter = false;
    V shading{0, 0, 0};
    bool in_medium() const { return mat < 0; }
    bool on_surface() const { return mat >= 0; }
};

struct Boundary {
    int mat, med;
    bool operator==(const Boundary& o) const { return mat == o.mat && med == o.med; }
};

// Priority stack: kernel/pstack.cuh:33-137 (insertion after the last element with
// priority <= the new one; pop removes the top-most matching element).
template <typename T, typename S, int N>
struct PStack {
    int n = 0;
    T el[N];
    S pr[N];
    const T& top() const { return el[n - 1]; }
    S top_priority() const { return pr[n - 1]; }
    const T& second() const { return el[n - 2]; }
    bool push(T e, S p)
    {
        if (n == N) return false;
        int pos = 0;
        for (int i = n - 1; i >= 0; i--)
            if (pr[i] <= p) {
                pos = i + 1;
                break;
            }
        for (int i = n; i > pos; i--) {
            el[i] = el[i - 1];
            pr[i] = pr[i - 1];
        }
        el[pos] = e;
        pr[pos] = p;
        n++;
        return true;
    }
    bool pop(T e, S p)
    {
        for (int i = n - 1; i >= 0 && pr[i] >= p; i--) {
            if (pr[i] == p && el[i] == e) {
                for (int j = i; j < n - 1; j++) {
                    el[j] = el[j + 1];
                    pr[j] = pr[j + 1];
                }
                n--;
                return true;
            }
        }
        return false;
    }
};
typedef PStack<Boundary, int, kMaxStack> BStack;

struct BsdfEval {
    V f;
    float cos_gen, for_pdf, rev_pdf, sin_theta;
};

void bsdf_setup(BSDF& b, V ray_dir, const Isect& is, const Scene& sc, bool from_light, float ior);
V bsdf_evaluate(const BSDF& b, V dir_gen, const Scene& sc, float& cos_gen, float* for_pdf, float* rev_pdf,
                float* sin_theta);
float bsdf_pdf(const BSDF& b, V dir_gen, const Scene& sc, bool reverse);
V bsdf_sample(const BSDF& b, float u0, float u1, float u2, V& dir_gen, float& pdf, float& cos_gen,
              const Scene& sc, unsigned& event, float& sin_theta);

// ---- lights: kernel/light.cuh ----------------------------------------------
V light_illuminate(const Light& l, const Scene& sc, V recv, float u0, float u1, V& dir_to_light, float& dist,
                   float& for_pdf_w, float* emission_pdf_w, float* cos_at_light);
V light_emit(const Light& l, const Scene& sc, float d0, float d1, float p0, float p1, V& pos, V& dir,
             float& emission_pdf_w, float* for_pdf_a, float* cos_light);
V light_radiance(const Light& l, const Scene& sc, V ray_dir, V hit, float* for_pdf_a, float* emission_pdf_w);
inline bool light_finite(const Light& l) { return l.type == L_AREA || l.type == L_POINT; }
inline bool light_delta(const Light& l) { return l.type == L_DIRECTIONAL || l.type == L_POINT; }

} // namespace orc

#endif
