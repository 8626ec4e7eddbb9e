// Oracle leaf functions: scene preparation, medium sampling, BSDF, lights.
// TEST INFRASTRUCTURE ONLY (see oracle.h). Citations: nolmoonen/gpuupbp.
#include "oracle_core.h"

#include <algorithm>

namespace orc {

static Mat4 mul(const Mat4& a, const Mat4& b)
{
    Mat4 r;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            float s = 0.f;
            for (int k = 0; k < 4; k++) s += a.m[i * 4 + k] * b.m[k * 4 + j];
            r.m[i * 4 + j] = s;
        }
    return r;
}

static Mat4 identity()
{
    Mat4 r;
    memset(r.m, 0, sizeof(r.m));
    r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.f;
    return r;
}

static Mat4 scale(float x, float y, float z)
{
    Mat4 r = identity();
    r.m[0] = x;
    r.m[5] = y;
    r.m[10] = z;
    return r;
}

static Mat4 translate(float x, float y, float z)
{
    Mat4 r = identity();
    r.m[3] = x;
    r.m[7] = y;
    r.m[11] = z;
    return r;
}

static Mat4 inverse(const Mat4& a)
{
    double m[4][8];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 8; j++) m[i][j] = j < 4 ? a.m[i * 4 + j] : (j - 4 == i ? 1.0 : 0.0);
    for (int c = 0; c < 4; c++) {
        int p = c;
        for (int r = c + 1; r < 4; r++)
            if (std::fabs(m[r][c]) > std::fabs(m[p][c])) p = r;
        for (int j = 0; j < 8; j++) std::swap(m[c][j], m[p][j]);
        const double inv = 1.0 / m[c][c];
        for (int j = 0; j < 8; j++) m[c][j] *= inv;
        for (int r = 0; r < 4; r++)
            if (r != c) {
                const double f = m[r][c];
                for (int j = 0; j < 8; j++) m[r][j] -= f * m[c][j];
            }
    }
    Mat4 o;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) o.m[i * 4 + j] = float(m[i][j + 4]);
    return o;
}

// shared/vec_math_ext.h:81-113 (homogeneous point transform)
V transform_point(const Mat4& mat, V p)
{
    const float* m = mat.m;
    float w = m[15];
    w += m[12] * p.x;
    w += m[13] * p.y;
    w += m[14] * p.z;
    const float inv = 1.f / w;
    V r;
    r.x = (m[3] + p.x * m[0] + p.y * m[1] + p.z * m[2]) * inv;
    r.y = (m[7] + p.x * m[4] + p.y * m[5] + p.z * m[6]) * inv;
    r.z = (m[11] + p.x * m[8] + p.y * m[9] + p.z * m[10]) * inv;
    return r;
}

// host/camera.hpp:33-76 with the perspective matrix of shared/vec_math_ext.h:46-74
static void init_camera(Camera& c, V origin, V target, V roll, float res_x, float res_y, float fov, int mat, int med)
{
    c.origin = origin;
    c.direction = normalize(target - origin);
    c.res_x = res_x;
    c.res_y = res_y;
    const V fwd = c.direction;
    const V up = normalize(cross(roll, -fwd));
    const V left = cross(-fwd, up);
    const V pos{dot(up, origin), dot(left, origin), dot(-fwd, origin)};
    Mat4 w2c = identity();
    const V rows[3] = {up, left, -fwd};
    const float t[3] = {-pos.x, -pos.y, -pos.z};
    for (int r = 0; r < 3; r++) {
        w2c.m[r * 4 + 0] = rows[r].x;
        w2c.m[r * 4 + 1] = rows[r].y;
        w2c.m[r * 4 + 2] = rows[r].z;
        w2c.m[r * 4 + 3] = t[r];
    }
    const float nearp = 0.1f, farp = 10000.f;
    const float f = 1.f / std::tan(fov * kPi / 360.f);
    const float d = 1.f / (nearp - farp);
    Mat4 persp;
    memset(persp.m, 0, sizeof(persp.m));
    persp.m[0] = f;
    persp.m[5] = -f;
    persp.m[10] = (nearp + farp) * d;
    persp.m[11] = 2.f * nearp * farp * d;
    persp.m[14] = -1.f;
    const Mat4 w2n = mul(persp, w2c);
    const Mat4 n2w = inverse(w2n);
    const float aspect = res_x / res_y;
    c.world_to_raster = mul(mul(scale(res_x * .5f, res_x * .5f, 0.f), translate(1.f, 1.f / aspect, 0.f)), w2n);
    c.raster_to_world = mul(mul(n2w, translate(-1.f, -1.f / aspect, 0.f)), scale(2.f / res_x, 2.f / res_x, 0.f));
    c.image_plane_dist = res_x / (2.f * std::tan(fov * kPi / 360.f));
    c.mat_id = mat;
    c.med_id = med;
}

// host/medium.hpp:39-85
static void init_medium(Medium& m, V absorption, V emission, V scattering, float cont, float g)
{
    m.cont_prob = cont;
    m.mean_cosine = g;
    auto absv = [](V a) { return V{std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)}; };
    m.absorption = absv(absorption);
    m.emission = absv(emission);
    m.scattering = absv(scattering);
    m.attenuation = m.absorption + m.scattering;
    m.min_pos_att = m.attenuation.x;
    m.min_pos_idx = 0;
    if (m.attenuation.y > 0.f && (m.min_pos_att == 0 || m.attenuation.y < m.min_pos_att)) {
        m.min_pos_att = m.attenuation.y;
        m.min_pos_idx = 1;
    }
    if (m.attenuation.z > 0.f && (m.min_pos_att == 0 || m.attenuation.z < m.min_pos_att)) {
        m.min_pos_att = m.attenuation.z;
        m.min_pos_idx = 2;
    }
    m.has_scattering = m.scattering.x > 0 || m.scattering.y > 0 || m.scattering.z > 0;
    m.has_attenuation = m.attenuation.x > 0 || m.attenuation.y > 0 || m.attenuation.z > 0;
    m.max_beam_length = m.has_attenuation ? -std::log(1e-9f) / maxc(m.attenuation) : kInf;
}

// ---- environment map: host/env_map.hpp:42-144, kernel/env_map.cuh:36-137, distribution.cuh ----
static void init_dist(Dist1D& d, const float* f, int n)
{
    d.func.assign(f, f + n);
    d.cdf.assign(size_t(n) + 1, 0.f);
    for (int i = 1; i < n + 1; ++i) d.cdf[size_t(i)] = d.cdf[size_t(i - 1)] + d.func[size_t(i - 1)] / float(n);
    d.integral = d.cdf[size_t(n)];
    for (int i = 1; i < n + 1; ++i)
        d.cdf[size_t(i)] = d.integral == 0.f ? float(i) / float(n) : d.cdf[size_t(i)] / d.integral;
}

void init_env_map(const upbp_env_map& in, EnvMap& m)
{
    m.width = int(in.width);
    m.height = int(in.height);
    m.img.resize(size_t(m.width) * size_t(m.height));
    for (size_t i = 0; i < m.img.size(); i++) m.img[i] = fromp(in.rgb + 3 * i);
    const int nv = m.height, nu = 2 * m.height; // height maps to pi, 2 * height to 2 pi
    std::vector<float> row(static_cast<size_t>(nu)), marg(static_cast<size_t>(nv));
    m.rows.resize(size_t(nv));
    for (int r = 0; r < nv; ++r) {
        const float v = (float(r) + .5f) / float(nv);
        const float st = std::sin(kPi * v);
        for (int c = 0; c < nu; ++c) {
            const V p = m.img[size_t(r) * size_t(m.width) + size_t(c)];
            row[size_t(c)] = st * (0.212671f * p.x + 0.715160f * p.y + 0.072169f * p.z);
        }
        init_dist(m.rows[size_t(r)], row.data(), nu);
        marg[size_t(r)] = m.rows[size_t(r)].integral;
    }
    init_dist(m.marginal, marg.data(), nv);
    m.norm = .5f * kInvPi * kInvPi;
}

static float sample_dist(const Dist1D& d, float u, float& pdf, int& off)
{
    const int n = int(d.func.size());
    // last cdf entry <= u
    off = int(std::upper_bound(d.cdf.begin(), d.cdf.end(), u) - d.cdf.begin()) - 1;
    off = std::max(0, std::min(off, n - 1));
    float du = u - d.cdf[size_t(off)];
    if (d.cdf[size_t(off + 1)] - d.cdf[size_t(off)] > 0.f) du /= d.cdf[size_t(off + 1)] - d.cdf[size_t(off)];
    pdf = d.integral > 0.f ? d.func[size_t(off)] / d.integral : 0.f;
    return (float(off) + du) / float(n);
}

static V env_radiance(const EnvMap& m, float u, float v)
{
    const float xf = u * float(m.width), yf = v * float(m.height);
    const int x1 = std::max(0, std::min(int(xf), m.width - 1)), y1 = std::max(0, std::min(int(yf), m.height - 1));
    const int x2 = x1 == m.width - 1 ? x1 : x1 + 1, y2 = y1 == m.height - 1 ? y1 : y1 + 1;
    const float tx = xf - float(x1), ty = yf - float(y1);
    auto at = [&](int x, int y) { return m.img[size_t(y) * size_t(m.width) + size_t(x)]; };
    return (1.f - ty) * ((1.f - tx) * at(x1, y1) + tx * at(x2, y1)) + ty * ((1.f - tx) * at(x1, y2) + tx * at(x2, y2));
}

static float env_sin_theta(float v, int height)
{
    const float h = float(height);
    return v < 1.f ? std::sin(kPi * (float(int(v * h)) + .5f) / h) : std::sin(kPi * ((h - 1.f) + .5f) / h);
}

V env_sample(const EnvMap& m, float u0, float u1, float& pdf_w, V& radiance)
{
    float pv, pu;
    int iv, iu;
    const float v = sample_dist(m.marginal, u1, pv, iv);
    const float u = sample_dist(m.rows[size_t(iv)], u0, pu, iu);
    pdf_w = m.norm * (pu * pv) / env_sin_theta(v, m.height);
    radiance = env_radiance(m, u, v);
    const float phi = u * 2.f * kPi, theta = v * kPi, st = std::sin(theta);
    return {-st * std::cos(phi), st * std::sin(phi), std::cos(theta)};
}

float env_pdf_uv(const EnvMap& m, float u, float v)
{
    const int nu = int(m.rows[0].func.size()), nv = int(m.marginal.func.size());
    const int iu = std::max(0, std::min(int(u * float(nu)), nu - 1));
    const int iv = std::max(0, std::min(int(v * float(nv)), nv - 1));
    const Dist1D& row = m.rows[size_t(iv)];
    if (row.integral * m.marginal.integral == 0.f) return 0.f;
    return row.func[size_t(iu)] * m.marginal.func[size_t(iv)] / (row.integral * m.marginal.integral);
}

V env_lookup(const EnvMap& m, V direction, float* pdf_w)
{
    const V d = normalize(direction);
    const float phi = (d.x != 0.f || d.y != 0.f) ? std::atan2(d.y, d.x) : 0.f;
    const float theta = std::acos(d.z);
    const float u = std::min(std::max(.5f - phi * .5f * kInvPi, 0.f), 1.f);
    const float v = std::min(std::max(theta * kInvPi, 0.f), 1.f);
    V rad = env_radiance(m, u, v);
    if (pdf_w) {
        *pdf_w = m.norm * env_pdf_uv(m, u, v) / env_sin_theta(v, m.height);
        if (*pdf_w == 0.f) rad = v3(0.f);
    }
    return rad;
}

Guide g_oracle_guide; // installed by oracle_set_guide_lobes, copied into every prepared scene

void prepare_scene(const upbp_scene* in, uint32_t res_x, uint32_t res_y, Scene& s)
{
    s.guide = g_oracle_guide;
    auto tris = [](const upbp_triangles& t, std::vector<Tri>& out) {
        out.resize(t.count);
        for (uint32_t i = 0; i < t.count; i++) {
            Tri& o = out[i];
            o.p0 = fromp(t.verts + 9 * i);
            o.p1 = fromp(t.verts + 9 * i + 3);
            o.p2 = fromp(t.verts + 9 * i + 6);
            o.n0 = fromp(t.normals + 12 * i);
            o.n1 = fromp(t.normals + 12 * i + 3);
            o.n2 = fromp(t.normals + 12 * i + 6);
            o.face = fromp(t.normals + 12 * i + 9);
            o.mat = t.mat_id[i];
            o.light = t.light_id[i];
        }
    };
    tris(in->real, s.real);
    tris(in->imag, s.imag);
    auto sphs = [](const upbp_spheres& t, std::vector<Sph>& out) {
        out.resize(t.count);
        for (uint32_t i = 0; i < t.count; i++)
            out[i] = {fromp(t.spheres + 4 * i), t.spheres[4 * i + 3], t.mat_id[i], t.light_id ? t.light_id[i] : -1};
    };
    sphs(in->sph_real, s.sreal);
    sphs(in->sph_imag, s.simag);
    s.mats.resize(in->mat_count);
    for (uint32_t i = 0; i < in->mat_count; i++) {
        const upbp_material& m = in->materials[i];
        s.mats[i] = {fromp(m.diffuse), fromp(m.phong), fromp(m.mirror), m.phong_exponent, m.ior, m.priority,
                     m.real != 0, m.med_idx};
    }
    s.media.resize(in->med_count);
    for (uint32_t i = 0; i < in->med_count; i++) {
        const upbp_medium& m = in->media[i];
        init_medium(s.media[i], fromp(m.absorption), fromp(m.emission), fromp(m.scattering), m.continuation_prob,
                    m.mean_cosine);
    }
    s.lights.resize(in->light_count);
    for (uint32_t i = 0; i < in->light_count; i++) {
        const upbp_light& l = in->lights[i];
        Light& o = s.lights[i];
        o = Light();
        o.type = l.type;
        o.mat_id = l.mat_id;
        o.med_id = l.med_id;
        o.intensity = fromp(l.intensity);
        if (l.type == UPBP_LIGHT_AREA) { // host/light.hpp:30-47
            o.p0 = fromp(l.p0);
            o.e1 = fromp(l.p1) - o.p0;
            o.e2 = fromp(l.p2) - o.p0;
            const V n = cross(o.e1, o.e2);
            o.inv_area = 2.f / length(n);
            o.frame.set_from_z(n);
        } else if (l.type == UPBP_LIGHT_DIRECTIONAL) { // host/light.hpp:60-68
            o.frame.set_from_z(fromp(l.position));
        } else if (l.type == UPBP_LIGHT_POINT) {
            o.position = fromp(l.position);
        } else {
            o.bg_color = fromp(l.intensity);
            o.scale = l.scale;
            o.uses_env = l.uses_env_map != 0 && in->env_map.rgb != nullptr;
        }
    }
    if (in->env_map.rgb && in->env_map.width) init_env_map(in->env_map, s.env);
    s.global_medium = in->global_medium_idx;
    s.background_light = in->background_light_idx;
    const upbp_camera& c = in->camera;
    init_camera(s.cam, fromp(c.origin), fromp(c.target), fromp(c.roll), float(res_x), float(res_y),
                c.horizontal_fov_deg, c.mat_id, c.med_id);
    s.center = fromp(in->scene_center);
    s.radius = in->scene_radius;
    s.inv_radius_sqr = 1.f / (s.radius * s.radius);
}

// kernel/medium.cuh:63-115
float sample_ray(const Medium& m, float dist_to_boundary, float rnd, float* for_pdf, unsigned flags, float* rev_pdf)
{
    if (m.min_pos_att != 0.f && m.has_scattering) {
        const float s = -std::log(rnd) / m.min_pos_att;
        if (s < dist_to_boundary) {
            const float att = std::exp(-m.min_pos_att * s);
            *for_pdf = m.min_pos_att * att;
            *rev_pdf = (flags & ORIGIN_IN_MEDIUM) ? *for_pdf : att;
            return s;
        }
        const float att = std::exp(-m.min_pos_att * dist_to_boundary);
        *for_pdf = att;
        *rev_pdf = (flags & ORIGIN_IN_MEDIUM) ? m.min_pos_att * att : att;
        return dist_to_boundary;
    }
    *for_pdf = 1.f;
    *rev_pdf = 1.f;
    return dist_to_boundary;
}

// kernel/medium.cuh:122-154
float ray_sample_pdf(const Medium& m, float dmin, float dmax, unsigned flags, float* rev_pdf)
{
    float for_pdf = 1.f;
    if (rev_pdf) *rev_pdf = 1.f;
    if (m.min_pos_att != 0.f && m.has_scattering) {
        const float att = std::fmax(std::exp(-m.min_pos_att * (dmax - dmin)), 1e-35f);
        const float minatt = std::fmax(m.min_pos_att * att, 1e-35f);
        for_pdf = (flags & END_IN_MEDIUM) ? minatt : att;
        if (rev_pdf) *rev_pdf = (flags & ORIGIN_IN_MEDIUM) ? minatt : att;
    }
    return for_pdf;
}

// kernel/bsdf.cuh:143-177
static void component_probabilities(BSDF& b, const Material& m)
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
    const V tmp = m.diffuse + m.phong + b.reflect_coeff * m.mirror;
    b.cont_prob = maxc(tmp) + (1.f - b.reflect_coeff);
    b.cont_prob = std::fmin(1.f, std::fmax(0.f, b.cont_prob));
}

// kernel/bsdf.cuh:179-223
void bsdf_setup(BSDF& b, V ray_dir, const Isect& is, const Scene& sc, bool from_light, float ior)
{
    b = BSDF();
    b.flags = from_light ? BSDF::FROM_LIGHT : 0;
    if (is.on_surface()) {
        const Material& m = sc.mats[size_t(is.mat)];
        b.mat = is.mat;
        b.frame.set_from_z(is.shading);
        b.local_dir_fix = b.frame.to_local(-ray_dir);
        b.geom_normal = b.frame.to_local(is.normal);
        b.cos_fix_geom = dot(b.local_dir_fix, b.geom_normal);
        if (std::fabs(b.local_dir_fix.z) < kEpsCosine || std::fabs(b.cos_fix_geom) < kEpsCosine) {
            b.flags = 0;
            return;
        }
        b.ior = ior;
        component_probabilities(b, m);
        if (b.diff_prob == 0.f && b.phong_prob == 0.f) b.flags |= BSDF::DELTA;
    } else {
        b.flags |= BSDF::IN_MEDIUM;
        const unsigned idx = unsigned(std::fmin(float(is.med), float(sc.media.size() - 1)));
        const Medium& m = sc.media[idx];
        b.med = int(idx);
        b.world_dir_fix = -ray_dir;
        b.frame.set_from_z(-b.world_dir_fix);
        b.cont_prob = m.cont_prob;
        b.mean_cosine = m.mean_cosine;
        b.scatter_coef = m.scattering;
        b.scatter_pos = !black_or_neg(m.scattering);
    }
    b.flags |= BSDF::VALID;
}

// kernel/bsdf.cuh:227-249
static V eval_diffuse(const BSDF& b, const Material& m, V lg, float* fp, float* rp)
{
    if (b.diff_prob == 0) return v3(0.f);
    if (b.local_dir_fix.z < kEpsCosine || lg.z < kEpsCosine) return v3(0.f);
    if (fp) *fp += b.diff_prob * std::fmax(0.f, lg.z * kInvPi);
    if (rp) *rp += b.diff_prob * std::fmax(0.f, b.local_dir_fix.z * kInvPi);
    return m.diffuse * kInvPi;
}

// kernel/bsdf.cuh:251-284
static V eval_phong(const BSDF& b, const Material& m, V lg, float* fp, float* rp)
{
    if (b.phong_prob == 0) return v3(0.f);
    if (b.local_dir_fix.z < kEpsCosine || lg.z < kEpsCosine) return v3(0.f);
    const V r = reflect_local(b.local_dir_fix);
    const float dr = dot(r, lg);
    if (dr <= kEpsPhong) return v3(0.f);
    const float pw = std::pow(dr, m.phong_exp);
    if (!positive(pw)) return v3(0.f);
    if (fp || rp) {
        const float pdf = b.phong_prob * power_cos_hemisphere_pdf_w(r, lg, m.phong_exp);
        if (fp) *fp += pdf;
        if (rp) *rp += pdf;
    }
    const V rho = m.phong * (m.phong_exp + 2.f) * .5f * kInvPi;
    return rho * pw;
}

// kernel/bsdf.cuh:290-326
static void pdf_diffuse(const BSDF& b, V lg, float* fp, float* rp)
{
    if (b.diff_prob == 0) return;
    if (fp) *fp += b.diff_prob * std::fmax(0.f, lg.z * kInvPi);
    if (rp) *rp += b.diff_prob * std::fmax(0.f, b.local_dir_fix.z * kInvPi);
}
static void pdf_phong(const BSDF& b, const Material& m, V lg, float* fp, float* rp)
{
    if (b.phong_prob == 0) return;
    const V r = reflect_local(b.local_dir_fix);
    const float dr = dot(r, lg);
    if (dr <= kEpsPhong) return;
    const float pdf = power_cos_hemisphere_pdf_w(r, lg, m.phong_exp) * b.phong_prob;
    if (fp) *fp += pdf;
    if (rp) *rp += pdf;
}

// ---- learned proposal (no reference counterpart) ------------------------------------------
void init_guide(Guide& g, const float* lobes, int res, const float* lo, const float* hi, float alpha)
{
    g = Guide();
    g.alpha = g.alpha_s = alpha;
    if (!(alpha > 0.f)) return;
    g.res = res;
    g.lo = fromp(lo);
    g.inv_cell = {float(res) / (hi[0] - lo[0]), float(res) / (hi[1] - lo[1]), float(res) / (hi[2] - lo[2])};
    const size_t n = size_t(res) * res * res;
    g.lobes.resize(n * 8);
    for (size_t r = 0; r < n; r++) {
        const float* in = lobes + r * 40;
        float wsum = 0.f;
        for (int i = 0; i < 8; i++) wsum += std::fmax(in[5 * i + 4], 0.f);
        for (int i = 0; i < 8; i++) {
            const float* l = in + 5 * i;
            GuideLobe& o = g.lobes[r * 8 + size_t(i)];
            const float len = std::sqrt(l[0] * l[0] + l[1] * l[1] + l[2] * l[2]);
            o.mu = len > 0.f ? V{l[0] / len, l[1] / len, l[2] / len} : V{0.f, 0.f, 1.f};
            o.kappa = std::fmin(std::fmax(l[3], 1e-2f), 1e4f);
            o.w = wsum > 0.f ? std::fmax(l[4], 0.f) / wsum : 1.f / 8.f;
            o.c = o.w * o.kappa / (2.f * kPi * -std::expm1(-2.f * o.kappa));
        }
    }
}

static int axis_cell(float f, int r) { return f < 0.f ? 0 : (f >= float(r) ? r - 1 : int(f)); }

const GuideLobe* guide_record(const Guide& g, V p)
{
    const int r = g.res;
    const int ix = axis_cell((p.x - g.lo.x) * g.inv_cell.x, r), iy = axis_cell((p.y - g.lo.y) * g.inv_cell.y, r),
              iz = axis_cell((p.z - g.lo.z) * g.inv_cell.z, r);
    return &g.lobes[size_t((iz * r + iy) * r + ix) * 8];
}

// the phase function of mean cosine g as three vMF lobes about the arrival direction (against
// it for g < 0): the fitted table's rows interpolated linearly in min(|g|, 0.99); a_j is the
// lobe weight over the lobe's sphere integral
static float sphere_integral(float k) { return k > 1e-6f ? 2.f * kPi * -std::expm1(-2.f * k) / k : 4.f * kPi; }

PhaseFit phase_fit(V arrive, float mean_cosine)
{
    PhaseFit h;
    h.mu = mean_cosine < 0.f ? -arrive : arrive;
    h.n = 3;
    const float t = std::fmin(std::fabs(mean_cosine), 0.99f) / kPhaseFitStep;
    const int row = std::min(int(t), kPhaseFitN - 2);
    const float f = t - float(row);
    for (int j = 0; j < 3; j++) {
        h.k[j] = (1.f - f) * kPhaseFit[row][j] + f * kPhaseFit[row + 1][j];
        h.a[j] = ((1.f - f) * kPhaseFit[row][3 + j] + f * kPhaseFit[row + 1][3 + j]) / sphere_integral(h.k[j]);
    }
    return h;
}

// the clamped cosine about a surface normal as its one fitted lobe
PhaseFit cosine_fit(V normal)
{
    PhaseFit h;
    h.mu = normal;
    h.n = 1;
    h.k[0] = kCosineKappa;
    h.a[0] = 1.f / sphere_integral(kCosineKappa);
    h.k[1] = h.k[2] = h.a[1] = h.a[2] = 0.f;
    return h;
}

// no mirror and no refraction: the BSDF's own samples are non-specular reflections
bool surface_guidable(const Material& m) { return luminance(m.mirror) == 0.f && !(m.ior > 0.f); }

// lobe (i, j) of the product: concentration |kappa_i mu_i + k_j mu_h| and its unnormalised mass
// c_i a_j 2 pi (exp(kp - kappa_i - k_j) - exp(-kp - kappa_i - k_j)) / kp, the second term dropped
// above kp = 8 where it is below 1e-7 of the first
static float prod_mass(const GuideLobe& l, const PhaseFit& h, int j)
{
    const float lh = dot(l.mu, h.mu);
    const float kp = std::sqrt(std::fmax(0.f, l.kappa * l.kappa + h.k[j] * h.k[j] + 2.f * l.kappa * h.k[j] * lh));
    const float s = l.kappa + h.k[j];
    const float m = kp > 8.f ? 2.f * kPi * std::exp(kp - s) / kp : std::exp(kp - s) * sphere_integral(kp);
    return l.c * h.a[j] * m;
}

// q(w) h(w) / Z
float guide_prod_pdf(const GuideLobe* rec, const PhaseFit& h, V w)
{
    float q = 0.f, hv = 0.f, z = 0.f;
    for (int j = 0; j < h.n; j++) hv += h.a[j] * std::exp(h.k[j] * (dot(h.mu, w) - 1.f));
    for (int i = 0; i < 8; i++) {
        q += rec[i].c * std::exp(rec[i].kappa * (dot(rec[i].mu, w) - 1.f));
        for (int j = 0; j < h.n; j++) z += prod_mass(rec[i], h, j);
    }
    return z > 0.f ? q * hv / z : 0.f;
}

// product lobe by mass with us; its direction by Jakob 2012 ("Numerically stable sampling of
// the von Mises Fisher distribution on S2") as cos = 1 + log1p((1 - u0) expm1(-2 k)) / k
V guide_prod_sample(const GuideLobe* rec, const PhaseFit& h, float us, float u0, float u1)
{
    float tot = 0.f;
    for (int i = 0; i < 8; i++)
        for (int j = 0; j < h.n; j++) tot += prod_mass(rec[i], h, j);
    int pi = 7, pj = h.n - 1;
    float acc = 0.f;
    bool found = false;
    for (int i = 0; i < 8 && !found; i++)
        for (int j = 0; j < h.n; j++) {
            acc += prod_mass(rec[i], h, j);
            if (us * tot < acc) {
                pi = i;
                pj = j;
                found = true;
                break;
            }
        }
    const GuideLobe& l = rec[pi];
    const V v = l.mu * l.kappa + h.mu * h.k[pj];
    const float kv = std::sqrt(dot(v, v));
    V mu{0.f, 0.f, 1.f};
    float k = 1e-6f;
    if (kv > 1e-6f) {
        k = kv;
        mu = v / kv;
    }
    const float ct = std::fmax(-1.f, 1.f + std::log1p((1.f - u0) * std::expm1(-2.f * k)) / k);
    const float st = std::sqrt(std::fmax(0.f, 1.f - ct * ct));
    const float phi = 2.f * kPi * u1;
    Frame f;
    f.set_from_z(mu);
    return f.to_world({std::cos(phi) * st, std::sin(phi) * st, ct});
}

float guide_pdf(const Guide& g, float mean_cosine, V p, V arrive, V gen)
{
    return guide_prod_pdf(guide_record(g, p), phase_fit(arrive, mean_cosine), gen);
}

float cam_medium_pdf(const Scene& sc, float mean_cosine, V p, V arrive, V gen)
{
    const float hg = hg_pdf(-arrive, gen, mean_cosine, nullptr);
    if (!(sc.guide.alpha > 0.f)) return hg;
    return sc.guide.alpha * guide_pdf(sc.guide, mean_cosine, p, arrive, gen) + (1.f - sc.guide.alpha) * hg;
}

// kernel/bsdf.cuh:334-396
V bsdf_evaluate(const BSDF& b, V dir_gen, const Scene& sc, float& cos_gen, float* for_pdf, float* rev_pdf,
                float* sin_theta)
{
    if (b.on_surface()) {
        V res = v3(0.f);
        if (for_pdf) *for_pdf = 0.f;
        if (rev_pdf) *rev_pdf = 0.f;
        if (sin_theta) *sin_theta = 0.f;
        const V lg = b.frame.to_local(dir_gen);
        const float cg = lg.z;
        const float cgg = dot(lg, b.geom_normal);
        cos_gen = std::fabs(cg);
        if (cos_gen < kEpsCosine || std::fabs(cgg) < kEpsCosine) return res;
        if (cg * b.local_dir_fix.z < 0.f || cgg * b.cos_fix_geom < 0.f) return res;
        const Material& m = sc.mats[size_t(b.mat)];
        res += eval_diffuse(b, m, lg, for_pdf, rev_pdf);
        res += eval_phong(b, m, lg, for_pdf, rev_pdf);
        // the camera technique's side follows the learned product at guidable surfaces: the
        // normal on the side the camera path leaves to (wo at a camera vertex, dir_gen at a
        // light vertex)
        const float a = sc.guide.alpha_s;
        if (a > 0.f && surface_guidable(m)) {
            const V nz = b.frame.z;
            if (!b.from_light() && for_pdf) {
                const V n = b.local_dir_fix.z < 0.f ? -nz : nz;
                *for_pdf = a * guide_prod_pdf(guide_record(sc.guide, b.pos), cosine_fit(n), dir_gen) + (1.f - a) * *for_pdf;
            } else if (b.from_light() && rev_pdf) {
                const V n = cg < 0.f ? -nz : nz;
                *rev_pdf = a * guide_prod_pdf(guide_record(sc.guide, b.pos), cosine_fit(n), b.dir_fix_world()) +
                           (1.f - a) * *rev_pdf;
            }
        }
        return res;
    }
    cos_gen = 1.f;
    if (!b.scatter_pos) {
        if (for_pdf) *for_pdf = 0.f;
        if (rev_pdf) *rev_pdf = 0.f;
        if (sin_theta) *sin_theta = 0.f;
        return v3(0.f);
    }
    const float p = hg_pdf(b.world_dir_fix, dir_gen, b.mean_cosine, sin_theta);
    float fp = p, rp = p; // the camera technique's side follows the learned mixture
    const float a = sc.guide.alpha;
    if (a > 0.f) {
        if (!b.from_light()) fp = a * guide_pdf(sc.guide, b.mean_cosine, b.pos, -b.world_dir_fix, dir_gen) + (1.f - a) * p;
        else rp = a * guide_pdf(sc.guide, b.mean_cosine, b.pos, -dir_gen, b.world_dir_fix) + (1.f - a) * p;
    }
    if (for_pdf) *for_pdf = fp;
    if (rev_pdf) *rev_pdf = rp;
    return b.scatter_coef * p;
}

// kernel/bsdf.cuh:402-447
float bsdf_pdf(const BSDF& b, V dir_gen, const Scene& sc, bool reverse)
{
    if (b.on_surface()) {
        const V lg = b.frame.to_local(dir_gen);
        const float cg = lg.z;
        const float cgg = dot(lg, b.geom_normal);
        if (std::fabs(cg) < kEpsCosine || std::fabs(cgg) < kEpsCosine) return 0.f;
        if (cg * b.local_dir_fix.z < 0.f || cgg * b.cos_fix_geom < 0.f) return 0.f;
        float fp = 0.f, rp = 0.f;
        const Material& m = sc.mats[size_t(b.mat)];
        pdf_diffuse(b, lg, &fp, &rp);
        pdf_phong(b, m, lg, &fp, &rp);
        const float a = sc.guide.alpha_s;
        const bool camera_technique = b.from_light() ? reverse : !reverse;
        if (a > 0.f && camera_technique && surface_guidable(m)) {
            const V nz = b.frame.z;
            if (!b.from_light())
                fp = a * guide_prod_pdf(guide_record(sc.guide, b.pos), cosine_fit(b.local_dir_fix.z < 0.f ? -nz : nz),
                                        dir_gen) + (1.f - a) * fp;
            else
                rp = a * guide_prod_pdf(guide_record(sc.guide, b.pos), cosine_fit(cg < 0.f ? -nz : nz), b.dir_fix_world()) +
                     (1.f - a) * rp;
        }
        return reverse ? rp : fp;
    }
    const float p = hg_pdf(b.world_dir_fix, dir_gen, b.mean_cosine, nullptr);
    const float a = sc.guide.alpha;
    // a camera vertex's forward pdf and a light vertex's reverse pdf belong to the camera
    // technique (the learned mixture); the other two to the phase function
    const bool camera_technique = b.from_light() ? reverse : !reverse;
    if (!(a > 0.f) || !camera_technique) return p;
    if (!b.from_light()) return a * guide_pdf(sc.guide, b.mean_cosine, b.pos, -b.world_dir_fix, dir_gen) + (1.f - a) * p;
    return a * guide_pdf(sc.guide, b.mean_cosine, b.pos, -dir_gen, b.world_dir_fix) + (1.f - a) * p;
}

// kernel/bsdf.cuh:455-559 (component samplers)
static V sample_diffuse(const BSDF& b, const Material& m, float u0, float u1, V& lg, float& pdf)
{
    if (b.local_dir_fix.z < kEpsCosine) return v3(0.f);
    float up;
    lg = sample_cos_hemisphere_w(u0, u1, &up);
    pdf += up * b.diff_prob;
    return m.diffuse * kInvPi;
}
static V sample_phong(const BSDF& b, const Material& m, float u0, float u1, V& lg, float& pdf)
{
    lg = sample_power_cos_hemisphere_w(u0, u1, m.phong_exp);
    const V r = reflect_local(b.local_dir_fix);
    {
        Frame f;
        f.set_from_z(r);
        lg = f.to_world(lg);
    }
    const float dr = dot(r, lg);
    if (dr <= kEpsPhong) return v3(0.f);
    pdf_phong(b, m, lg, &pdf, nullptr);
    const V rho = m.phong * (m.phong_exp + 2.f) * .5f * kInvPi;
    return rho * std::pow(dr, m.phong_exp);
}
static V sample_reflect(const BSDF& b, const Material& m, V& lg, float& pdf)
{
    lg = reflect_local(b.local_dir_fix);
    pdf += b.refl_prob;
    return b.reflect_coeff * m.mirror / std::fabs(lg.z);
}
static V sample_refract(const BSDF& b, V& lg, float& pdf)
{
    if (b.ior < 0) return v3(0.f);
    float cos_i = b.local_dir_fix.z;
    float cos_t;
    const float eta = b.ior;
    if (cos_i < 0.f) {
        cos_i = -cos_i;
        cos_t = 1.f;
    } else {
        cos_t = -1.f;
    }
    const float sin_i2 = 1.f - cos_i * cos_i;
    const float sin_t2 = eta * eta * sin_i2;
    if (sin_t2 < 1.f) {
        cos_t *= std::sqrt(std::fmax(0.f, 1.f - sin_t2));
        lg = {-eta * b.local_dir_fix.x, -eta * b.local_dir_fix.y, cos_t};
        pdf += b.refr_prob;
        const float rc = 1.f - b.reflect_coeff;
        if (!b.from_light()) return v3(rc * eta * eta / std::fabs(cos_t));
        return v3(rc / std::fabs(cos_t));
    }
    return v3(0.f);
}

// kernel/bsdf.cuh:567-660, surface part
static V bsdf_sample_surface(const BSDF& b, float u0, float u1, float u2, V& dir_gen, float& pdf, float& cos_gen,
                             const Scene& sc, unsigned& event, float& sin_theta)
{
    {
        unsigned ev;
        if (u2 < b.diff_prob) ev = BSDF::DIFFUSE;
        else if (u2 < b.diff_prob + b.phong_prob) ev = BSDF::PHONG;
        else if (u2 < b.diff_prob + b.phong_prob + b.refl_prob) ev = BSDF::REFLECT;
        else ev = BSDF::REFRACT;
        event = ev;
        sin_theta = 0.f;
        pdf = 0.f;
        V res = v3(0.f);
        V lg{0, 0, 0};
        const Material& m = sc.mats[size_t(b.mat)];
        if (ev == BSDF::DIFFUSE) {
            res += sample_diffuse(b, m, u0, u1, lg, pdf);
            if (black_or_neg(res)) return v3(0.f);
            res += eval_phong(b, m, lg, &pdf, nullptr);
        } else if (ev == BSDF::PHONG) {
            res += sample_phong(b, m, u0, u1, lg, pdf);
            if (black_or_neg(res)) return v3(0.f);
            res += eval_diffuse(b, m, lg, &pdf, nullptr);
        } else if (ev == BSDF::REFLECT) {
            res += sample_reflect(b, m, lg, pdf);
            if (black_or_neg(res)) return v3(0.f);
        } else {
            res += sample_refract(b, lg, pdf);
            if (black_or_neg(res)) return v3(0.f);
        }
        dir_gen = b.frame.to_world(lg);
        lg = b.frame.to_local(dir_gen);
        const float cg = lg.z;
        const float cgg = dot(lg, b.geom_normal);
        cos_gen = std::fabs(cg);
        if (cos_gen < kEpsCosine || std::fabs(cgg) < kEpsCosine) return v3(0.f);
        const bool refr = ev == BSDF::REFRACT;
        if ((refr && cg * b.local_dir_fix.z > 0.f) || (!refr && cg * b.local_dir_fix.z < 0.f) ||
            (refr && cgg * b.cos_fix_geom > 0.f) || (!refr && cgg * b.cos_fix_geom < 0.f))
            return v3(0.f);
        return res;
    }
}

// kernel/bsdf.cuh:567-660
V bsdf_sample(const BSDF& b, float u0, float u1, float u2, V& dir_gen, float& pdf, float& cos_gen,
              const Scene& sc, unsigned& event, float& sin_theta)
{
    if (b.on_surface()) {
        const float a = sc.guide.alpha_s;
        if (!(a > 0.f) || b.from_light() || !surface_guidable(sc.mats[size_t(b.mat)]))
            return bsdf_sample_surface(b, u0, u1, u2, dir_gen, pdf, cos_gen, sc, event, sin_theta);
        // camera path at a guidable surface: the learned product with the cosine lobe with
        // probability alpha_s (variate u2, rescaled for the BSDF's component choice otherwise)
        const V n = b.local_dir_fix.z < 0.f ? -b.frame.z : b.frame.z;
        const GuideLobe* rec = guide_record(sc.guide, b.pos);
        if (u2 < a) {
            dir_gen = guide_prod_sample(rec, cosine_fit(n), u2 / a, u0, u1);
            event = BSDF::DIFFUSE;
            sin_theta = 0.f;
            pdf = 0.f;
            return bsdf_evaluate(b, dir_gen, sc, cos_gen, &pdf, nullptr, nullptr);
        }
        const V f = bsdf_sample_surface(b, u0, u1, (u2 - a) / (1.f - a), dir_gen, pdf, cos_gen, sc, event, sin_theta);
        if (black_or_neg(f)) return f;
        pdf = a * guide_prod_pdf(rec, cosine_fit(n), dir_gen) + (1.f - a) * pdf;
        return f;
    }
    cos_gen = 1.f;
    event = BSDF::SCATTER;
    if (!b.scatter_pos) {
        dir_gen = v3(0.f);
        pdf = 0.f;
        sin_theta = 0.f;
        return v3(0.f);
    }
    const float a = sc.guide.alpha;
    if (!(a > 0.f) || b.from_light()) {
        dir_gen = hg_sample(b.world_dir_fix, b.mean_cosine, u0, u1, b.frame, pdf, &sin_theta);
        return b.scatter_coef * pdf;
    }
    // camera path: the learned product with probability alpha (variate u2), else the phase function
    const GuideLobe* rec = guide_record(sc.guide, b.pos);
    const PhaseFit h = phase_fit(-b.world_dir_fix, b.mean_cosine);
    float hg;
    if (u2 < a) {
        dir_gen = guide_prod_sample(rec, h, u2 / a, u0, u1);
        hg = hg_pdf(b.world_dir_fix, dir_gen, b.mean_cosine, &sin_theta);
    } else {
        dir_gen = hg_sample(b.world_dir_fix, b.mean_cosine, u0, u1, b.frame, hg, &sin_theta);
    }
    pdf = a * guide_prod_pdf(rec, h, dir_gen) + (1.f - a) * hg;
    return b.scatter_coef * hg;
}

// ---- lights: kernel/light.cuh:41-543 ---------------------------------------
V light_illuminate(const Light& l, const Scene& sc, V recv, float u0, float u1, V& dir, float& dist, float& for_pdf_w,
                   float* emission_pdf_w, float* cos_at_light)
{
    switch (l.type) {
    case L_AREA: { // :43-72
        float b0, b1;
        sample_uniform_triangle(u0, u1, b0, b1);
        const V p = l.p0 + l.e1 * b0 + l.e2 * b1;
        dir = p - recv;
        const float d2 = dot(dir, dir);
        dist = std::sqrt(d2);
        dir = dir / dist;
        const float cn = dot(l.frame.z, -dir);
        if (cn < kEpsCosine) return v3(0.f);
        for_pdf_w = l.inv_area * d2 / cn;
        if (cos_at_light) *cos_at_light = cn;
        if (emission_pdf_w) *emission_pdf_w = l.inv_area * cn * kInvPi;
        return l.intensity;
    }
    case L_BACKGROUND: { // :128-151
        V rad;
        if (l.uses_env) {
            dir = env_sample(sc.env, u0, u1, for_pdf_w, rad);
            rad = rad * l.scale;
        } else {
            dir = sample_uniform_sphere_w(u0, u1, &for_pdf_w);
            rad = l.bg_color * l.scale;
        }
        dist = kInf;
        if (emission_pdf_w) *emission_pdf_w = for_pdf_w * concentric_disc_pdf_a() * sc.inv_radius_sqr;
        if (cos_at_light) *cos_at_light = 1.f;
        return rad;
    }
    case L_DIRECTIONAL: { // :230-248
        dir = -l.frame.z;
        dist = kInf;
        for_pdf_w = 1.f;
        if (cos_at_light) *cos_at_light = 1.f;
        if (emission_pdf_w) *emission_pdf_w = concentric_disc_pdf_a() * sc.inv_radius_sqr;
        return l.intensity;
    }
    default: { // point, :301-317
        dir = l.position - recv;
        const float d2 = dot(dir, dir);
        for_pdf_w = d2;
        dist = std::sqrt(d2);
        dir = dir / dist;
        if (cos_at_light) *cos_at_light = 1.f;
        if (emission_pdf_w) *emission_pdf_w = uniform_sphere_pdf_w();
        return l.intensity;
    }
    }
}

V light_emit(const Light& l, const Scene& sc, float d0, float d1, float p0, float p1, V& pos, V& dir,
             float& emission_pdf_w, float* for_pdf_a, float* cos_light)
{
    switch (l.type) {
    case L_AREA: { // :74-99
        float b0, b1;
        sample_uniform_triangle(p0, p1, b0, b1);
        pos = l.p0 + l.e1 * b0 + l.e2 * b1;
        V lo = sample_cos_hemisphere_w(d0, d1, &emission_pdf_w);
        emission_pdf_w *= l.inv_area;
        lo.z = std::fmax(lo.z, kEpsCosine);
        dir = l.frame.to_world(lo);
        if (for_pdf_a) *for_pdf_a = l.inv_area;
        if (cos_light) *cos_light = lo.z;
        return l.intensity * lo.z;
    }
    case L_BACKGROUND: { // :153-190
        float fp;
        V rad;
        if (l.uses_env) {
            dir = -env_sample(sc.env, d0, d1, fp, rad);
            rad = rad * l.scale;
        } else {
            dir = sample_uniform_sphere_w(d0, d1, &fp);
            rad = l.bg_color * l.scale;
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
    case L_DIRECTIONAL: { // :250-272
        float x, y;
        sample_concentric_disc(p0, p1, x, y);
        pos = sc.center + sc.radius * (-l.frame.z + l.frame.x * x + l.frame.y * y);
        dir = l.frame.z;
        emission_pdf_w = concentric_disc_pdf_a() * sc.inv_radius_sqr;
        if (for_pdf_a) *for_pdf_a = 1.f;
        if (cos_light) *cos_light = 1.f;
        return l.intensity;
    }
    default: { // point, :319-337
        pos = l.position;
        dir = sample_uniform_sphere_w(d0, d1, &emission_pdf_w);
        if (for_pdf_a) *for_pdf_a = 1.f;
        if (cos_light) *cos_light = 1.f;
        return l.intensity;
    }
    }
}

V light_radiance(const Light& l, const Scene& sc, V ray_dir, V hit, float* for_pdf_a, float* emission_pdf_w)
{
    (void)hit;
    switch (l.type) {
    case L_AREA: { // :101-122
        const float c = std::fmax(0.f, dot(l.frame.z, -ray_dir));
        if (c == 0.f) return v3(0.f);
        if (for_pdf_a) *for_pdf_a = l.inv_area;
        if (emission_pdf_w) *emission_pdf_w = cos_hemisphere_pdf_w(l.frame.z, -ray_dir) * l.inv_area;
        return l.intensity;
    }
    case L_BACKGROUND: { // :192-217
        float fp;
        V rad;
        if (l.uses_env) {
            rad = l.scale * env_lookup(sc.env, ray_dir, &fp);
        } else {
            fp = uniform_sphere_pdf_w();
            rad = l.bg_color * l.scale;
        }
        const float pos_pdf = concentric_disc_pdf_a() * sc.inv_radius_sqr;
        if (for_pdf_a) *for_pdf_a = fp;
        if (emission_pdf_w) *emission_pdf_w = fp * pos_pdf;
        return rad;
    }
    default: return v3(0.f);
    }
}

// The roots of |o + t d - c| = r in FP32: b = o.d, c = o.o - r^2, t = -b -/+ sqrt(b^2 - c). When
// the near root lies more than a radius away, the solve is repeated from the point it gives and
// the two steps are added, as the reference does (gpuupbp/kernel/intersector.cu:56-75); the far
// root is then that point's far root plus the first step (:86). An all-hits caller takes both,
// a closest-hit caller the first.
int sphere_hit(const Sph& s, V ro, V rd, float tmin, float tmax, float* th, V* nh)
{
    const V o = ro - s.c;
    const float l = 1.f / std::sqrt(dot(rd, rd));
    const V d = rd * l;
    float b = dot(o, d), c = dot(o, o) - s.r * s.r, disc = b * b - c;
    if (!(disc > 0.f)) return 0;
    float sq = std::sqrt(disc);
    const float first = -b - sq;
    float second = 0.f;
    const bool again = std::fabs(first) > s.r;
    if (again) {
        const V p = o + first * d;
        b = dot(p, d);
        c = dot(p, p) - s.r * s.r;
        disc = b * b - c;
        if (disc > 0.f) {
            sq = std::sqrt(disc);
            second = -b - sq;
        }
    }
    int n = 0;
    const float tn = first + second;
    if (tn * l > tmin && tn * l < tmax) {
        th[n] = tn * l;
        nh[n++] = (o + tn * d) / s.r;
    }
    const float tf = (-b + sq) + (again ? first : 0.f);
    if (tf * l > tmin && tf * l < tmax) {
        th[n] = tf * l;
        nh[n++] = (o + tf * d) / s.r;
    }
    return n;
}

} // namespace orc
