// Predefined Cornell-box scenes (-s 0..40): the scene table of init_scene_configs
// (nolmoonen/gpuupbp gpuupbp/misc/config.cpp:100-518, enums scene_config.hpp:32-96) and the
// builder SceneLoader::load_predefined (gpuupbp/misc/scene_loader_predef.cpp:126-603).
//
// Every scene is a light, a global medium and a list of elements (geometry, optional material,
// optional medium). The builder follows the reference's rules:
//  * the camera is fixed (:133-139) and sits in the global medium;
//  * base materials 0..9 are shared by walls; a sphere or the large box gets a material of its
//    own keyed by priority (1 + material) * 10 + 1 + medium, created once per priority as a
//    copy of its base material (or a blank imaginary one) (:343-359);
//  * an element whose geometry was already placed is skipped (:330-336);
//  * a big ceiling light replaces the ceiling quad (:389-394); the small ceiling light is a
//    thin open box whose floor emits (:464-504);
//  * the scene sphere is the half-diagonal of the bounding box of all triangles and spheres
//    (scene_loader.hpp:65-71).
// Base materials keep medium index 0 unless a wall element sets it: the reference value-
// initialises its Material and reset_material leaves med_idx alone (material.hpp:32-41).
#include "scene_handle.h"

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace {

// material ids (scene_config.hpp:64-77)
enum PMat : int { M_NONE = -1, M_WATER, M_ICE, M_GLASS, M_RED, M_GREEN, M_BLUE, M_WHITE, M_GLOSSY, M_MIRROR,
                  M_LIGHT, M_COUNT };
// medium ids (scene_config.hpp:79-95)
enum PMed : int { D_NONE = -1, D_CLEAR, D_WATER, D_WHITE_ISO, D_WHITE_ANISO, D_WEAK_WHITE_ISO, D_WEAK_YELLOW_ISO,
                  D_WEAK_WHITE_ANISO, D_RED_ABS, D_YELLOW_EMIT, D_YG_ABS_EMIT_SCAT, D_BLUE_ABS_EMIT, D_REDDISH,
                  D_ABS_ANISO, D_COUNT };
// geometry ids (scene_config.hpp:46-62); the order matters: ids above G_FLOOR get own materials
enum PGeom : int { G_LEFT, G_RIGHT, G_FRONT_FACING_FRONT, G_BACK_FACING_FRONT, G_BACK, G_CEILING, G_FLOOR,
                   G_LARGE_SPHERE, G_SMALL_LEFT, G_SMALL_RIGHT, G_SMALL_BOTTOM_LEFT, G_SMALL_BOTTOM_RIGHT,
                   G_SMALL_TOP, G_VERY_LARGE_SPHERE, G_VERY_LARGE_BOX, G_COUNT };
// light ids (scene_config.hpp:33-44)
enum PLight : int { L_CEIL_BIG, L_CEIL_SMALL, L_CEIL_SMALL_DISTANT, L_CEIL_POINT, L_FACING_SMALL, L_FACING_POINT,
                    L_SUN, L_BACKGROUND, L_NONE };

struct Elem {
    int geom, mat, med;
};

struct SceneDef {
    const char* short_name;
    const char* long_name;
    int light, global_medium;
    int box;                     // 0 none, 1 diffuse floor, 2 glossy floor, 3 white back wall
    std::vector<Elem> extra;     // elements after the box
};

// the box variants (config.cpp:105-130)
std::vector<Elem> box_elements(int box)
{
    if (box == 0) return {};
    std::vector<Elem> e;
    e.push_back({G_LEFT, M_GREEN, D_NONE});
    e.push_back({G_RIGHT, M_RED, D_NONE});
    e.push_back({G_BACK, box == 3 ? M_WHITE : M_BLUE, D_NONE});
    e.push_back({G_CEILING, M_WHITE, D_NONE});
    e.push_back({G_FLOOR, box == 2 ? M_GLOSSY : M_WHITE, D_NONE});
    return e;
}

const std::vector<SceneDef>& scene_table()
{
    // three overlapping spheres: glass, water (with water medium), ice
    const std::vector<Elem> gwi = {{G_SMALL_BOTTOM_LEFT, M_GLASS, D_NONE},
                                   {G_SMALL_BOTTOM_RIGHT, M_WATER, D_WATER},
                                   {G_SMALL_TOP, M_ICE, D_NONE}};
    auto with = [](std::vector<Elem> a, std::vector<Elem> b) {
        a.insert(a.end(), b.begin(), b.end());
        return a;
    };
    const std::vector<Elem> spec = {{G_SMALL_LEFT, M_MIRROR, D_NONE}, {G_SMALL_RIGHT, M_GLASS, D_NONE}};
    const std::vector<Elem> iwr = {{G_SMALL_BOTTOM_LEFT, M_NONE, D_WHITE_ISO},
                                   {G_SMALL_BOTTOM_RIGHT, M_NONE, D_WATER},
                                   {G_SMALL_TOP, M_NONE, D_RED_ABS}};
    const std::vector<Elem> media_spheres = {{G_SMALL_LEFT, M_NONE, D_BLUE_ABS_EMIT},
                                             {G_SMALL_RIGHT, M_NONE, D_YG_ABS_EMIT_SCAT}};
    static const std::vector<SceneDef> t = {
        {"abssph", "large absorbing sphere + big area light", L_CEIL_BIG, D_CLEAR, 1,
         {{G_LARGE_SPHERE, M_NONE, D_RED_ABS}}},
        {"isosph", "large isoscattering sphere + small area light", L_CEIL_SMALL, D_CLEAR, 1,
         {{G_LARGE_SPHERE, M_NONE, D_WHITE_ISO}}},
        {"mirsphinfisov", "large mirror sphere + weak isoscattering + small area light", L_CEIL_SMALL,
         D_WEAK_WHITE_ISO, 2, {{G_LARGE_SPHERE, M_MIRROR, D_NONE}}},
        {"smswiso", "small media spheres + weak isoscattering + small area light", L_CEIL_SMALL, D_WEAK_WHITE_ISO,
         1, media_spheres},
        {"smswani", "small media spheres + weak anisoscattering + small area light", L_CEIL_SMALL,
         D_WEAK_WHITE_ANISO, 1, media_spheres},
        {"ssssun", "specular small spheres + sun", L_SUN, D_CLEAR, 2, spec},
        {"ssspoint", "specular small spheres + point light", L_CEIL_POINT, D_CLEAR, 2, spec},
        {"sssback", "specular small spheres + background light", L_BACKGROUND, D_CLEAR, 2, spec},
        {"mirsph", "large mirror sphere + small area light", L_CEIL_SMALL, D_CLEAR, 2,
         {{G_LARGE_SPHERE, M_MIRROR, D_NONE}}},
        {"emisph", "large emitting sphere + small area light", L_CEIL_SMALL, D_CLEAR, 1,
         {{G_LARGE_SPHERE, M_NONE, D_YELLOW_EMIT}}},
        {"gemisph", "large glass emitting sphere", L_NONE, D_CLEAR, 1, {{G_LARGE_SPHERE, M_GLASS, D_YELLOW_EMIT}}},
        {"nogeoisobig", "no geometry + isoscattering + big area light", L_CEIL_BIG, D_WHITE_ISO, 0, {}},
        {"nogeoisosmall", "no geometry + isoscattering + small area light", L_CEIL_SMALL, D_WHITE_ISO, 0, {}},
        {"nogeoisopt", "no geometry + isoscattering + point light", L_CEIL_POINT, D_WHITE_ISO, 0, {}},
        {"nogeoanibig", "no geometry + anisoscattering + big area light", L_CEIL_BIG, D_WHITE_ANISO, 0, {}},
        {"nogeoanismall", "no geometry + anisoscattering + small area light", L_CEIL_SMALL, D_WHITE_ANISO, 0, {}},
        {"nogeoanipt", "no geometry + anisoscattering + point light", L_CEIL_POINT, D_WHITE_ANISO, 0, {}},
        {"faceisosmall", "no geometry + isoscattering + facing small area light", L_FACING_SMALL, D_WHITE_ISO, 0,
         {}},
        {"faceisopt", "no geometry + isoscattering + facing point light", L_FACING_POINT, D_WHITE_ISO, 0, {}},
        {"faceanismall", "no geometry + anisoscattering + facing small area light", L_FACING_SMALL, D_WHITE_ANISO,
         0, {}},
        {"faceanipt", "no geometry + anisoscattering + facing point light", L_FACING_POINT, D_WHITE_ANISO, 0, {}},
        {"overgwi", "three overlapping spheres (glass, water, ice) + big area light", L_CEIL_BIG, D_CLEAR, 3, gwi},
        {"rovergwi", "three overlapping spheres (glass, water, ice) + one big red sphere + big area light",
         L_CEIL_BIG, D_CLEAR, 3, with(gwi, {{G_VERY_LARGE_SPHERE, M_NONE, D_RED_ABS}})},
        {"wovergwi", "three overlapping spheres (glass, water, ice) + one big water sphere + big area light",
         L_CEIL_BIG, D_CLEAR, 3, with(gwi, {{G_VERY_LARGE_SPHERE, M_WATER, D_WATER}})},
        {"iovergwi", "three overlapping spheres (glass, water, ice) + one big ice sphere + big area light",
         L_CEIL_BIG, D_CLEAR, 3, with(gwi, {{G_VERY_LARGE_SPHERE, M_ICE, D_NONE}})},
        {"govergwi", "three overlapping spheres (glass, water, ice) + one big glass sphere + big area light",
         L_CEIL_BIG, D_CLEAR, 3, with(gwi, {{G_VERY_LARGE_SPHERE, M_GLASS, D_NONE}})},
        {"movergwi", "three overlapping spheres (glass, water, ice) + one big mirror sphere + big area light",
         L_CEIL_BIG, D_CLEAR, 3, with(gwi, {{G_VERY_LARGE_SPHERE, M_MIRROR, D_NONE}})},
        {"aniovergwi", "three overlapping spheres (glass, water, ice) + weak anisoscattering + small area light",
         L_CEIL_SMALL, D_WEAK_WHITE_ANISO, 3, gwi},
        {"overiwr", "three overlapping glass spheres (iso, water, red absorb) + big area light", L_CEIL_BIG, D_CLEAR,
         3,
         {{G_SMALL_BOTTOM_LEFT, M_GLASS, D_WHITE_ISO},
          {G_SMALL_BOTTOM_RIGHT, M_GLASS, D_WATER},
          {G_SMALL_TOP, M_GLASS, D_RED_ABS}}},
        {"isooveriwr", "three overlapping spheres (iso, water, red absorb) + weak isoscattering + big area light",
         L_CEIL_BIG, D_WEAK_WHITE_ISO, 0, iwr},
        {"woveriwr",
         "three overlapping spheres (iso, water, red absorb) + one big water sphere + background light",
         L_BACKGROUND, D_CLEAR, 3, with(iwr, {{G_VERY_LARGE_SPHERE, M_NONE, D_WATER}})},
        {"backg", "only background light and global medium", L_BACKGROUND, D_REDDISH, 0, {}},
        {"backas", "only background light and absorbing sphere", L_BACKGROUND, D_CLEAR, 0,
         {{G_VERY_LARGE_SPHERE, M_NONE, D_RED_ABS}}},
        {"backss", "only background light and scattering sphere", L_BACKGROUND, D_CLEAR, 0,
         {{G_VERY_LARGE_SPHERE, M_NONE, D_WHITE_ISO}}},
        {"bigss", "only big area light and scattering sphere", L_CEIL_BIG, D_CLEAR, 0,
         {{G_VERY_LARGE_SPHERE, M_NONE, D_WHITE_ISO}}},
        {"smallsb", "only small area light and scattering box", L_CEIL_SMALL_DISTANT, D_CLEAR, 0,
         {{G_VERY_LARGE_BOX, M_NONE, D_WEAK_YELLOW_ISO}}},
        {"ssssuninfisov", "specular small spheres + weak anisoscattering + sun", L_SUN, D_WEAK_WHITE_ISO, 2, spec},
        {"ssspointinfisov", "specular small spheres + weak anisoscattering + point light", L_CEIL_POINT,
         D_WEAK_WHITE_ISO, 2, spec},
        {"sssbackinfisov", "specular small spheres + weak anisoscattering + background light", L_BACKGROUND,
         D_WEAK_WHITE_ISO, 2, spec},
        {"glasssphbck", "large glass medium sphere + background light", L_BACKGROUND, D_CLEAR, 2,
         {{G_LARGE_SPHERE, M_GLASS, D_ABS_ANISO}}},
        {"glasssphbck", "large glass medium sphere + large area light", L_CEIL_BIG, D_CLEAR, 2,
         {{G_LARGE_SPHERE, M_GLASS, D_ABS_ANISO}}},
    };
    return t;
}

constexpr float kInvPiF = 0.318309886183790671538f;

struct P3 {
    float x, y, z;
};
P3 operator+(P3 a, P3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
P3 operator-(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
P3 operator*(P3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
P3 cross3(P3 a, P3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
P3 unit(P3 a)
{
    const float l = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    return {a.x / l, a.y / l, a.z / l};
}

upbp_material blank_material() // reset_material (material.hpp:32-41) over a value-initialised one
{
    upbp_material m;
    memset(&m, 0, sizeof(m));
    m.phong_exponent = 1.f;
    m.ior = -1.f;
    m.priority = -1;
    m.real = 1;
    m.med_idx = 0;
    return m;
}

void set3(float* d, float x, float y, float z)
{
    d[0] = x;
    d[1] = y;
    d[2] = z;
}

upbp_medium medium(P3 absorption, P3 emission, P3 scattering, float cont, float g = 0.f)
{
    upbp_medium m;
    memset(&m, 0, sizeof(m));
    set3(m.absorption, absorption.x, absorption.y, absorption.z);
    set3(m.emission, emission.x, emission.y, emission.z);
    set3(m.scattering, scattering.x, scattering.y, scattering.z);
    m.continuation_prob = cont;
    m.mean_cosine = g;
    return m;
}

struct Builder {
    upbp_scene_handle* h;
    P3 bmin{1e36f, 1e36f, 1e36f}, bmax{-1e36f, -1e36f, -1e36f};

    void grow(P3 p)
    {
        bmin = {std::fmin(bmin.x, p.x), std::fmin(bmin.y, p.y), std::fmin(bmin.z, p.z)};
        bmax = {std::fmax(bmax.x, p.x), std::fmax(bmax.y, p.y), std::fmax(bmax.z, p.z)};
    }
    // flat-shaded triangle (PredefTriangle, scene_loader_predef.cpp:33-56)
    void tri(bool real, P3 a, P3 b, P3 c, int mat, int light = -1)
    {
        const P3 n = unit(cross3(b - a, c - a));
        std::vector<float>& V = real ? h->rv : h->iv;
        std::vector<float>& N = real ? h->rn : h->in;
        for (P3 p : {a, b, c}) {
            V.insert(V.end(), {p.x, p.y, p.z});
            grow(p);
        }
        for (int k = 0; k < 4; k++) N.insert(N.end(), {n.x, n.y, n.z});
        (real ? h->rm : h->im).push_back(mat);
        (real ? h->rl : h->il).push_back(light);
    }
    void sphere(bool real, P3 c, float r, int mat)
    {
        (real ? h->rs : h->is).insert((real ? h->rs : h->is).end(), {c.x, c.y, c.z, r});
        (real ? h->rsm : h->ism).push_back(mat);
        (real ? h->rsl : h->isl).push_back(-1);
        grow(c - P3{r, r, r});
        grow(c + P3{r, r, r});
    }
    void area_light(P3 a, P3 b, P3 c, float e)
    {
        upbp_light l;
        memset(&l, 0, sizeof(l));
        l.type = UPBP_LIGHT_AREA;
        l.mat_id = -1;
        l.med_id = -1;
        set3(l.p0, a.x, a.y, a.z);
        set3(l.p1, b.x, b.y, b.z);
        set3(l.p2, c.x, c.y, c.z);
        set3(l.intensity, e, e, e);
        h->lights.push_back(l);
    }
    void other_light(int type, P3 pos, P3 e, float scale = 0.f)
    {
        upbp_light l;
        memset(&l, 0, sizeof(l));
        l.type = type;
        l.mat_id = -1;
        l.med_id = -1;
        set3(l.position, pos.x, pos.y, pos.z);
        set3(l.intensity, e.x, e.y, e.z);
        l.scale = scale;
        h->lights.push_back(l);
    }
};

} // namespace

extern "C" int32_t upbp_predefined_scene_count(void) { return static_cast<int32_t>(scene_table().size()); }

extern "C" const char* upbp_predefined_scene_name(int32_t id, int32_t long_name)
{
    if (id < 0 || id >= upbp_predefined_scene_count()) return nullptr;
    const SceneDef& d = scene_table()[size_t(id)];
    return long_name ? d.long_name : d.short_name;
}

extern "C" int upbp_scene_load_predefined(int32_t scene_id, uint32_t res_x, uint32_t res_y, upbp_scene_handle** out)
{
    (void)res_x;
    (void)res_y;
    if (!out) {
        upbp_set_loader_error("null argument");
        return UPBP_E_ARG;
    }
    *out = nullptr;
    if (scene_id < 0 || scene_id >= upbp_predefined_scene_count()) {
        upbp_set_loader_error("predefined scene id out of range");
        return UPBP_E_ARG;
    }
    const SceneDef& def = scene_table()[size_t(scene_id)];
    auto* h = new upbp_scene_handle();
    upbp_scene& s = h->scene;
    memset(&s, 0, sizeof(s));
    s.background_light_idx = -1;
    Builder B{h};

    // camera (scene_loader_predef.cpp:133-139)
    set3(s.camera.origin, -0.0439815f, -4.12529f, 0.222539f);
    set3(s.camera.target, -0.03709525f, -3.126785f, 0.1683229f);
    set3(s.camera.roll, 3.73896e-4f, 0.0542148f, 0.998529f);
    s.camera.horizontal_fov_deg = 45.f;
    s.camera.focal_dist = 1.f;
    s.camera.mat_id = -1;
    s.camera.med_id = -1;

    // base materials (:143-197)
    h->mats.assign(M_COUNT, blank_material());
    set3(h->mats[M_RED].diffuse, 0.803922f, 0.152941f, 0.152941f);
    set3(h->mats[M_GREEN].diffuse, 0.156863f, 0.803922f, 0.172549f);
    set3(h->mats[M_BLUE].diffuse, 0.156863f, 0.172549f, 0.803922f);
    set3(h->mats[M_WHITE].diffuse, 0.803922f, 0.803922f, 0.803922f);
    set3(h->mats[M_GLOSSY].diffuse, 0.1f, 0.1f, 0.1f);
    set3(h->mats[M_GLOSSY].phong, 0.7f, 0.7f, 0.7f);
    h->mats[M_GLOSSY].phong_exponent = 90.f;
    set3(h->mats[M_MIRROR].mirror, 1.f, 1.f, 1.f);
    set3(h->mats[M_WATER].mirror, 0.7f, 0.7f, 0.7f);
    h->mats[M_WATER].ior = 1.33f;
    set3(h->mats[M_ICE].mirror, 0.5f, 0.5f, 0.5f);
    h->mats[M_ICE].ior = 1.31f;
    set3(h->mats[M_GLASS].mirror, 1.f, 1.f, 1.f);
    h->mats[M_GLASS].ior = 1.6f;

    // media (:200-293); survival probability 0.8 (host/medium.hpp:37), clear 1
    const P3 z{0.f, 0.f, 0.f};
    h->media.resize(D_COUNT);
    h->media[D_CLEAR] = medium(z, z, z, 1.f);
    h->media[D_WATER] = medium({0.7f, 0.6f, 0.f}, z, z, .8f);
    h->media[D_RED_ABS] = medium({0.f, 1.f, 1.f}, z, z, .8f);
    h->media[D_YELLOW_EMIT] = medium(z, {0.7f, 0.7f, 0.f}, z, .8f);
    h->media[D_YG_ABS_EMIT_SCAT] = medium({0.5f, 0.f, 0.5f}, {1.f, 1.f, 0.f}, {0.1f, 0.1f, 0.f}, .8f);
    h->media[D_BLUE_ABS_EMIT] = medium({0.1f, 0.1f, 0.f}, {0.f, 0.f, 1.f}, z, .8f);
    h->media[D_WHITE_ISO] = medium(z, z, {0.9f, 0.9f, 0.9f}, .8f);
    h->media[D_WHITE_ANISO] = medium(z, z, {0.9f, 0.9f, 0.9f}, .8f, 0.6f);
    h->media[D_WEAK_WHITE_ISO] = medium(z, z, {0.1f, 0.1f, 0.1f}, .8f);
    h->media[D_WEAK_YELLOW_ISO] = medium(z, z, {0.1f, 0.1f, 0.f}, .8f);
    h->media[D_WEAK_WHITE_ANISO] = medium(z, z, {0.1f, 0.1f, 0.1f}, .8f, 0.6f);
    h->media[D_REDDISH] = medium({0.f, 0.02f, 0.f}, z, {0.002f, 0.002f, 0.f}, .8f, 0.6f);
    h->media[D_ABS_ANISO] = medium({0.02f, 2.f, 2.f}, z, {12.f, 20.f, 20.f}, .8f, -0.3f);
    s.global_medium_idx = static_cast<uint32_t>(def.global_medium);

    // box corners and derived sphere placements (:296-326)
    const P3 cb[8] = {{-1.27029f, 1.30455f, -1.28002f}, {1.28975f, 1.30455f, -1.28002f},
                      {1.28975f, 1.30455f, 1.28002f},   {-1.27029f, 1.30455f, 1.28002f},
                      {-1.27029f, -1.25549f, -1.28002f}, {1.28975f, -1.25549f, -1.28002f},
                      {1.28975f, -1.25549f, 1.28002f},  {-1.27029f, -1.25549f, 1.28002f}};
    const P3 floor_c = (cb[0] + cb[1] + cb[4] + cb[5]) * (1.f / 4.f);
    const P3 box_c = (cb[0] + cb[1] + cb[2] + cb[3] + cb[4] + cb[5] + cb[6] + cb[7]) * (1.f / 8.f);
    const float r_large = 0.8f, r_small = 0.5f, r_very_large = 1.f;
    const P3 large_c = floor_c + P3{0, 0, r_large + 0.01f};
    const P3 left_wall_c = (cb[0] + cb[4]) * (1.f / 2.f) + P3{0, 0, r_small + 0.01f};
    const P3 right_wall_c = (cb[1] + cb[5]) * (1.f / 2.f) + P3{0, 0, r_small + 0.01f};
    const float xlen = right_wall_c.x - left_wall_c.x;
    const P3 small_left_c = left_wall_c + P3{2.f * xlen / 7.f, 0, 0};
    const P3 small_right_c = right_wall_c - P3{2.f * xlen / 7.f, 0, 0};
    const float overlap = 0.5f * r_small;
    const P3 bottom_left_c = box_c + P3{-overlap, 0, -overlap};
    const P3 bottom_right_c = box_c + P3{overlap, 0, -overlap};
    const P3 top_c = box_c + P3{0, 0, overlap};

    std::vector<Elem> elems = box_elements(def.box);
    elems.insert(elems.end(), def.extra.begin(), def.extra.end());
    std::vector<bool> used(G_COUNT, false);
    for (const Elem& e : elems) {
        if (used[size_t(e.geom)]) continue;
        used[size_t(e.geom)] = true;
        const bool real = e.mat >= 0;
        int mat_id = e.mat;
        if (e.geom > G_FLOOR) { // own material per (material, medium) priority (:343-359)
            const int priority = (1 + e.mat) * 10 + 1 + e.med;
            mat_id = M_COUNT;
            while (size_t(mat_id) < h->mats.size() && h->mats[size_t(mat_id)].priority != priority) mat_id++;
            if (size_t(mat_id) == h->mats.size()) {
                upbp_material m = real ? h->mats[size_t(e.mat)] : blank_material();
                m.priority = priority;
                m.real = real ? 1 : 0;
                h->mats.push_back(m);
            }
        }
        h->mats[size_t(mat_id)].med_idx = e.med;
        switch (e.geom) {
        case G_LEFT:
            B.tri(real, cb[3], cb[7], cb[4], mat_id);
            B.tri(real, cb[4], cb[0], cb[3], mat_id);
            break;
        case G_RIGHT:
            B.tri(real, cb[1], cb[5], cb[6], mat_id);
            B.tri(real, cb[6], cb[2], cb[1], mat_id);
            break;
        case G_FRONT_FACING_FRONT:
            B.tri(real, cb[4], cb[5], cb[6], mat_id);
            B.tri(real, cb[6], cb[7], cb[4], mat_id);
            break;
        case G_BACK_FACING_FRONT:
            B.tri(real, cb[6], cb[5], cb[4], mat_id);
            B.tri(real, cb[4], cb[7], cb[6], mat_id);
            break;
        case G_BACK:
            B.tri(real, cb[0], cb[1], cb[2], mat_id);
            B.tri(real, cb[2], cb[3], cb[0], mat_id);
            break;
        case G_CEILING:
            if (def.light != L_CEIL_BIG) {
                B.tri(real, cb[2], cb[6], cb[7], mat_id);
                B.tri(real, cb[7], cb[3], cb[2], mat_id);
            }
            break;
        case G_FLOOR:
            B.tri(real, cb[0], cb[4], cb[5], mat_id);
            B.tri(real, cb[5], cb[1], cb[0], mat_id);
            break;
        case G_LARGE_SPHERE: B.sphere(real, large_c, r_large, mat_id); break;
        case G_SMALL_LEFT: B.sphere(real, small_left_c, r_small, mat_id); break;
        case G_SMALL_RIGHT: B.sphere(real, small_right_c, r_small, mat_id); break;
        case G_SMALL_BOTTOM_LEFT: B.sphere(real, bottom_left_c, r_small, mat_id); break;
        case G_SMALL_BOTTOM_RIGHT: B.sphere(real, bottom_right_c, r_small, mat_id); break;
        case G_SMALL_TOP: B.sphere(real, top_c, r_small, mat_id); break;
        case G_VERY_LARGE_SPHERE: B.sphere(real, box_c, r_very_large, mat_id); break;
        case G_VERY_LARGE_BOX: { // the whole box, inward faces (:430-444)
            const int q[12][3] = {{7, 3, 4}, {0, 4, 3}, {5, 1, 6}, {2, 6, 1}, {4, 5, 6}, {6, 7, 4},
                                  {1, 0, 2}, {3, 2, 0}, {6, 2, 7}, {3, 7, 2}, {4, 0, 5}, {1, 5, 0}};
            for (const auto& t : q) B.tri(real, cb[t[0]], cb[t[1]], cb[t[2]], mat_id);
        } break;
        default: break;
        }
    }

    // the light (:448-530)
    const float small_e = 25.03329895614464f;
    switch (def.light) {
    case L_CEIL_BIG:
        B.area_light(cb[2], cb[6], cb[7], 0.95492965f);
        B.area_light(cb[7], cb[3], cb[2], 0.95492965f);
        B.tri(true, cb[2], cb[6], cb[7], M_LIGHT, 0);
        B.tri(true, cb[7], cb[3], cb[2], M_LIGHT, 1);
        break;
    case L_CEIL_SMALL:
    case L_CEIL_SMALL_DISTANT: {
        const float y = def.light == L_CEIL_SMALL ? 0.f : -1.5f;
        const P3 lb[8] = {{-0.25f, 0.25f + y, 1.26002f},  {0.25f, 0.25f + y, 1.26002f},
                          {0.25f, 0.25f + y, 1.28002f},   {-0.25f, 0.25f + y, 1.28002f},
                          {-0.25f, -0.25f + y, 1.26002f}, {0.25f, -0.25f + y, 1.26002f},
                          {0.25f, -0.25f + y, 1.28002f},  {-0.25f, -0.25f + y, 1.28002f}};
        B.area_light(lb[0], lb[5], lb[4], small_e);
        B.area_light(lb[5], lb[0], lb[1], small_e);
        if (def.light == L_CEIL_SMALL) { // the lamp housing: back, left, right, front walls
            const int q[8][3] = {{0, 2, 1}, {2, 0, 3}, {3, 4, 7}, {4, 3, 0},
                                 {1, 6, 5}, {6, 1, 2}, {4, 5, 6}, {6, 7, 4}};
            for (const auto& t : q) B.tri(true, lb[t[0]], lb[t[1]], lb[t[2]], M_WHITE);
        }
        B.tri(true, lb[0], lb[5], lb[4], M_LIGHT, 0);
        B.tri(true, lb[5], lb[0], lb[1], M_LIGHT, 1);
    } break;
    case L_CEIL_POINT:
        B.other_light(UPBP_LIGHT_POINT, {0.f, -0.5f, 1.f}, P3{1.f, 1.f, 1.f} * (70.f * (kInvPiF * 0.25f)));
        break;
    case L_FACING_SMALL: {
        const P3 lr[4] = {{-0.25f, 0, 0.25f}, {0.25f, 0, 0.25f}, {-0.25f, 0, -0.25f}, {0.25f, 0, -0.25f}};
        B.area_light(lr[2], lr[3], lr[0], small_e);
        B.area_light(lr[1], lr[0], lr[3], small_e);
        B.tri(true, lr[2], lr[3], lr[0], M_LIGHT, 0);
        B.tri(true, lr[1], lr[0], lr[3], M_LIGHT, 1);
    } break;
    case L_FACING_POINT:
        B.other_light(UPBP_LIGHT_POINT, {0.f, 0.f, 0.f}, P3{1.f, 1.f, 1.f} * (70.f * (kInvPiF * 0.25f)));
        break;
    case L_SUN: B.other_light(UPBP_LIGHT_DIRECTIONAL, {-1.f, 1.5f, -1.f}, P3{0.5f, 0.2f, 0.f} * 20.f); break;
    case L_BACKGROUND:
        B.other_light(UPBP_LIGHT_BACKGROUND, z, P3{135.f / 255.f, 206.f / 255.f, 250.f / 255.f}, 1.f);
        s.background_light_idx = 0;
        break;
    default: break;
    }

    // scene sphere (scene_loader.hpp:65-71)
    const P3 diag = B.bmax - B.bmin;
    float r2 = diag.x * diag.x + diag.y * diag.y + diag.z * diag.z;
    if (r2 == 0.f || r2 > 1e36f) r2 = 1.f;
    const P3 c = (B.bmax + B.bmin) * .5f;
    set3(s.scene_center, c.x, c.y, c.z);
    s.scene_radius = std::sqrt(r2) * .5f;
    h->bind();
    *out = h;
    return UPBP_OK;
}
