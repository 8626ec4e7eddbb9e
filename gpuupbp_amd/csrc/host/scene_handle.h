// Owner of the arrays an upbp_scene points into (both scene loaders fill it).
#ifndef UPBP_HOST_SCENE_HANDLE_H
#define UPBP_HOST_SCENE_HANDLE_H

#include "upbp_capi.h"

#include <string>
#include <vector>

struct upbp_scene_handle {
    upbp_scene scene;
    std::vector<float> rv, rn, iv, in;        // triangle vertices / normals, real and imaginary
    std::vector<int32_t> rm, rl, im, il;      // triangle material / light ids
    std::vector<float> rs, is;                // spheres: centre, radius
    std::vector<int32_t> rsm, rsl, ism, isl;  // sphere material / light ids
    std::vector<upbp_material> mats;
    std::vector<upbp_medium> media;
    std::vector<upbp_light> lights;
    std::vector<float> env; // environment map RGB
    uint32_t env_w = 0, env_h = 0;

    // point the scene's views at the vectors (after they stop growing)
    void bind()
    {
        scene.real = {static_cast<uint32_t>(rm.size()), rv.data(), rn.data(), rm.data(), rl.data()};
        scene.imag = {static_cast<uint32_t>(im.size()), iv.data(), in.data(), im.data(), il.data()};
        scene.sph_real = {static_cast<uint32_t>(rsm.size()), rs.data(), rsm.data(), rsl.data()};
        scene.sph_imag = {static_cast<uint32_t>(ism.size()), is.data(), ism.data(), isl.data()};
        scene.mat_count = static_cast<uint32_t>(mats.size());
        scene.materials = mats.data();
        scene.med_count = static_cast<uint32_t>(media.size());
        scene.media = media.data();
        scene.light_count = static_cast<uint32_t>(lights.size());
        scene.lights = lights.data();
        scene.env_map = {env_w, env_h, env.empty() ? nullptr : env.data()};
    }
};

// error text of the last failed scene load on this thread (upbp_global_error)
void upbp_set_loader_error(const std::string& msg);

// OpenEXR into RGB floats (rows from the top); empty string on success, else the error
std::string upbp_read_exr_rgb(const std::string& path, std::vector<float>& rgb, uint32_t& width, uint32_t& height);

#endif
