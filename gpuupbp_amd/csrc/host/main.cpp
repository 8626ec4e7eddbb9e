// gpuupbp command-line driver over libupbp: the drop-in for the reference's main
// (nolmoonen/gpuupbp gpuupbp/gpuupbp.cpp:57-148). Same options (upbp_parse_commandline),
// same iteration- or time-based loop with percentage lines and continuous output
// (gpuupbp.cpp:33-55), same final image. Times are wall-clock (std::chrono) rather
// than the reference's clock() CPU time, so -t means seconds of rendering.
#include "upbp_capi.h"

#include <chrono>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

namespace {

std::string describe(const upbp_config& c)
{
    std::string s = "techniques from upbp:";
    const struct {
        uint32_t bit;
        const char* name;
    } tech[] = {{UPBP_BPT, "bpt"},   {UPBP_SURF, "surf"}, {UPBP_PP3D, "pp3d"},
                {UPBP_PB2D, "pb2d"}, {UPBP_BB1D, "bb1d"}, {UPBP_BP2D, "bp2d"}};
    for (const auto& t : tech)
        if (c.algorithm_flags & t.bit) s += std::string(" ") + t.name;
    char line[160];
    snprintf(line, sizeof(line), "\n          resolution:        %dx%d\n          max path length:   %u\n", c.res_x,
             c.res_y, c.max_path_length);
    s += line;
    const struct {
        uint32_t bit;
        const char* name;
        float r0, alpha;
    } rad[] = {{UPBP_BB1D, "bb1d", c.bb1d_radius_initial, c.bb1d_radius_alpha},
               {UPBP_PP3D, "pp3d", c.pp3d_radius_initial, c.pp3d_radius_alpha},
               {UPBP_SURF, "surf", c.surf_radius_initial, c.surf_radius_alpha},
               {UPBP_PB2D, "pb2d", c.pb2d_radius_initial, c.pb2d_radius_alpha},
               {UPBP_BP2D, "bp2d", c.bp2d_radius_initial, c.bp2d_radius_alpha}};
    for (const auto& r : rad) {
        if (!(c.algorithm_flags & r.bit)) continue;
        snprintf(line, sizeof(line), "          %s radius init:  %g\n               radius alpha: %g\n", r.name,
                 double(r.r0), double(r.alpha));
        s += line;
    }
    return s;
}

// mean image so far (FP64 on the device) -> FP32 rows -> file
int save_mean(upbp_renderer* r, const upbp_config& c, int iters, const std::string& path)
{
    const size_t n = size_t(c.res_x) * size_t(c.res_y) * 3;
    std::vector<double> acc(n);
    int e = upbp_read_accum(r, acc.data(), iters);
    if (e) return e;
    std::vector<float> rgb(n);
    for (size_t i = 0; i < n; i++) rgb[i] = float(acc[i]);
    e = upbp_save_image(path.c_str(), rgb.data(), uint32_t(c.res_x), uint32_t(c.res_y));
    if (e == UPBP_E_ARG) fprintf(stderr, "Error: used unknown extension of %s\n", path.c_str());
    return e;
}

// -log: one CSV line per iteration into <timestamp>_log.txt with the columns of the
// reference's Logger (gpuupbp/misc/logger.cpp:31-86); times in seconds (HIP events, wall
// clock), memory in MiB, the two utilisation columns 0 as in a non-debug reference build.
struct CsvLog {
    FILE* f = nullptr;
    explicit CsvLog(bool on)
    {
        if (!on) return;
        char name[64];
        const time_t now = time(nullptr);
        strftime(name, sizeof(name), "%Y-%m-%d-%H-%M-%S_log.txt", localtime(&now));
        f = fopen(name, "w");
        if (f)
            fprintf(f, "%40s,%6s,%12s,%12s,%12s,%12s,%12s,%12s,%12s,%12s,%12s,%12s\n", "scene", "iter", "lighttime",
                    "buildtime", "cameratime", "res_x", "res_y", "mem_used", "mem_free", "mem_total", "light_util",
                    "camera_util");
    }
    ~CsvLog()
    {
        if (f) fclose(f);
    }
    void line(const upbp_config& c, int iter, const upbp_iter_stats& st)
    {
        if (!f) return;
        uint64_t fr = 0, tot = 0;
        upbp_memory_info(&fr, &tot);
        const double mib = 1.0 / (1024.0 * 1024.0);
        fprintf(f, "%40s,%6d,%12g,%12g,%12g,%12u,%12u,%12g,%12g,%12g,%12g,%12g\n", c.scene_obj_file, iter,
                st.ms_light * 1e-3, st.ms_build * 1e-3, st.ms_camera * 1e-3, c.res_x, c.res_y, double(tot - fr) * mib,
                double(fr) * mib, double(tot) * mib, 0.0, 0.0);
        fflush(f);
    }
};

} // namespace

int main(int argc, const char* argv[])
{
    upbp_config cfg;
    // the reference exits with 0 both on parse errors and after printing help
    if (upbp_parse_commandline(argc, argv, &cfg) != UPBP_OK) return 0;
    upbp_scene_handle* sh = nullptr;
    int e = cfg.scene_id > -1
                ? upbp_scene_load_predefined(cfg.scene_id, uint32_t(cfg.res_x), uint32_t(cfg.res_y), &sh)
                : upbp_scene_load_obj(cfg.scene_obj_file, uint32_t(cfg.res_x), uint32_t(cfg.res_y), &sh);
    if (e) {
        fprintf(stderr, "Error: %s\n", upbp_global_error());
        return 2;
    }
    printf("Scene:    %s\n", cfg.scene_obj_file);
    if (cfg.max_time > 0.f) printf("Target:   %g seconds render time\n", double(cfg.max_time));
    else printf("Target:   %d iteration(s)\n", cfg.iteration_count);
    printf("Running:  %s", describe(cfg).c_str());

    upbp_renderer* r = nullptr;
    e = upbp_create(upbp_scene_get(sh), &cfg, &r);
    if (e) {
        fprintf(stderr, "Error: %s\n", upbp_last_error(nullptr));
        upbp_scene_free(sh);
        return 1;
    }
    if (cfg.guide_surface_alpha > 0.f || cfg.guide_media_alpha > 0.f) {
        upbp_guide_config g;
        upbp_guide_defaults(&g);
        g.surface_alpha = cfg.guide_surface_alpha;
        g.alpha = cfg.guide_media_alpha;
        if (upbp_set_guide(r, &g)) {
            fprintf(stderr, "Error: %s\n", upbp_last_error(r));
            upbp_destroy(r);
            upbp_scene_free(sh);
            return 1;
        }
    }
    const std::string out = cfg.output_name;
    const std::string name = out.substr(0, out.size() >= 4 ? out.size() - 4 : 0);
    const std::string ext = out.size() >= 3 ? out.substr(out.size() - 3) : "";
    auto continuous = [&](int iter) { // gpuupbp.cpp:33-55
        if (cfg.continuous_output > 0 && iter % cfg.continuous_output == 0) {
            const std::string fn = name + "-" + std::to_string(iter) + "." + ext;
            if (save_mean(r, cfg, iter, fn)) fprintf(stderr, "Error: could not write %s\n", fn.c_str());
        }
    };

    CsvLog log(cfg.do_log != 0);
    upbp_iter_stats st;
    upbp_iter_stats* stp = cfg.do_log ? &st : nullptr;
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    int iter = 0;
    if (cfg.max_time > 0.f) {
        const auto end = t0 + std::chrono::duration<double>(cfg.max_time);
        while (clk::now() < end) {
            if ((e = upbp_run_iteration(r, iter, stp))) break;
            if (stp) log.line(cfg, iter, st);
            iter++;
            continuous(iter);
            if (cfg.continuous_output <= 0) upbp_sync(r); // keep the time check close to the device
        }
    } else {
        int percent = -1;
        for (iter = 0; iter < cfg.iteration_count; iter++) {
            if ((e = upbp_run_iteration(r, iter, stp))) break;
            if (stp) log.line(cfg, iter, st);
            const int p = int((float(iter + 1) / float(cfg.iteration_count)) * 100.f);
            if (p != percent) {
                percent = p;
                printf("%d%%\n", p);
                fflush(stdout);
            }
            continuous(iter + 1);
        }
    }
    if (!e) e = upbp_sync(r);
    const double secs = std::chrono::duration<double>(clk::now() - t0).count();
    if (e) fprintf(stderr, "Error: %s\n", upbp_last_error(r));
    printf("done in %.2f s (%i iterations)\n", secs, iter);
    int rc = 0;
    if (!e && iter > 0 && save_mean(r, cfg, iter, out)) {
        fprintf(stderr, "Error: could not write %s\n", out.c_str());
        rc = 2;
    }
    upbp_destroy(r);
    upbp_scene_free(sh);
    return e ? 1 : rc;
}
