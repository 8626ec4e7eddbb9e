// Sanitizer driver for the oracle (test infrastructure): small renders of generated and predefined
// scenes with several technique sets, the MIS log, the training samples and the CPU-baseline sample.
// Built and run by `make asan` (AddressSanitizer + UndefinedBehaviorSanitizer) and `make tsan`
// (ThreadSanitizer: the oracle renders on 4 threads). Exit status 0 and no sanitizer report pass.
#include "oracle.h"
#include <cstdio>
#include <cstdlib>
#include <string>
#include <sys/stat.h>
#include <vector>
extern "C" int oracle_training_samples(const upbp_scene* scene, const upbp_config* cfg, int iteration,
                                       int32_t train_media, int32_t train_surfaces, int n_threads, float* out,
                                       uint64_t max_samples, uint64_t* n_samples);

static int run(const upbp_scene* sc, const char* name, std::vector<const char*> args)
{
    upbp_config cfg;
    upbp_config_defaults(&cfg);
    std::vector<const char*> av = {"gpuupbp"};
    for (auto a : args) av.push_back(a);
    if (upbp_parse_commandline(int(av.size()), av.data(), &cfg) != 0) {
        printf("%s: parse failed\n", name);
        return 1;
    }
    std::vector<double> img(size_t(cfg.res_x) * cfg.res_y * 3);
    oracle_stats st{};
    int rc = oracle_render(sc, &cfg, 0, 2, 4, img.data(), nullptr, &st);
    double m = 0;
    for (double v : img) m += v;
    printf("%s %s: rc %d mean %.4f errors %llu\n", name, args.size() > 5 ? args[5] : "", rc, m / img.size(),
           (unsigned long long)st.path_errors);
    uint32_t px[3] = {0, 100, 500};
    std::vector<double> mis(6 * 200000);
    uint32_t nr = 0;
    rc |= oracle_mis_check(sc, &cfg, 0, px, 3, 64, 4, mis.data(), 200000, &nr);
    std::vector<float> ts(8 * 100000);
    uint64_t ns = 0;
    rc |= oracle_training_samples(sc, &cfg, 0, 1, 1, 4, ts.data(), 100000, &ns);
    double times[3];
    rc |= oracle_render_sample(sc, &cfg, 0, 256, 256, 2, times, &st);
    fflush(stdout);
    return rc;
}

int main()
{
    int bad = 0;
    const char* algs[] = {"upbp_all", "upbp_bpt", "upbp_surf+pp3d", "upbp_bb1d+bpt", "upbp_pb2d+bp2d"};
    for (const char* kind : {"box", "boxfog", "furnace"}) {
        const std::string d = std::string("build/scene_") + kind;
        mkdir(d.c_str(), 0755);
        const std::string obj = d + "/" + kind + ".obj";
        if (upbp_generate_scene(kind, 0, obj.c_str()) != 0) return 3;
        upbp_scene_handle* h = nullptr;
        if (upbp_scene_load_obj(obj.c_str(), 32, 32, &h) != 0) return 4;
        for (const char* a : algs)
            bad |= run(upbp_scene_get(h), kind, {"-s", "-1", obj.c_str(), "-r", "32x32", "-a", a, "-l", "8", "-r_initial", "0.03"});
        bad |= run(upbp_scene_get(h), kind, {"-s", "-1", obj.c_str(), "-r", "32x32", "-a", "upbp_all", "-l", "8", "-qbt", "short", "-previous"});
        bad |= run(upbp_scene_get(h), kind, {"-s", "-1", obj.c_str(), "-r", "32x32", "-a", "upbp_all", "-l", "8", "-compatible"});
        upbp_scene_free(h);
    }
    for (int id : {0, 5, 12, 20, 33, 40}) {
        upbp_scene_handle* h = nullptr;
        if (upbp_scene_load_predefined(id, 32, 32, &h) != 0) continue;
        char ids[8];
        snprintf(ids, sizeof ids, "%d", id);
        bad |= run(upbp_scene_get(h), upbp_predefined_scene_name(id, 0), {"-s", ids, "-r", "32x32", "-a", "upbp_all", "-l", "10", "-r_initial", "0.03"});
        upbp_scene_free(h);
    }
    printf("oracle run done, bad %d\n", bad);
    return bad;
}
