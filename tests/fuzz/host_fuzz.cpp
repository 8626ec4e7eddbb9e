// Fuzz driver for the library's host code, built with AddressSanitizer and UndefinedBehaviorSanitizer
// by tests/test_host_fuzz_cpu.py: generated scenes with corrupted OBJ, MTL and AUX files, an EXR file
// with corrupted or truncated bytes, and random command lines. Every call must return (an error
// code is fine); a sanitizer report or an exception through the C ABI ends the run.
// Usage: host_fuzz ROUNDS WORK_DIR [KIND,KIND,...]   (scene kinds of upbp_generate_scene)
#include "upbp_capi.h"
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <vector>

static std::string slurp(const std::string& p)
{
    std::ifstream f(p, std::ios::binary);
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}
static void spit(const std::string& p, const std::string& s)
{
    std::ofstream f(p, std::ios::binary);
    f << s;
}
static std::vector<std::string> lines(const std::string& s)
{
    std::vector<std::string> v;
    std::stringstream ss(s);
    std::string l;
    while (std::getline(ss, l)) v.push_back(l);
    return v;
}
static const char* kTokens[] = {"-1", "0", "1", "2147483647", "-2147483648", "99999999999", "nan", "inf", "-inf",
                                "1e39", "-0", "", "/", "//", "1//", "a", "3/4/5/6", "-5/-5/-5", "1e-45", "#"};

static std::string mutate_text(const std::string& s, std::mt19937& g)
{
    auto v = lines(s);
    const int n = 1 + int(g() % 4);
    for (int k = 0; k < n && !v.empty(); k++) {
        const size_t i = g() % v.size();
        switch (g() % 6) {
        case 0: v.erase(v.begin() + long(i)); break;
        case 1: v[i] = v[i].substr(0, g() % (v[i].size() + 1)); break;
        case 2: { // replace a token
            std::stringstream ss(v[i]);
            std::vector<std::string> t;
            std::string w;
            while (ss >> w) t.push_back(w);
            if (!t.empty()) t[g() % t.size()] = kTokens[g() % (sizeof(kTokens) / sizeof(*kTokens))];
            std::string o;
            for (auto& x : t) o += x + " ";
            v[i] = o;
            break;
        }
        case 3: v.insert(v.begin() + long(i), v[g() % v.size()]); break;
        case 4: v[i] += std::string(" ") + kTokens[g() % (sizeof(kTokens) / sizeof(*kTokens))]; break;
        case 5: if (!v[i].empty()) v[i][g() % v[i].size()] = char(g() % 256); break;
        }
    }
    std::string o;
    for (auto& l : v) o += l + "\n";
    if (g() % 8 == 0) o = o.substr(0, g() % (o.size() + 1));
    return o;
}

int main(int argc, char** argv)
{
    const int rounds = argc > 1 ? atoi(argv[1]) : 200;
    const std::string dir = argc > 2 ? argv[2] : "/tmp/upbp_fuzz";
    std::vector<std::string> kinds;
    {
        std::stringstream ks(argc > 3 ? argv[3] : "box,boxfog,furnace");
        std::string k;
        while (std::getline(ks, k, ',')) kinds.push_back(k);
    }
    std::mt19937 g(12345);
    int loads_ok = 0, loads_err = 0;
    // scenes
    for (const std::string& kind : kinds) {
        const std::string d = dir + "/" + kind;
        if (mkdir(d.c_str(), 0755) != 0 && errno != EEXIST) return 2;
        const std::string obj = d + "/" + kind + ".obj";
        if (upbp_generate_scene(kind.c_str(), 0, obj.c_str()) != 0) {
            printf("generate %s failed: %s\n", kind.c_str(), upbp_global_error());
            return 4;
        }
        const std::string mtl = d + "/" + kind + ".mtl", aux = d + "/" + kind + ".aux";
        const std::string o0 = slurp(obj), m0 = slurp(mtl), a0 = slurp(aux);
        const bool big = o0.size() > (8u << 20);
        const int nr = big ? rounds / 20 : rounds;
        for (int r = 0; r < nr; r++) {
            const int which = int(g() % 3);
            spit(obj, which == 0 ? mutate_text(o0, g) : o0);
            spit(mtl, which == 1 ? mutate_text(m0, g) : m0);
            spit(aux, which == 2 ? mutate_text(a0, g) : a0);
            upbp_scene_handle* h = nullptr;
            const int rc = upbp_scene_load_obj(obj.c_str(), 16, 16, &h);
            if (rc == 0) {
                loads_ok++;
                upbp_scene_free(h);
            } else
                loads_err++;
        }
        spit(obj, o0);
        spit(mtl, m0);
        spit(aux, a0);
        printf("scene %s: %d rounds\n", kind.c_str(), nr);
        fflush(stdout);
    }
    printf("\nscene loads: %d ok, %d rejected\n", loads_ok, loads_err);
    // a scene lit by an environment map whose EXR file is corrupted
    {
        const std::string d = dir + "/env";
        if (mkdir(d.c_str(), 0755) != 0 && errno != EEXIST) return 2;
        const std::string obj = d + "/env.obj";
        if (upbp_generate_scene("box", 0, obj.c_str()) != 0) return 4;
        spit(obj + ".aux", slurp(obj + ".aux") + "\nlight_background_em 2 0.25 env.exr\n");
        const uint32_t W = 32, H = 16;
        std::vector<float> sky(3 * W * H);
        for (size_t i = 0; i < sky.size(); i++) sky[i] = 0.5f + float(i % 13) * 0.1f;
        if (upbp_save_exr((d + "/env.exr").c_str(), sky.data(), W, H) != 0) return 3;
        const std::string e0 = slurp(d + "/env.exr");
        int ok = 0, bad = 0;
        for (int r = 0; r < rounds; r++) {
            std::string e = e0;
            const int n = 1 + int(g() % 4);
            for (int k = 0; k < n && !e.empty(); k++) {
                const size_t i = g() % e.size();
                if (g() % 5 == 0) e = e.substr(0, i);
                else e[i] = char(g() % 256);
            }
            spit(d + "/env.exr", r == 0 ? e0 : e);
            upbp_scene_handle* h = nullptr;
            if (upbp_scene_load_obj(obj.c_str(), 16, 16, &h) == 0) {
                ok++;
                upbp_scene_free(h);
            } else
                bad++;
        }
        printf("\nenv scene loads: %d ok, %d rejected\n", ok, bad);
    }
    // predefined scenes
    for (int id = -1; id < 48; id++) {
        upbp_scene_handle* h = nullptr;
        if (upbp_scene_load_predefined(id, 8, 8, &h) == 0) upbp_scene_free(h);
        (void)upbp_predefined_scene_name(id, id & 1);
    }
    // EXR
    {
        const uint32_t W = 37, H = 11;
        std::vector<float> img(3 * W * H);
        for (size_t i = 0; i < img.size(); i++) img[i] = float(i % 97) * 0.25f;
        const std::string p = dir + "/img.exr";
        if (upbp_save_exr(p.c_str(), img.data(), W, H) != 0) return 3;
        const std::string e0 = slurp(p);
        std::vector<float> out(3 * 4096);
        int ok = 0, bad = 0;
        for (int r = 0; r < rounds * 20; r++) {
            std::string e = e0;
            const int n = 1 + int(g() % 6);
            for (int k = 0; k < n; k++) {
                const size_t i = g() % e.size();
                switch (g() % 4) {
                case 0: e[i] = char(g() % 256); break;
                case 1: e[i] = char(0xff); break;
                case 2: e[i] = 0; break;
                case 3: e = e.substr(0, i); break;
                }
                if (e.empty()) break;
            }
            spit(p, e);
            uint32_t w = 0, h = 0;
            if (upbp_read_exr(p.c_str(), out.data(), out.size(), &w, &h) == 0) ok++;
            else bad++;
        }
        printf("\nexr: %d ok, %d rejected\n", ok, bad);
    }
    // command lines
    {
        const char* opts[] = {"-s", "-r", "-l", "-a", "-r_initial", "-r_alpha", "-i", "-t", "-o", "-ioff",
                              "-previous", "-compatible", "-speconly", "-ignorespec", "-log", "-qbt", "-pbc", "-pbt",
                              "-pcpi", "-guide", "-guide_media", "-assert", "-r_initial_pb2d", "-r_initial_bb1d",
                              "-r_alpha_pp3d", "-r_alpha_surf", "-r_alpha_", "-r_initial_", "-sn", "-continuous_output",
                              "-h", "-hf"};
        const char* vals[] = {"0",  "-1", "1",  "64x64", "0x0",  "99999999x99999999", "x",   "upbp_all", "upbp_bpt+surf",
                              "nan", "inf", "1e40", "-5",  "abc",  "",  "2147483648", "long", "short", "a.exr",
                              "0.5", "4294967296", "-2147483649", "1x", "x1", "8x0", "65536x65536", "upbp_",
                              "upbp_all+", "+", "bpt"};
        int ok = 0, bad = 0;
        for (int r = 0; r < rounds * 50; r++) {
            std::vector<const char*> av = {"gpuupbp"};
            const int n = int(g() % 8);
            for (int k = 0; k < n; k++) {
                av.push_back(opts[g() % (sizeof(opts) / sizeof(*opts))]);
                if (g() % 4) av.push_back(vals[g() % (sizeof(vals) / sizeof(*vals))]);
            }
            upbp_config cfg;
            upbp_config_defaults(&cfg);
            if (upbp_parse_commandline(int(av.size()), av.data(), &cfg) == 0) ok++;
            else bad++;
        }
        printf("\nargv: %d ok, %d rejected\n", ok, bad);
    }
    printf("\nfuzz done\n");
    return 0;
}
