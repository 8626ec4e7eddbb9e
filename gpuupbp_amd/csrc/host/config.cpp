// Command-line configuration of the UPBP renderer.
//
// Same option set, defaults, error messages and output-name rules as
// parse_commandline in gpuupbp/misc/config.cpp:735-1370 (defaults :738-785,
// relative -pcpi :1340-1346, default output name :520-605 and :1348-1367).
// Written table-driven: the ten radius options share one handler.
#include "../../../include/upbp_capi.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

namespace {

void copy_str(char* dst, size_t cap, const std::string& s)
{
    size_t n = s.size() < cap - 1 ? s.size() : cap - 1;
    memcpy(dst, s.data(), n);
    dst[n] = '\0';
}

std::string two_digits(long v)
{
    char buf[32];
    snprintf(buf, sizeof(buf), "%02ld", v);
    return buf;
}

// Output name without scene prefix (config.cpp:520-573).
std::string name_tail(const upbp_config& c)
{
    std::string s;
    if (c.iteration_count > 0) s += "_i" + two_digits(c.iteration_count);
    s += "_l" + two_digits(static_cast<long>(c.max_path_length));
    s += "_a";
    static const struct {
        uint32_t bit;
        const char* tag;
    } techs[] = {{UPBP_BPT, "+bpt"},
                 {UPBP_SURF, "+surf"},
                 {UPBP_PP3D, "+pp3d"},
                 {UPBP_PB2D, "+pb2d"},
                 {UPBP_BB1D, "+bb1d"},
                 {UPBP_BP2D, "+bp2d"}};
    for (const auto& t : techs)
        if (c.algorithm_flags & t.bit) s += t.tag;
    if (c.algorithm_flags & (UPBP_PB2D | UPBP_BB1D)) s += c.query_beam_type == UPBP_BEAM_SHORT ? "_qbtS" : "_qbtL";
    if (c.algorithm_flags & (UPBP_BP2D | UPBP_BB1D)) s += c.photon_beam_type == UPBP_BEAM_SHORT ? "_pbtS" : "_pbtL";
    s += c.additional_args;
    s += ".exr";
    return s;
}

std::vector<std::string> split_plus(const std::string& s)
{
    std::vector<std::string> out;
    std::string cur;
    for (char ch : s) {
        if (ch == '+') {
            out.push_back(cur);
            cur.clear();
        } else {
            cur += ch;
        }
    }
    if (!cur.empty()) out.push_back(cur);
    return out;
}

int fail(const char* msg)
{
    fprintf(stderr, "%s", msg);
    return UPBP_E_ARG;
}

void short_help(const char* prog)
{
    printf("\nUsage: %s [ -s <scene_id> | -a <algorithm> | -l <path length> |\n", prog);
    printf("           -t <time> | -i <iteration> | -o <output_name> ]\n\n");
    printf("    -s  Selects the scene (default 0):\n");
    printf("        -1    user must supply additional argument which contains path to an obj scene file\n");
    for (int i = 0; i < 5; i++) printf("          %-2d    %s\n", i, upbp_predefined_scene_name(i, 1));
    printf("          5..%d other predefined simple scenes, for complete list, please see full help (-hf)\n",
           upbp_predefined_scene_count() - 1);
    printf("    -a  Selects the rendering algorithm (default upbp_all):\n");
    printf("        upbp_<tech>[+<tech>]*  custom selection of techniques from upbp "
           "(tech=bpt|surf|pp3d|pb2d|bb1d|bp2d)\n");
    printf("        upbp_all    upbp_bpt+surf+pp3d+pb2d+bb1d\n\n");
    printf("    -l  Maximum length of traced paths (default 10).\n");
    printf("    -t  Number of seconds to run the algorithm.\n");
    printf("    -i  Number of iterations to run the algorithm (default 1).\n");
    printf("    -o  User specified output name, with extension .bmp or .exr (default .exr).\n");
    printf("\n    Note: Time (-t) takes precedence over iterations (-i) if both are defined.\n");
    printf("\n    For more options, please see full help (-hf)\n");
}

void full_help(const char* prog)
{
    short_help(prog);
    printf("\n    Predefined scenes (-s):\n");
    for (int i = 0; i < upbp_predefined_scene_count(); i++)
        printf("        %-2d    %s\n", i, upbp_predefined_scene_name(i, 1));
    printf("\n    Radius options:\n");
    printf("    -r_alpha <alpha>  -r_alpha_{surf,pp3d,pb2d,bb1d,bp2d} <alpha>\n");
    printf("    -r_initial <init> -r_initial_{surf,pp3d,pb2d,bb1d,bp2d} <init>"
           " (positive absolute, negative relative to scene size)\n");
    printf("\n    Light transport options:\n");
    printf("    -previous  -compatible  -speconly  -ignorespec <0|1>\n");
    printf("\n    Beams options:\n");
    printf("    -qbt <S|L>  -pbt <S|L>  -pbc <count>\n");
    printf("\n    Other options:\n");
    printf("    -continuous_output <iter_count>  -pcpi <path_count>  -sn <0|1>  -assert  -ioff <iter>  -log\n");
    printf("    -r <WIDTHxHEIGHT>  -ioff <iter>\n");
    printf("    -guide <alpha>  -guide_media <alpha>   learned scattering proposal at surfaces / in media\n");
    printf("                    (mixture probability in [0, 1]; 0 = off, the default; not in gpuupbp)\n");
}

} // namespace

extern "C" void upbp_config_defaults(upbp_config* c)
{
    memset(c, 0, sizeof(*c));
    c->algorithm_flags = UPBP_BPT | UPBP_SURF | UPBP_PP3D | UPBP_PB2D | UPBP_BB1D;
    c->iteration_count = 1;
    c->max_time = -1.f;
    c->res_x = 256;
    c->res_y = 256;
    c->max_path_length = 10;
    c->min_path_length = 0;
    c->path_count_per_iter = -1.f;
    c->query_beam_type = UPBP_BEAM_LONG;
    c->photon_beam_type = UPBP_BEAM_SHORT;
    c->surf_radius_initial = -0.0015f;
    c->surf_radius_alpha = 0.75f;
    c->pp3d_radius_initial = -0.001f;
    c->pp3d_radius_alpha = 1.0f;
    c->pb2d_radius_initial = -0.001f;
    c->pb2d_radius_alpha = 1.0f;
    c->bb1d_radius_initial = -0.001f;
    c->bb1d_radius_alpha = 1.0f;
    c->bb1d_used_light_subpath_count = -1.f;
    c->bp2d_radius_initial = -0.001f;
    c->bp2d_radius_alpha = 1.0f;
    c->continuous_output = 0;
    c->ignore_fully_spec_paths = 0;
    c->gpu_assert = 0;
    c->iteration_offset = 0;
    c->do_log = 0;
    c->use_shading_normal = 1;
    c->scene_id = 0;
}

extern "C" int upbp_parse_commandline(int argc, const char** argv, upbp_config* cfg)
{
    upbp_config& c = *cfg;
    upbp_config_defaults(&c);
    std::ostringstream extra;

    // radius options: {option suffix, initial, alpha, name tag}
    struct RadiusOpt {
        const char* tech;
        float* init;
        float* alpha;
        bool init_set;
        bool alpha_set;
    } radius[5] = {{"surf", &c.surf_radius_initial, &c.surf_radius_alpha, false, false},
                   {"pp3d", &c.pp3d_radius_initial, &c.pp3d_radius_alpha, false, false},
                   {"pb2d", &c.pb2d_radius_initial, &c.pb2d_radius_alpha, false, false},
                   {"bb1d", &c.bb1d_radius_initial, &c.bb1d_radius_alpha, false, false},
                   {"bp2d", &c.bp2d_radius_initial, &c.bp2d_radius_alpha, false, false}};

    for (int i = 1; i < argc; i++) {
        const std::string arg(argv[i]);
        if (arg == "-h" || arg == "--help" || arg == "/?") {
            short_help(argv[0]);
            fflush(stdout);
            return UPBP_HELP;
        }
        if (arg == "-hf" || arg == "--help_full" || arg == "/??") {
            full_help(argv[0]);
            fflush(stdout);
            return UPBP_HELP;
        }
        if (arg.empty() || arg[0] != '-') continue;

        // every remaining option except the flags takes one value
        auto need_value = [&](const char* opt, const char* help) -> const char* {
            if (++i == argc) {
                fprintf(stderr, "missing argument of %s option, please see help (%s)", opt, help);
                return nullptr;
            }
            return argv[i];
        };

        if (arg == "-s") {
            const char* v = need_value("-s", "-h");
            if (!v) return UPBP_E_ARG;
            std::istringstream iss(v);
            iss >> c.scene_id;
            if (iss.fail() || c.scene_id >= upbp_predefined_scene_count()) return fail("invalid argument of -s option, please see help (-h)");
            if (c.scene_id == 3 || c.scene_id == 4 || c.scene_id == 9 || c.scene_id == 10)
                return fail("scenes 3, 4, 9, and 10 cannot be rendered with UPBP, please see help (-h)");
            if (c.scene_id == -1) {
                if (++i == argc) return fail("missing file name argument of -s option, please see help (-h)");
                copy_str(c.scene_obj_file, sizeof(c.scene_obj_file), argv[i]);
            }
        } else if (arg == "-a") {
            const char* v = need_value("-a", "-h");
            if (!v) return UPBP_E_ARG;
            const std::string alg(v);
            if (alg == "upbp_all") {
                // BP2D is not part of upbp_all (config.cpp:858-860)
                c.algorithm_flags |= UPBP_BPT | UPBP_SURF | UPBP_PP3D | UPBP_PB2D | UPBP_BB1D;
            } else if (alg.size() >= 8 && alg.compare(0, 4, "upbp") == 0 && (alg[4] == '+' || alg[4] == '_')) {
                c.algorithm_flags = alg[4] == '+' ? (UPBP_BPT | UPBP_SURF) : 0u;
                for (const std::string& t : split_plus(alg.substr(5))) {
                    if (t == "bpt") c.algorithm_flags |= UPBP_BPT;
                    else if (t == "surf") c.algorithm_flags |= UPBP_SURF;
                    else if (t == "pp3d") c.algorithm_flags |= UPBP_PP3D;
                    else if (t == "pb2d") c.algorithm_flags |= UPBP_PB2D;
                    else if (t == "bb1d") c.algorithm_flags |= UPBP_BB1D;
                    else if (t == "bp2d") c.algorithm_flags |= UPBP_BP2D;
                    else return fail("invalid argument of -a option, please see help (-h)");
                }
            }
        } else if (arg == "-l") {
            const char* v = need_value("-l", "-h");
            if (!v) return UPBP_E_ARG;
            std::istringstream iss(v);
            long l = 0;
            iss >> l;
            if (iss.fail() || l < 1) return fail("invalid argument of -l option, please see help (-h)");
            c.max_path_length = static_cast<uint32_t>(l);
        } else if (arg == "-t") {
            const char* v = need_value("-t", "-h");
            if (!v) return UPBP_E_ARG;
            std::istringstream iss(v);
            iss >> c.max_time;
            if (iss.fail() || c.max_time < 0) return fail("invalid argument of -t option, please see help (-h)");
            extra << "_t" << v;
        } else if (arg == "-i") {
            const char* v = need_value("-i", "-h");
            if (!v) return UPBP_E_ARG;
            std::istringstream iss(v);
            iss >> c.iteration_count;
            if (iss.fail() || c.iteration_count < 1)
                return fail("invalid argument of -i option, please see help (-h)");
        } else if (arg == "-o") {
            const char* v = need_value("-o", "-h");
            if (!v) return UPBP_E_ARG;
            if (v[0] == '\0') return fail("invalid argument of -o option, please see help (-h)");
            copy_str(c.output_name, sizeof(c.output_name), v);
        } else if (arg == "-r") {
            const char* v = need_value("-r", "-hf");
            if (!v) return UPBP_E_ARG;
            int w = -1, h = -1;
            sscanf(v, "%dx%d", &w, &h);
            if (w <= 0 || h <= 0) return fail("invalid argument of -r option, please see help (-hf)");
            c.res_x = static_cast<uint32_t>(w);
            c.res_y = static_cast<uint32_t>(h);
            extra << "_r" << v;
        } else if (arg == "-r_alpha" || arg == "-r_initial") {
            const bool is_alpha = arg == "-r_alpha";
            const char* v = need_value(arg.c_str(), "-hf");
            if (!v) return UPBP_E_ARG;
            float val = 0.f;
            sscanf(v, "%f", &val);
            if (is_alpha ? val <= 0.f : val == 0.f) {
                fprintf(stderr, "invalid argument of %s option, please see help (-hf)", arg.c_str());
                return UPBP_E_ARG;
            }
            for (auto& r : radius) {
                if (is_alpha && !r.alpha_set) *r.alpha = val;
                if (!is_alpha && !r.init_set) *r.init = val;
            }
            extra << (is_alpha ? "_ralpha" : "_rinit") << v;
        } else if (arg.compare(0, 9, "-r_alpha_") == 0 || arg.compare(0, 11, "-r_initial_") == 0) {
            const bool is_alpha = arg[3] == 'a';
            const std::string tech = arg.substr(is_alpha ? 9 : 11);
            RadiusOpt* r = nullptr;
            for (auto& ro : radius)
                if (tech == ro.tech) r = &ro;
            if (!r) {
                fprintf(stderr, "invalid option \"%s\", please see help (-hf)", arg.c_str());
                return UPBP_E_ARG;
            }
            const char* v = need_value(arg.c_str(), "-hf");
            if (!v) return UPBP_E_ARG;
            float* dst = is_alpha ? r->alpha : r->init;
            sscanf(v, "%f", dst);
            if (is_alpha ? *dst <= 0.f : *dst == 0.f) {
                fprintf(stderr, "invalid argument of %s option, please see help (-hf)", arg.c_str());
                return UPBP_E_ARG;
            }
            (is_alpha ? r->alpha_set : r->init_set) = true;
            extra << (is_alpha ? "_ralpha" : "_rinit") << tech << v;
        } else if (arg == "-previous") {
            c.algorithm_flags |= UPBP_PREVIOUS;
            extra << "_prev";
        } else if (arg == "-compatible") {
            c.algorithm_flags |= UPBP_COMPATIBLE;
            extra << "_comp";
        } else if (arg == "-speconly") {
            c.algorithm_flags |= UPBP_SPECULAR_ONLY;
            extra << "_speconly";
        } else if (arg == "-ignorespec") {
            const char* v = need_value("-ignorespec", "-hf");
            if (!v) return UPBP_E_ARG;
            const std::string o(v);
            if (o == "1") c.ignore_fully_spec_paths = 1;
            else if (o == "0") c.ignore_fully_spec_paths = 0;
            else return fail("invalid argument of -ignorespec option, please see help (-hf)");
        } else if (arg == "-qbt" || arg == "-pbt") {
            const char* v = need_value(arg.c_str(), "-hf");
            if (!v) return UPBP_E_ARG;
            const uint32_t t = (v[0] == 'L' || v[0] == 'l') ? UPBP_BEAM_LONG : UPBP_BEAM_SHORT;
            (arg == "-qbt" ? c.query_beam_type : c.photon_beam_type) = t;
        } else if (arg == "-pbc") {
            const char* v = need_value("-pbc", "-hf");
            if (!v) return UPBP_E_ARG;
            sscanf(v, "%f", &c.bb1d_used_light_subpath_count);
            if (c.bb1d_used_light_subpath_count == 0.f)
                return fail("invalid argument of -pbc option, please see help (-hf)");
            extra << "_pbc" << v;
        } else if (arg == "-continuous_output") {
            const char* v = need_value("-continuous_output", "-hf");
            if (!v) return UPBP_E_ARG;
            std::istringstream iss(v);
            iss >> c.continuous_output;
            if (iss.fail()) return fail("invalid argument of -continuous_output option, please see help (-hf)");
        } else if (arg == "-pcpi") {
            const char* v = need_value("-pcpi", "-hf");
            if (!v) return UPBP_E_ARG;
            sscanf(v, "%f", &c.path_count_per_iter);
            if (std::floor(c.path_count_per_iter) == 0.f)
                return fail("invalid argument of -pcpi option, please see help (-hf)");
            extra << "_pcpi" << v;
        } else if (arg == "-sn") {
            const char* v = need_value("-sn", "-hf");
            if (!v) return UPBP_E_ARG;
            int a = 1;
            sscanf(v, "%d", &a);
            c.use_shading_normal = a == 0 ? 0 : 1;
            extra << "_sn" << v;
        } else if (arg == "-assert") {
            c.gpu_assert = 1;
            extra << "_assert";
        } else if (arg == "-ioff") {
            const char* v = need_value("-ioff", "-hf");
            if (!v) return UPBP_E_ARG;
            std::istringstream iss(v);
            long long off = 0;
            iss >> off;
            if (iss.fail() || off < 0) return fail("invalid argument of -ioff option, please see help (-hf)");
            c.iteration_offset = static_cast<uint64_t>(off);
        } else if (arg == "-log") {
            c.do_log = 1;
            extra << "_log";
        } else if (arg == "-guide" || arg == "-guide_media") {
            const char* v = need_value(arg.c_str(), "-hf");
            if (!v) return UPBP_E_ARG;
            char* end = nullptr;
            const float a = std::strtof(v, &end);
            if (end == v || *end != '\0' || !(a >= 0.f && a <= 1.f))
                return fail(("invalid argument of " + arg + " option, please see help (-hf)").c_str());
            (arg == "-guide" ? c.guide_surface_alpha : c.guide_media_alpha) = a;
            extra << (arg == "-guide" ? "_guide" : "_guidem") << v;
        } else {
            fprintf(stderr, "invalid option \"%s\", please see help (-hf)", arg.c_str());
            return UPBP_E_ARG;
        }
    }

    copy_str(c.additional_args, sizeof(c.additional_args), extra.str());

    // relative light path count (config.cpp:1340-1346)
    if (c.path_count_per_iter < 0)
        c.path_count_per_iter = std::floor(-c.path_count_per_iter * static_cast<float>(c.res_x) * c.res_y);
    else
        c.path_count_per_iter = std::floor(c.path_count_per_iter);

    if (c.output_name[0] == '\0') {
        std::string name;
        if (c.scene_id > -1) {
            name = "s" + two_digits(c.scene_id);
        } else {
            const std::string p(c.scene_obj_file);
            const size_t slash = p.find_last_of("\\/");
            const size_t dot = p.find_last_of('.');
            const size_t begin = slash == std::string::npos ? 0 : slash + 1;
            name = "s-" + p.substr(begin, dot == std::string::npos ? std::string::npos : dot - begin);
        }
        name += name_tail(c);
        copy_str(c.output_name, sizeof(c.output_name), name);
    }
    const std::string out(c.output_name);
    const std::string ext = out.size() > 4 ? out.substr(out.size() - 4) : "";
    if (ext != ".bmp" && ext != ".exr") copy_str(c.output_name, sizeof(c.output_name), out + ".exr");
    return UPBP_OK;
}
