// OBJ + MTL + AUX scene loader.
//
// Same file semantics as SceneLoader::load_from_obj and its helpers in
// gpuupbp/misc/scene_loader_obj.cpp:272-1104 and SceneLoader::load_scene in
// gpuupbp/misc/scene_loader.hpp:44-76:
//  * keywords are dispatched on their leading characters (so "scaterring" in the
//    reference's own scenes still sets the scattering coefficient);
//  * faces with more than three corners are fanned around the first corner;
//  * every triangle with an emissive material becomes one area light, after the
//    AUX lights, in file order (:692-702);
//  * a medium with continuation probability -1 gets max(max albedo, 0.8) (:1075-1088);
//  * without globalMediumId a clear global medium is appended (:1097-1103);
//  * the scene sphere is the half-diagonal of the triangle bounding box (scene_loader.hpp:65-71).
// Errors return UPBP_E_IO / UPBP_E_PARSE instead of exit(2).
#include "../../../include/upbp_capi.h"
#include "scene_handle.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <sys/stat.h>
#include <vector>

namespace {

struct V3 {
    float x, y, z;
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 normalize(V3 a)
{
    const float inv = 1.f / std::sqrt(dot(a, a));
    return {a.x * inv, a.y * inv, a.z * inv};
}

struct Mat {
    std::string name;
    bool emissive = false;
    V3 kd{0, 0, 0}, ks{0, 0, 0}, ke{0, 0, 0}, mirror{0, 0, 0};
    float ns = 0.f;
    int medium = -1;
    int enclosing = -1;
    bool real = true;
    float ior = -1.f;
    int priority = -1;
};

struct Med {
    std::string name;
    V3 absorption{0, 0, 0}, emission{0, 0, 0}, scattering{.5f, .5f, .5f};
    float cont = -1.f;
    float g = 0.f;
};

struct Tri {
    uint32_t v[3];
    uint32_t n[3]; // UINT32_MAX when absent
    int mat;
};

struct AuxLight {
    int type; // 0 directional, 1 point, 2 background
    V3 pos{0, 0, 0}, emission{0, 0, 0};
    std::string env_map;
    float env_scale = 1.f, env_rotate = 0.f;
};

struct Loader {
    std::string err;
    std::vector<float> verts, norms;
    std::vector<Mat> mats;
    std::vector<Med> media;
    std::vector<Tri> tris;
    std::vector<AuxLight> lights;
    float tm[4][3] = {{0}};
    float fov = 0.f, tdist = 0.f;
    int cam_mat = -1;
    int global_medium = -1;

    int find_mat(const std::string& name)
    {
        for (size_t i = 0; i < mats.size(); i++)
            if (mats[i].name == name) return static_cast<int>(i);
        err = "material not found: " + name;
        return -1;
    }
    int find_or_add_medium(const std::string& name)
    {
        for (size_t i = 0; i < media.size(); i++)
            if (media[i].name == name) return static_cast<int>(i);
        Med m;
        m.name = name;
        media.push_back(m);
        return static_cast<int>(media.size() - 1);
    }
};

std::string dir_of(const std::string& p)
{
    const size_t i = p.find_last_of('/');
    return i == std::string::npos ? std::string() : p.substr(0, i + 1);
}

// Whitespace tokenizer over a whole file, with "rest of line" support.
struct Tok {
    std::string s;
    size_t pos = 0;
    bool read_file(const std::string& path)
    {
        FILE* f = fopen(path.c_str(), "rb");
        if (!f) return false;
        struct stat st;
        if (fstat(fileno(f), &st) != 0 || !S_ISREG(st.st_mode)) { // a directory opens, but is no file
            fclose(f);
            return false;
        }
        fseek(f, 0, SEEK_END);
        long n = ftell(f);
        fseek(f, 0, SEEK_SET);
        s.resize(n > 0 ? static_cast<size_t>(n) : 0);
        size_t got = n > 0 ? fread(&s[0], 1, static_cast<size_t>(n), f) : 0;
        s.resize(got);
        fclose(f);
        return true;
    }
    static bool ws(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\f' || c == '\v'; }
    bool next(std::string& out)
    {
        while (pos < s.size() && ws(s[pos])) pos++;
        if (pos >= s.size()) return false;
        size_t b = pos;
        while (pos < s.size() && !ws(s[pos])) pos++;
        out.assign(s, b, pos - b);
        return true;
    }
    void skip_line()
    {
        while (pos < s.size() && s[pos] != '\n') pos++;
    }
    // first whitespace-delimited word of the rest of the current line, consumes the line
    std::string word_of_line()
    {
        size_t e = pos;
        while (e < s.size() && s[e] != '\n') e++;
        std::string line(s, pos, e - pos);
        pos = e;
        size_t b = 0;
        while (b < line.size() && ws(line[b])) b++;
        size_t q = b;
        while (q < line.size() && !ws(line[q])) q++;
        return line.substr(b, q - b);
    }
    bool peek_number()
    {
        size_t p = pos;
        while (p < s.size() && ws(s[p])) p++;
        if (p >= s.size()) return false;
        char c = s[p];
        return (c >= '0' && c <= '9') || c == '-' || c == '+' || c == '.';
    }
    float f()
    {
        std::string t;
        if (!next(t)) return 0.f;
        return strtof(t.c_str(), nullptr);
    }
    int i()
    {
        std::string t;
        if (!next(t)) return 0;
        return static_cast<int>(strtol(t.c_str(), nullptr, 10));
    }
    V3 v3()
    {
        V3 r;
        r.x = f();
        r.y = f();
        r.z = f();
        return r;
    }
};

bool read_mtl(Loader& L, const std::string& path)
{
    Tok t;
    if (!t.read_file(path)) {
        L.err = "could not open " + path;
        return false;
    }
    std::string w;
    Mat* m = nullptr;
    while (t.next(w)) {
        switch (w[0]) {
        case '#': t.skip_line(); break;
        case 'n': { // newmtl
            Mat nm;
            nm.name = t.word_of_line();
            L.mats.push_back(nm);
            m = &L.mats.back();
            break;
        }
        case 'N':
            if (w.size() > 1 && w[1] == 's' && m) m->ns = t.f();
            break;
        case 'K':
            if (!m) {
                t.skip_line();
                break;
            }
            if (w.size() > 1 && w[1] == 'd') m->kd = t.v3();
            else if (w.size() > 1 && w[1] == 's') m->ks = t.v3();
            else if (w.size() > 1 && w[1] == 'a') t.v3();
            else if (w.size() > 1 && w[1] == 'e') {
                m->ke = t.v3();
                m->emissive = m->ke.x > 0.f || m->ke.y > 0.f || m->ke.z > 0.f;
            } else t.skip_line();
            break;
        default: t.skip_line(); break;
        }
    }
    return true;
}

bool read_obj(Loader& L, const std::string& path)
{
    Tok t;
    if (!t.read_file(path)) {
        L.err = "could not open " + path;
        return false;
    }
    std::string w;
    int cur_mat = -1;
    while (t.next(w)) {
        switch (w[0]) {
        case '#': t.skip_line(); break;
        case 'v':
            if (w.size() == 1) {
                V3 p = t.v3();
                L.verts.insert(L.verts.end(), {p.x, p.y, p.z});
            } else if (w[1] == 'n') {
                V3 n = t.v3();
                L.norms.insert(L.norms.end(), {n.x, n.y, n.z});
            } else if (w[1] == 't') {
                t.f();
                t.f();
            }
            break;
        case 'm': {
            const std::string lib = t.word_of_line();
            if (!read_mtl(L, dir_of(path) + lib)) return false;
            break;
        }
        case 'u': {
            const std::string name = t.word_of_line();
            cur_mat = L.find_mat(name);
            if (cur_mat < 0) return false;
            break;
        }
        case 'f': {
            if (cur_mat < 0) {
                L.err = "face before usemtl";
                return false;
            }
            // corners until the next non-numeric token
            std::vector<uint32_t> cv, cn;
            while (t.peek_number()) {
                std::string c;
                t.next(c);
                long v = 0, vt = 0, vn = 0;
                uint32_t ni = UINT32_MAX;
                if (c.find("//") != std::string::npos) {
                    sscanf(c.c_str(), "%ld//%ld", &v, &vn);
                    ni = static_cast<uint32_t>(vn - 1);
                } else if (sscanf(c.c_str(), "%ld/%ld/%ld", &v, &vt, &vn) == 3) {
                    ni = static_cast<uint32_t>(vn - 1);
                } else {
                    sscanf(c.c_str(), "%ld", &v);
                }
                cv.push_back(static_cast<uint32_t>(v - 1));
                cn.push_back(ni);
            }
            if (cv.size() < 3) {
                L.err = "face with fewer than three corners";
                return false;
            }
            for (size_t k = 2; k < cv.size(); k++) {
                Tri tr;
                tr.v[0] = cv[0];
                tr.v[1] = cv[k - 1];
                tr.v[2] = cv[k];
                tr.n[0] = cn[0];
                tr.n[1] = cn[k - 1];
                tr.n[2] = cn[k];
                tr.mat = cur_mat;
                L.tris.push_back(tr);
            }
            break;
        }
        default: t.skip_line(); break;
        }
    }
    return true;
}

bool read_aux(Loader& L, const std::string& path)
{
    Tok t;
    if (!t.read_file(path)) {
        L.err = "could not open " + path;
        return false;
    }
    std::string w;
    int mat = -1, med = -1;
    auto need_mat = [&](const char* opt) {
        if (mat < 0) L.err = std::string("using ") + opt + " option without material selection";
        return mat >= 0;
    };
    auto need_med = [&](const char* opt) {
        if (med < 0) L.err = std::string("using ") + opt + " option without medium selection";
        return med >= 0;
    };
    while (t.next(w)) {
        const char c1 = w.size() > 1 ? w[1] : '\0';
        switch (w[0]) {
        case '#': t.skip_line(); break;
        case 'T': { // TM_ROWk
            const int row = w.size() > 6 ? w[6] - '0' : -1;
            V3 r = t.v3();
            if (row >= 0 && row < 4) {
                L.tm[row][0] = r.x;
                L.tm[row][1] = r.y;
                L.tm[row][2] = r.z;
            }
            break;
        }
        case 'C': { // CAMERA_*
            const char k = w.size() > 7 ? w[7] : '\0';
            if (k == 'F') L.fov = t.f();
            else if (k == 'T') L.tdist = t.f();
            else if (k == 'M') {
                L.cam_mat = L.find_mat(t.word_of_line());
                if (L.cam_mat < 0) return false;
            } else if (k == 'R') {
                t.i(); // the render resolution comes from the command line
            } else {
                L.err = "unknown camera option: " + w;
                return false;
            }
            break;
        }
        case 'm':
            if (c1 == 'a') { // material
                mat = L.find_mat(t.word_of_line());
                if (mat < 0) return false;
            } else if (c1 == 'e') {
                const char k = w.size() > 6 ? w[6] : '\0';
                if (k == 'I') { // mediumId
                    if (!need_mat("mediumId")) return false;
                    L.mats[mat].medium = L.find_or_add_medium(t.word_of_line());
                } else if (k == '\0') { // medium
                    med = L.find_or_add_medium(t.word_of_line());
                } else {
                    t.skip_line();
                }
            } else if (c1 == 'i') { // mirror
                if (!need_mat("mirror")) return false;
                L.mats[mat].mirror = t.v3();
            } else {
                t.skip_line();
            }
            break;
        case 'i':
            if (!need_mat("IOR")) return false;
            L.mats[mat].ior = t.f();
            break;
        case 'p':
            if (!need_mat("priority")) return false;
            L.mats[mat].priority = t.i();
            break;
        case 'g':
            if (c1 == '\0') {
                if (!need_med("g")) return false;
                L.media[med].g = t.f();
            } else if (c1 == 'e') {
                if (!need_mat("geometryType")) return false;
                const std::string gt = t.word_of_line();
                if (!gt.empty() && gt[0] == 'r') L.mats[mat].real = true;
                else if (!gt.empty() && gt[0] == 'i') L.mats[mat].real = false;
                else {
                    L.err = "unknown geometry type";
                    return false;
                }
            } else if (c1 == 'l') {
                L.global_medium = L.find_or_add_medium(t.word_of_line());
            } else {
                t.skip_line();
            }
            break;
        case 'K':
            if (!need_mat("Ke")) return false;
            L.mats[mat].ke = t.v3();
            L.mats[mat].emissive = L.mats[mat].ke.x > 0.f || L.mats[mat].ke.y > 0.f || L.mats[mat].ke.z > 0.f;
            break;
        case 'a':
            if (!need_med("absorption")) return false;
            L.media[med].absorption = t.v3();
            break;
        case 'c':
            if (!need_med("continuation_probability")) return false;
            L.media[med].cont = t.f();
            break;
        case 'e':
            if (c1 == 'm') {
                if (!need_med("emission")) return false;
                L.media[med].emission = t.v3();
            } else if (c1 == 'n') {
                if (!need_mat("enclosingMatId")) return false;
                L.mats[mat].enclosing = L.find_mat(t.word_of_line());
                if (L.mats[mat].enclosing < 0) return false;
            }
            break;
        case 's':
            if (!need_med("scattering")) return false;
            L.media[med].scattering = t.v3();
            break;
        case 'l': {
            const char k = w.size() > 6 ? w[6] : '\0';
            AuxLight al;
            if (k == 'p' || k == 'd') {
                al.type = k == 'p' ? 1 : 0;
                al.pos = t.v3();
                al.emission = t.v3();
                L.lights.push_back(al);
            } else if (k == 'b') {
                al.type = 2;
                const char kk = w.size() > 17 ? w[17] : '\0';
                if (kk == 'c') {
                    al.emission = t.v3();
                } else if (kk == 'e') {
                    al.env_scale = t.f();
                    al.env_rotate = t.f();
                    al.env_map = t.word_of_line();
                }
                L.lights.push_back(al);
            } else {
                t.skip_line();
            }
            break;
        }
        default: t.skip_line(); break;
        }
    }
    return true;
}

} // namespace

static thread_local std::string g_loader_error;

void upbp_set_loader_error(const std::string& msg) { g_loader_error = msg; }

extern "C" const char* upbp_global_error(void) { return g_loader_error.c_str(); }

static int load_obj(const char* obj_path, uint32_t res_x, uint32_t res_y, upbp_scene_handle** out)
{
    (void)res_x;
    (void)res_y;
    if (!obj_path || !out) {
        g_loader_error = "null argument";
        return UPBP_E_ARG;
    }
    *out = nullptr;
    Loader L;
    const std::string path(obj_path);
    FILE* probe = fopen(path.c_str(), "rb");
    if (!probe) {
        g_loader_error = "could not open " + path;
        return UPBP_E_IO;
    }
    fclose(probe);
    if (!read_obj(L, path) || !read_aux(L, path + ".aux")) {
        g_loader_error = L.err;
        return L.err.compare(0, 14, "could not open") == 0 ? UPBP_E_IO : UPBP_E_PARSE;
    }

    std::unique_ptr<upbp_scene_handle> own(new upbp_scene_handle());
    upbp_scene_handle* h = own.get();
    upbp_scene& s = h->scene;
    memset(&s, 0, sizeof(s));
    s.background_light_idx = -1;

    // AUX lights first (scene_loader_obj.cpp:1001-1033)
    for (const AuxLight& al : L.lights) {
        upbp_light l;
        memset(&l, 0, sizeof(l));
        l.mat_id = -1;
        l.med_id = -1;
        if (al.type == 1) {
            l.type = UPBP_LIGHT_POINT;
            memcpy(l.position, &al.pos, sizeof(l.position));
            memcpy(l.intensity, &al.emission, sizeof(l.intensity));
        } else if (al.type == 0) {
            l.type = UPBP_LIGHT_DIRECTIONAL;
            memcpy(l.position, &al.pos, sizeof(l.position));
            memcpy(l.intensity, &al.emission, sizeof(l.intensity));
        } else {
            l.type = UPBP_LIGHT_BACKGROUND;
            memcpy(l.intensity, &al.emission, sizeof(l.intensity));
            l.scale = 1.f;
            // only the first background light counts, and only it loads its map
            // (scene_loader_obj.cpp:994-1013)
            if (s.background_light_idx < 0) {
                s.background_light_idx = static_cast<int32_t>(h->lights.size());
                if (!al.env_map.empty()) {
                    const std::string file = al.env_map[0] == '/' ? al.env_map : dir_of(path) + al.env_map;
                    std::vector<float> img;
                    uint32_t w = 0, ht = 0;
                    const std::string err = upbp_read_exr_rgb(file, img, w, ht);
                    if (!err.empty()) {
                        g_loader_error = "environment map loading failed: " + err;
                        return UPBP_E_IO;
                    }
                    if (w < 2 * ht) { // the distribution reads 2 * height columns per row
                        g_loader_error = "environment map must be at least twice as wide as high";
                        return UPBP_E_PARSE;
                    }
                    // rotation about the vertical axis by whole columns, then the scale
                    // (tinyexr_wrapper.cpp:117-130)
                    h->env.assign(img.size(), 0.f);
                    const int rot = static_cast<int>(al.env_rotate * static_cast<float>(w));
                    for (uint32_t y = 0; y < ht; y++)
                        for (uint32_t x = 0; x < w; x++) {
                            int xs = static_cast<int>(x) + rot;
                            if (xs >= static_cast<int>(w)) xs -= static_cast<int>(w);
                            if (xs < 0 || xs >= static_cast<int>(w)) continue; // rotations outside [0, 2)
                            for (int k = 0; k < 3; k++)
                                h->env[3 * (size_t(y) * w + size_t(xs)) + k] = img[3 * (size_t(y) * w + x) + k] * al.env_scale;
                        }
                    h->env_w = w;
                    h->env_h = ht;
                    l.uses_env_map = 1;
                }
            }
        }
        h->lights.push_back(l);
    }

    V3 bmin{1e36f, 1e36f, 1e36f}, bmax{-1e36f, -1e36f, -1e36f};
    const size_t nv = L.verts.size() / 3, nn = L.norms.size() / 3;
    for (const Tri& t : L.tris) {
        const Mat& m = L.mats[static_cast<size_t>(t.mat)];
        V3 p[3];
        for (int k = 0; k < 3; k++) {
            if (t.v[k] >= nv) {
                g_loader_error = "face references a missing vertex";
                return UPBP_E_PARSE;
            }
            p[k] = {L.verts[3 * t.v[k]], L.verts[3 * t.v[k] + 1], L.verts[3 * t.v[k] + 2]};
            bmin = {std::fmin(bmin.x, p[k].x), std::fmin(bmin.y, p[k].y), std::fmin(bmin.z, p[k].z)};
            bmax = {std::fmax(bmax.x, p[k].x), std::fmax(bmax.y, p[k].y), std::fmax(bmax.z, p[k].z)};
        }
        V3 face = normalize(cross(sub(p[1], p[0]), sub(p[2], p[0])));
        V3 n[3];
        const bool has_n = t.n[0] != UINT32_MAX && t.n[1] != UINT32_MAX && t.n[2] != UINT32_MAX;
        if (has_n) {
            for (int k = 0; k < 3; k++) {
                if (t.n[k] >= nn) {
                    g_loader_error = "face references a missing normal";
                    return UPBP_E_PARSE;
                }
                n[k] = {L.norms[3 * t.n[k]], L.norms[3 * t.n[k] + 1], L.norms[3 * t.n[k] + 2]};
            }
            V3 avg = add(add(n[0], n[1]), n[2]);
            avg = {avg.x / 3.f, avg.y / 3.f, avg.z / 3.f};
            if (dot(face, avg) < 0.f) face = {-face.x, -face.y, -face.z};
        } else {
            n[0] = n[1] = n[2] = face;
        }
        int light_id = -1;
        if (m.emissive) {
            upbp_light l;
            memset(&l, 0, sizeof(l));
            l.type = UPBP_LIGHT_AREA;
            l.mat_id = m.enclosing < 0 ? -1 : m.enclosing;
            l.med_id = l.mat_id < 0 ? -1 : L.mats[static_cast<size_t>(l.mat_id)].medium;
            memcpy(l.p0, &p[0], 12);
            memcpy(l.p1, &p[1], 12);
            memcpy(l.p2, &p[2], 12);
            memcpy(l.intensity, &m.ke, 12);
            light_id = static_cast<int>(h->lights.size());
            h->lights.push_back(l);
        }
        std::vector<float>& V = m.real ? h->rv : h->iv;
        std::vector<float>& N = m.real ? h->rn : h->in;
        for (int k = 0; k < 3; k++) V.insert(V.end(), {p[k].x, p[k].y, p[k].z});
        for (int k = 0; k < 3; k++) N.insert(N.end(), {n[k].x, n[k].y, n[k].z});
        N.insert(N.end(), {face.x, face.y, face.z});
        (m.real ? h->rm : h->im).push_back(t.mat);
        (m.real ? h->rl : h->il).push_back(light_id);
    }

    for (const Mat& m : L.mats) {
        upbp_material o;
        memset(&o, 0, sizeof(o));
        memcpy(o.diffuse, &m.kd, 12);
        if (m.ns != 0.f) memcpy(o.phong, &m.ks, 12);
        o.phong_exponent = m.ns;
        memcpy(o.mirror, &m.mirror, 12);
        o.ior = m.ior;
        o.priority = m.priority;
        o.real = m.real ? 1 : 0;
        o.med_idx = m.medium;
        h->mats.push_back(o);
    }
    for (const Med& m : L.media) {
        upbp_medium o;
        memset(&o, 0, sizeof(o));
        memcpy(o.absorption, &m.absorption, 12);
        memcpy(o.emission, &m.emission, 12);
        memcpy(o.scattering, &m.scattering, 12);
        o.mean_cosine = m.g;
        float cont = m.cont;
        if (cont == -1.f) {
            const float ax = m.scattering.x / (m.scattering.x + m.absorption.x);
            const float ay = m.scattering.y / (m.scattering.y + m.absorption.y);
            const float az = m.scattering.z / (m.scattering.z + m.absorption.z);
            const float mx = std::fmax(ax, std::fmax(ay, az));
            if (std::isnan(mx) || std::isinf(mx)) cont = 1.f;
            else cont = mx > 0.8f ? mx : 0.8f;
        }
        o.continuation_prob = cont;
        h->media.push_back(o);
    }
    if (L.global_medium < 0) {
        upbp_medium clear;
        memset(&clear, 0, sizeof(clear));
        clear.continuation_prob = 1.f;
        h->media.push_back(clear);
        s.global_medium_idx = static_cast<uint32_t>(h->media.size() - 1);
    } else {
        s.global_medium_idx = static_cast<uint32_t>(L.global_medium);
    }

    // camera from the AUX transform rows (scene_loader_obj.cpp:224-253)
    upbp_camera& cam = s.camera;
    cam.origin[0] = L.tm[3][0];
    cam.origin[1] = L.tm[3][1];
    cam.origin[2] = L.tm[3][2];
    cam.target[0] = L.tm[3][0] - L.tm[2][0];
    cam.target[1] = L.tm[3][1] - L.tm[2][1];
    cam.target[2] = L.tm[3][2] - L.tm[2][2];
    V3 roll{L.tm[1][0], L.tm[1][1], L.tm[1][2]};
    const float rl = 1.f / std::sqrt(dot(roll, roll));
    cam.roll[0] = roll.x * rl;
    cam.roll[1] = roll.y * rl;
    cam.roll[2] = roll.z * rl;
    cam.horizontal_fov_deg = L.fov * 57.295779f;
    cam.focal_dist = L.tdist;
    cam.mat_id = L.cam_mat < 0 ? -1 : L.cam_mat;
    cam.med_id = cam.mat_id < 0 ? -1 : L.mats[static_cast<size_t>(cam.mat_id)].medium;

    V3 diag = sub(bmax, bmin);
    float r2 = dot(diag, diag);
    if (r2 == 0.f || r2 > 1e36f) r2 = 1.f;
    s.scene_center[0] = (bmax.x + bmin.x) * .5f;
    s.scene_center[1] = (bmax.y + bmin.y) * .5f;
    s.scene_center[2] = (bmax.z + bmin.z) * .5f;
    s.scene_radius = std::sqrt(r2) * .5f;

    h->bind();
    *out = own.release();
    return UPBP_OK;
}

// The C ABI lets no C++ exception through: a scene whose sizes exhaust memory is an error code.
extern "C" int upbp_scene_load_obj(const char* obj_path, uint32_t res_x, uint32_t res_y, upbp_scene_handle** out)
{
    try {
        return load_obj(obj_path, res_x, res_y, out);
    } catch (const std::exception& e) {
        g_loader_error = std::string("scene loading failed: ") + e.what();
        if (out) *out = nullptr;
        return UPBP_E_PARSE;
    }
}

extern "C" const upbp_scene* upbp_scene_get(const upbp_scene_handle* h) { return h ? &h->scene : nullptr; }

extern "C" void upbp_scene_free(upbp_scene_handle* h) { delete h; }
