// Seeded procedural scene generators. Each writes an OBJ + MTL + AUX triple
// (the reference's scene format, gpuupbp/misc/scene_loader_obj.cpp) so that the
// drop-in loader path is what reads every benchmark and test scene.
//
//   "box"   : open Cornell-style box [-1,1]^3, per-wall diffuse albedo U[0.5,0.8]^3,
//             0.5x0.5 quad light (Ke 10) under the ceiling, imaginary icosphere
//             (subdivision 3, r = 0.3) holding a medium (sigma_a 0.1, sigma_s 0.5, g 0).
//   "blobs" : 8x8 grid of noise-displaced icospheres (subdivision 5, 20480 triangles
//             each, 1.31M in total) on a floor, three overlapping imaginary media
//             spheres with priorities 1..3, two quad lights and a point light.
//   "stress": the same construction on a 10x10 grid (2.05M triangles) with eight
//             overlapping media spheres (priorities 1..8, g from -0.3 to 0.95).
//   "furnace": closed box whose floor is an emitter (Ke 1, black) and whose other five walls are
//              white diffuse (albedo 1), camera inside (tests only). The enclosure is isothermal:
//              every wall and the emitter have radiosity pi, so the radiance everywhere is exactly
//              1 for any enclosure shape (energy balance); paths of up to l segments miss only the
//              paths that bounce l - 1 times off white walls without reaching the floor.
//   "blobs_small": the blob scene with a 1x1 to 3x3 grid and 1 to 4 media (by the seed).
//   "slabroom": room of ~196k triangles with a dense forward-scattering medium slab, lit
//             only through a 5% aperture in one wall ("slabroom_small": coarse walls).
#include "../../../include/upbp_capi.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

namespace {

struct P3 {
    double x, y, z;
};
P3 operator+(P3 a, P3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
P3 operator-(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
P3 operator*(P3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double dot(P3 a, P3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
P3 cross(P3 a, P3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
P3 unit(P3 a) { return a * (1.0 / std::sqrt(dot(a, a))); }

// splitmix64: deterministic, platform independent
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x1234567ull) {}
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform(double a, double b) { return a + (b - a) * (static_cast<double>(next() >> 11) * 0x1.0p-53); }
};

struct Mesh {
    std::vector<P3> v;
    std::vector<P3> n; // optional per-vertex normals, same indexing as v
    std::vector<uint32_t> f;
};

Mesh icosphere(int subdiv)
{
    Mesh m;
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    const double base[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                                {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (auto& b : base) m.v.push_back(unit({b[0], b[1], b[2]}));
    m.f = {0, 11, 5, 0, 5,  1,  0,  1,  7,  0,  7,  10, 0, 10, 11, 1, 5, 9, 5, 11, 4,  11, 10, 2,  10, 7, 6, 7, 1, 8,
           3, 9,  4, 3, 4,  2,  3,  2,  6,  3,  6,  8,  3, 8,  9,  4, 9, 5, 2, 4,  11, 6,  2,  10, 8,  6, 7, 9, 8, 1};
    for (int s = 0; s < subdiv; s++) {
        std::map<uint64_t, uint32_t> mid;
        auto midpoint = [&](uint32_t a, uint32_t b) {
            const uint64_t key = a < b ? (uint64_t(a) << 32 | b) : (uint64_t(b) << 32 | a);
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            m.v.push_back(unit((m.v[a] + m.v[b]) * 0.5));
            const uint32_t id = static_cast<uint32_t>(m.v.size() - 1);
            mid[key] = id;
            return id;
        };
        std::vector<uint32_t> nf;
        nf.reserve(m.f.size() * 4);
        for (size_t i = 0; i < m.f.size(); i += 3) {
            const uint32_t a = m.f[i], b = m.f[i + 1], c = m.f[i + 2];
            const uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            nf.insert(nf.end(), {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca});
        }
        m.f.swap(nf);
    }
    return m;
}

// value noise on an integer lattice with a seeded hash, smooth-step interpolated
double lattice(int64_t x, int64_t y, int64_t z, uint64_t seed)
{
    uint64_t h = seed ^ (static_cast<uint64_t>(x) * 0x8da6b343ull) ^ (static_cast<uint64_t>(y) * 0xd8163841ull) ^
                 (static_cast<uint64_t>(z) * 0xcb1ab31full);
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdull;
    h ^= h >> 33;
    h *= 0xc4ceb9fe1a85ec53ull;
    h ^= h >> 33;
    return static_cast<double>(h >> 11) * 0x1.0p-53 * 2.0 - 1.0;
}

double value_noise(P3 p, uint64_t seed)
{
    const double fx = std::floor(p.x), fy = std::floor(p.y), fz = std::floor(p.z);
    const int64_t ix = static_cast<int64_t>(fx), iy = static_cast<int64_t>(fy), iz = static_cast<int64_t>(fz);
    auto sm = [](double t) { return t * t * (3.0 - 2.0 * t); };
    const double u = sm(p.x - fx), v = sm(p.y - fy), w = sm(p.z - fz);
    double acc = 0.0;
    for (int dz = 0; dz < 2; dz++)
        for (int dy = 0; dy < 2; dy++)
            for (int dx = 0; dx < 2; dx++) {
                const double wt = (dx ? u : 1 - u) * (dy ? v : 1 - v) * (dz ? w : 1 - w);
                acc += wt * lattice(ix + dx, iy + dy, iz + dz, seed);
            }
    return acc;
}

void compute_vertex_normals(Mesh& m)
{
    m.n.assign(m.v.size(), {0, 0, 0});
    for (size_t i = 0; i < m.f.size(); i += 3) {
        const P3 a = m.v[m.f[i]], b = m.v[m.f[i + 1]], c = m.v[m.f[i + 2]];
        const P3 fn = cross(b - a, c - a); // area weighted
        for (int k = 0; k < 3; k++) m.n[m.f[i + k]] = m.n[m.f[i + k]] + fn;
    }
    for (auto& n : m.n) n = unit(n);
}

struct Writer {
    FILE* obj = nullptr;
    uint32_t vbase = 0, nbase = 0;
    void usemtl(const char* name) { fprintf(obj, "usemtl %s\n", name); }
    // quad p0 p1 p2 p3 (counter-clockwise seen from the side the face normal points to)
    void quad(P3 a, P3 b, P3 c, P3 d)
    {
        for (P3 p : {a, b, c, d}) fprintf(obj, "v %.6f %.6f %.6f\n", p.x, p.y, p.z);
        fprintf(obj, "f %u %u %u\n", vbase + 1, vbase + 2, vbase + 3);
        fprintf(obj, "f %u %u %u\n", vbase + 1, vbase + 3, vbase + 4);
        vbase += 4;
    }
    void mesh(const Mesh& m, P3 center, double scale, bool with_normals)
    {
        for (const P3& p : m.v) {
            const P3 q = center + p * scale;
            fprintf(obj, "v %.6f %.6f %.6f\n", q.x, q.y, q.z);
        }
        if (with_normals)
            for (const P3& n : m.n) fprintf(obj, "vn %.5f %.5f %.5f\n", n.x, n.y, n.z);
        for (size_t i = 0; i < m.f.size(); i += 3) {
            if (with_normals)
                fprintf(obj, "f %u//%u %u//%u %u//%u\n", vbase + m.f[i] + 1, nbase + m.f[i] + 1,
                        vbase + m.f[i + 1] + 1, nbase + m.f[i + 1] + 1, vbase + m.f[i + 2] + 1,
                        nbase + m.f[i + 2] + 1);
            else
                fprintf(obj, "f %u %u %u\n", vbase + m.f[i] + 1, vbase + m.f[i + 1] + 1, vbase + m.f[i + 2] + 1);
        }
        vbase += static_cast<uint32_t>(m.v.size());
        if (with_normals) nbase += static_cast<uint32_t>(m.n.size());
    }
};

std::string base_name(const std::string& p)
{
    const size_t s = p.find_last_of('/');
    return s == std::string::npos ? p : p.substr(s + 1);
}

std::string strip_ext(const std::string& p)
{
    const size_t d = p.find_last_of('.');
    return d == std::string::npos ? p : p.substr(0, d);
}

void camera_aux(FILE* aux, P3 origin, P3 target, P3 up, double fov_deg)
{
    const P3 back = unit(origin - target); // TM_ROW2: origin - target direction
    fprintf(aux, "TM_ROW0 1 0 0\n");
    fprintf(aux, "TM_ROW1 %.6f %.6f %.6f\n", up.x, up.y, up.z);
    fprintf(aux, "TM_ROW2 %.6f %.6f %.6f\n", back.x, back.y, back.z);
    fprintf(aux, "TM_ROW3 %.6f %.6f %.6f\n", origin.x, origin.y, origin.z);
    fprintf(aux, "CAMERA_FOV %.7f\n", fov_deg * M_PI / 180.0);
    fprintf(aux, "CAMERA_TDIST 1.0\n");
}

int gen_box(const std::string& obj_path, uint32_t seed, bool furnace, bool fog = false)
{
    const std::string mtl_path = strip_ext(obj_path) + ".mtl";
    FILE* obj = fopen(obj_path.c_str(), "w");
    FILE* mtl = fopen(mtl_path.c_str(), "w");
    FILE* aux = fopen((obj_path + ".aux").c_str(), "w");
    if (!obj || !mtl || !aux) return UPBP_E_IO;
    Rng rng(seed);
    const char* walls[6] = {"wall_floor", "wall_ceiling", "wall_back", "wall_left", "wall_right", "wall_front"};
    for (int i = 0; i < 6; i++) {
        if (furnace && i == 0) continue; // the floor is the emitter
        double r = 1, g = 1, b = 1;
        if (!furnace) {
            r = rng.uniform(0.5, 0.8);
            g = rng.uniform(0.5, 0.8);
            b = rng.uniform(0.5, 0.8);
        }
        if (fog && i == 2) // glossy back wall
            fprintf(mtl, "newmtl %s\nNs 30\nKd %.6f %.6f %.6f\nKs 0.25 0.25 0.25\nKe 0 0 0\n\n", walls[i], 0.6 * r,
                    0.6 * g, 0.6 * b);
        else
            fprintf(mtl, "newmtl %s\nNs 0\nKd %.6f %.6f %.6f\nKs 0 0 0\nKe 0 0 0\n\n", walls[i], r, g, b);
    }
    if (furnace) fprintf(mtl, "newmtl light\nNs 0\nKd 0 0 0\nKs 0 0 0\nKe 1 1 1\n\n");
    else fprintf(mtl, "newmtl light\nNs 0\nKd 0 0 0\nKs 0 0 0\nKe 10 10 10\n\n");
    if (!furnace) fprintf(mtl, "newmtl sphere_boundary\nNs 0\nKd 0 0 0\nKs 0 0 0\nKe 0 0 0\n\n");

    Writer w;
    w.obj = obj;
    fprintf(obj, "mtllib %s\n", base_name(mtl_path).c_str());
    // inward-facing walls of [-1,1]^3
    w.usemtl(furnace ? "light" : walls[0]);
    w.quad({-1, -1, -1}, {-1, -1, 1}, {1, -1, 1}, {1, -1, -1}); // floor, normal +y
    w.usemtl(walls[1]);
    w.quad({-1, 1, -1}, {1, 1, -1}, {1, 1, 1}, {-1, 1, 1}); // ceiling, normal -y
    w.usemtl(walls[2]);
    w.quad({-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}); // back, normal +z
    w.usemtl(walls[3]);
    w.quad({-1, -1, -1}, {-1, 1, -1}, {-1, 1, 1}, {-1, -1, 1}); // left, normal +x
    w.usemtl(walls[4]);
    w.quad({1, -1, -1}, {1, -1, 1}, {1, 1, 1}, {1, 1, -1}); // right, normal -x
    if (furnace) {
        w.usemtl(walls[5]);
        w.quad({-1, -1, 1}, {-1, 1, 1}, {1, 1, 1}, {1, -1, 1}); // front, normal -z
    }
    // 0.5 x 0.5 light under the ceiling, facing down
    if (!furnace) {
        w.usemtl("light");
        w.quad({-0.25, 0.999, -0.25}, {0.25, 0.999, -0.25}, {0.25, 0.999, 0.25}, {-0.25, 0.999, 0.25});
    }
    if (!furnace) {
        Mesh s = icosphere(3);
        w.usemtl("sphere_boundary");
        w.mesh(s, {0, 0, 0}, fog ? 0.55 : 0.3, false);
    }
    fclose(obj);
    fclose(mtl);

    if (furnace)
        camera_aux(aux, {0, 0, 0.9}, {0, 0, 0}, {0, 1, 0}, 40.0);
    else
        camera_aux(aux, {0, 0, 3.2}, {0, 0, 0}, {0, 1, 0}, 40.0);
    if (!furnace) {
        if (fog)
            fprintf(aux, "medium sphere_medium\nabsorption 0.3 0.3 0.3\nemission 0 0 0\nscattering 2.5 2.2 2.0\ng 0.5\n\n");
        else
            fprintf(aux, "medium sphere_medium\nabsorption 0.1 0.1 0.1\nemission 0 0 0\nscattering 0.5 0.5 0.5\ng 0\n\n");
        fprintf(aux, "material sphere_boundary\ngeometryType imaginary\nmediumId sphere_medium\npriority 1\n");
    }
    fclose(aux);
    return UPBP_OK;
}

int gen_blobs(const std::string& obj_path, uint32_t seed, int grid_n = 8, int n_media = 3)
{
    const std::string mtl_path = strip_ext(obj_path) + ".mtl";
    FILE* obj = fopen(obj_path.c_str(), "w");
    FILE* mtl = fopen(mtl_path.c_str(), "w");
    FILE* aux = fopen((obj_path + ".aux").c_str(), "w");
    if (!obj || !mtl || !aux) return UPBP_E_IO;
    Rng rng(seed);

    // 4 blob materials (diffuse and glossy), a floor, lights, 3 medium boundaries
    const int n_blob_mats = 4;
    for (int i = 0; i < n_blob_mats; i++) {
        const double r = rng.uniform(0.3, 0.8), g = rng.uniform(0.3, 0.8), b = rng.uniform(0.3, 0.8);
        if (i % 2 == 0)
            fprintf(mtl, "newmtl blob%d\nNs 0\nKd %.5f %.5f %.5f\nKs 0 0 0\nKe 0 0 0\n\n", i, r, g, b);
        else
            fprintf(mtl, "newmtl blob%d\nNs 40\nKd %.5f %.5f %.5f\nKs 0.2 0.2 0.2\nKe 0 0 0\n\n", i, r * 0.7,
                    g * 0.7, b * 0.7);
    }
    fprintf(mtl, "newmtl floor\nNs 0\nKd 0.6 0.6 0.6\nKs 0 0 0\nKe 0 0 0\n\n");
    fprintf(mtl, "newmtl light_a\nNs 0\nKd 0 0 0\nKs 0 0 0\nKe 14 12 10\n\n");
    fprintf(mtl, "newmtl light_b\nNs 0\nKd 0 0 0\nKs 0 0 0\nKe 6 8 12\n\n");
    for (int i = 0; i < n_media; i++) fprintf(mtl, "newmtl fog%d\nNs 0\nKd 0 0 0\nKs 0 0 0\nKe 0 0 0\n\n", i);

    Writer w;
    w.obj = obj;
    fprintf(obj, "mtllib %s\n", base_name(mtl_path).c_str());
    w.usemtl("floor");
    w.quad({-6, -0.5, -6}, {-6, -0.5, 6}, {6, -0.5, 6}, {6, -0.5, -6});
    w.usemtl("light_a");
    w.quad({-2.5, 4.0, -1.0}, {-1.0, 4.0, -1.0}, {-1.0, 4.0, 0.5}, {-2.5, 4.0, 0.5});
    w.usemtl("light_b");
    w.quad({1.5, 3.5, 0.5}, {2.5, 3.5, 0.5}, {2.5, 3.5, 1.5}, {1.5, 3.5, 1.5});

    const Mesh base = icosphere(5);
    const double half = 0.5 * double(grid_n - 1);
    for (int gy = 0; gy < grid_n; gy++) {
        for (int gx = 0; gx < grid_n; gx++) {
            Mesh m = base;
            const uint64_t nseed = rng.next();
            const double amp = rng.uniform(0.08, 0.2);
            const double freq = rng.uniform(1.5, 3.0);
            for (P3& p : m.v) p = p * (1.0 + amp * value_noise(p * freq, nseed));
            compute_vertex_normals(m);
            char name[16];
            snprintf(name, sizeof(name), "blob%d", static_cast<int>(rng.next() % n_blob_mats));
            w.usemtl(name);
            const P3 c = {(gx - half) * 1.0, 0.0, (gy - half) * 1.0};
            w.mesh(m, c, 0.38, true);
        }
    }
    const Mesh fog = icosphere(3);
    const P3 fog_c3[3] = {{-1.2, 0.2, -0.8}, {1.0, 0.4, 0.3}, {0.0, 0.0, 1.4}};
    const double fog_r3[3] = {2.2, 1.8, 1.6};
    for (int i = 0; i < n_media; i++) {
        char name[16];
        snprintf(name, sizeof(name), "fog%d", i);
        w.usemtl(name);
        if (n_media == 3) {
            w.mesh(fog, fog_c3[i], fog_r3[i], false);
        } else { // a ring of overlapping spheres over the grid
            const double a = 2.0 * 3.14159265358979 * double(i) / double(n_media);
            w.mesh(fog, {2.6 * std::cos(a), 0.1 + 0.1 * (i % 3), 2.6 * std::sin(a)}, 1.5 + 0.1 * (i % 4), false);
        }
    }
    fclose(obj);
    fclose(mtl);

    if (grid_n == 8)
        camera_aux(aux, {0.0, 4.5, 7.5}, {0.0, -0.3, 0.0}, {0, 1, 0}, 55.0);
    else
        camera_aux(aux, {0.0, 5.5, 9.5}, {0.0, -0.3, 0.0}, {0, 1, 0}, 55.0);
    const double gs3[3] = {0.0, 0.5, 0.9}, gs8[8] = {-0.3, 0.0, 0.3, 0.5, 0.7, 0.8, 0.9, 0.95};
    for (int i = 0; i < n_media; i++) {
        const double g = n_media == 3 ? gs3[i] : gs8[i % 8];
        const double st = rng.uniform(0.2, 2.0);
        const double albedo = rng.uniform(0.6, 0.95);
        fprintf(aux, "medium fogmed%d\nabsorption %.5f %.5f %.5f\nemission 0 0 0\nscattering %.5f %.5f %.5f\ng %.2f\n\n",
                i, st * (1 - albedo), st * (1 - albedo), st * (1 - albedo), st * albedo, st * albedo * 0.9,
                st * albedo * 0.8, g);
        fprintf(aux, "material fog%d\ngeometryType imaginary\nmediumId fogmed%d\npriority %d\n\n", i, i, i + 1);
    }
    fprintf(aux, "light_point 2.5 2.0 3.0 6.0 6.0 6.0\n");
    fclose(aux);
    return UPBP_OK;
}

// n x n grid of quads spanning origin + u * [0,1] + v * [0,1] (normal along u x v), shared
// vertices; cells whose centre satisfies skip(u, v) are left out
template <typename Skip>
void grid(Writer& w, P3 origin, P3 eu, P3 ev, int n, Skip skip)
{
    for (int j = 0; j <= n; j++)
        for (int i = 0; i <= n; i++) {
            const P3 p = origin + eu * (double(i) / n) + ev * (double(j) / n);
            fprintf(w.obj, "v %.6f %.6f %.6f\n", p.x, p.y, p.z);
        }
    auto id = [&](int i, int j) { return w.vbase + uint32_t(j * (n + 1) + i) + 1; };
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            if (skip((i + .5) / n, (j + .5) / n)) continue;
            fprintf(w.obj, "f %u %u %u\n", id(i, j), id(i + 1, j), id(i + 1, j + 1));
            fprintf(w.obj, "f %u %u %u\n", id(i, j), id(i + 1, j + 1), id(i, j + 1));
        }
    w.vbase += uint32_t((n + 1) * (n + 1));
}

// Room x in [-2, 2], y in [0, 3], z in [-3, 3] of walls tessellated n x n per face (n = 128:
// about 196k triangles), with a square aperture of 5% of the x = +2 wall's area centred at
// y = 2, z = 0; the only light is a quad outside that wall facing the aperture. A dense
// forward-scattering medium slab (imaginary box x in [-1.5, 1.5], y in [0.5, 2.5],
// z in [-0.5, 0.5]; sigma_t ~ U[2, 10], albedo 0.95, g 0.9) sits between the camera and the
// far wall, so most of what the camera sees arrives through the aperture after several
// scattering events in the slab.
int gen_slabroom(const std::string& obj_path, uint32_t seed, int n)
{
    const std::string mtl_path = strip_ext(obj_path) + ".mtl";
    FILE* obj = fopen(obj_path.c_str(), "w");
    FILE* mtl = fopen(mtl_path.c_str(), "w");
    FILE* aux = fopen((obj_path + ".aux").c_str(), "w");
    if (!obj || !mtl || !aux) return UPBP_E_IO;
    Rng rng(seed);
    fprintf(mtl, "newmtl walls\nNs 0\nKd 0.6 0.6 0.6\nKs 0 0 0\nKe 0 0 0\n\n");
    fprintf(mtl, "newmtl floor\nNs 0\nKd 0.45 0.4 0.35\nKs 0 0 0\nKe 0 0 0\n\n");
    fprintf(mtl, "newmtl light\nNs 0\nKd 0 0 0\nKs 0 0 0\nKe 400 380 340\n\n");
    fprintf(mtl, "newmtl slab_boundary\nNs 0\nKd 0 0 0\nKs 0 0 0\nKe 0 0 0\n\n");
    Writer w;
    w.obj = obj;
    fprintf(obj, "mtllib %s\n", base_name(mtl_path).c_str());
    auto none = [](double, double) { return false; };
    // inward-facing walls
    w.usemtl("floor");
    grid(w, {-2, 0, -3}, {0, 0, 6}, {4, 0, 0}, n, none); // y = 0, normal +y
    w.usemtl("walls");
    grid(w, {-2, 3, -3}, {4, 0, 0}, {0, 0, 6}, n, none);  // ceiling, normal -y
    grid(w, {-2, 0, -3}, {4, 0, 0}, {0, 3, 0}, n, none);  // z = -3, normal +z
    grid(w, {-2, 0, 3}, {0, 3, 0}, {4, 0, 0}, n, none);   // z = +3, normal -z
    grid(w, {-2, 0, -3}, {0, 3, 0}, {0, 0, 6}, n, none);  // x = -2, normal +x
    // x = +2 (normal -x), with the aperture: 5% of 3 x 6 is 0.9, a square of side 0.9487
    const double half = 0.5 * std::sqrt(0.05 * 18.0);
    grid(w, {2, 0, -3}, {0, 0, 6}, {0, 3, 0}, n, [&](double u, double v) {
        const double z = -3 + 6 * u, y = 3 * v;
        return std::fabs(z) < half && std::fabs(y - 2.0) < half;
    });
    // the light outside, facing the aperture (normal -x)
    w.usemtl("light");
    w.quad({3.2, 1.25, -0.75}, {3.2, 1.25, 0.75}, {3.2, 2.75, 0.75}, {3.2, 2.75, -0.75});
    // the slab, outward normals
    const P3 a{-1.5, 0.5, -0.5}, b{1.5, 2.5, 0.5};
    w.usemtl("slab_boundary");
    w.quad({a.x, a.y, a.z}, {b.x, a.y, a.z}, {b.x, a.y, b.z}, {a.x, a.y, b.z}); // -y
    w.quad({a.x, b.y, a.z}, {a.x, b.y, b.z}, {b.x, b.y, b.z}, {b.x, b.y, a.z}); // +y
    w.quad({a.x, a.y, a.z}, {a.x, b.y, a.z}, {b.x, b.y, a.z}, {b.x, a.y, a.z}); // -z
    w.quad({a.x, a.y, b.z}, {b.x, a.y, b.z}, {b.x, b.y, b.z}, {a.x, b.y, b.z}); // +z
    w.quad({a.x, a.y, a.z}, {a.x, a.y, b.z}, {a.x, b.y, b.z}, {a.x, b.y, a.z}); // -x
    w.quad({b.x, a.y, a.z}, {b.x, b.y, a.z}, {b.x, b.y, b.z}, {b.x, a.y, b.z}); // +x
    fclose(obj);
    fclose(mtl);
    camera_aux(aux, {0.0, 1.5, -2.9}, {0.0, 1.5, 0.0}, {0, 1, 0}, 60.0);
    const double st = rng.uniform(2.0, 10.0);
    fprintf(aux, "medium slab\nabsorption %.5f %.5f %.5f\nemission 0 0 0\nscattering %.5f %.5f %.5f\ng 0.9\n\n",
            0.05 * st, 0.05 * st, 0.05 * st, 0.95 * st, 0.95 * st, 0.95 * st);
    fprintf(aux, "material slab_boundary\ngeometryType imaginary\nmediumId slab\npriority 1\n");
    fclose(aux);
    return UPBP_OK;
}

} // namespace

extern "C" int upbp_generate_scene(const char* kind, uint32_t seed, const char* obj_path)
{
    if (!kind || !obj_path) return UPBP_E_ARG;
    const std::string k(kind);
    if (k == "box") return gen_box(obj_path, seed, false);
    if (k == "furnace") return gen_box(obj_path, seed, true);
    if (k == "boxfog") return gen_box(obj_path, seed, false, true);
    if (k == "blobs") return gen_blobs(obj_path, seed);
    if (k == "stress") return gen_blobs(obj_path, seed, 10, 8);
    // batched-jobs mix: 1 to 9 blobs (20k to 184k triangles) and 1 to 4 media, set by the seed
    if (k == "blobs_small") return gen_blobs(obj_path, seed, 1 + int(seed % 3), 1 + int((seed / 3) % 4));
    if (k == "slabroom") return gen_slabroom(obj_path, seed, 128);
    if (k == "slabroom_small") return gen_slabroom(obj_path, seed, 8); // tests: same room, coarse walls
    return UPBP_E_ARG;
}
