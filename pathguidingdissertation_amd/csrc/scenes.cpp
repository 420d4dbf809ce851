// scenes.cpp — procedural scenes of the benchmark configurations.
//
// The reference loads scene.json + .gem meshes (SceneLoader.h:222-269); the
// benchmark configurations are procedural instead, generated here from a seed.
// Geometry uses only +,-,*,/ and sqrt plus self-contained polynomial sin/cos/exp,
// so the same seed yields bit-identical float32 triangles on every IEEE machine.
// Shared vertices are computed once, so meshes are watertight.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/pg_tracer.h"
#include "host_scene.hpp"


namespace {

constexpr double kPi = 3.14159265358979323846;

// ------------------------------------------------------------------ deterministic math
double dpoly_sin(double r) {  // |r| <= pi/4, Taylor to r^17
    double r2 = r * r, t = r, s = r;
    for (int k = 1; k <= 8; ++k) { t = -t * r2 / (double)((2 * k) * (2 * k + 1)); s += t; }
    return s;
}
double dpoly_cos(double r) {
    double r2 = r * r, t = 1.0, s = 1.0;
    for (int k = 1; k <= 8; ++k) { t = -t * r2 / (double)((2 * k - 1) * (2 * k)); s += t; }
    return s;
}
void dsincos(double x, double& s, double& c) {
    double k = std::nearbyint(x * (2.0 / kPi));
    double r = (x - k * 1.5707963267341256) - k * 6.077100506506192e-11;
    long q = (long)k & 3;
    double ps = dpoly_sin(r), pc = dpoly_cos(r);
    switch (q) {
    case 0: s = ps; c = pc; break;
    case 1: s = pc; c = -ps; break;
    case 2: s = -ps; c = -pc; break;
    default: s = -pc; c = ps; break;
    }
}
double dsin(double x) { double s, c; dsincos(x, s, c); return s; }
double dcos(double x) { double s, c; dsincos(x, s, c); return c; }
double dexp(double x) {  // exp for x in [-750, 50]
    if (x < -700.0) return 0.0;
    double k = std::nearbyint(x / 0.6931471805599453);
    double r = x - k * 0.6931471805599453;
    double t = 1.0, s = 1.0;
    for (int i = 1; i <= 20; ++i) { t = t * r / (double)i; s += t; }
    return std::ldexp(s, (int)k);
}

// splitmix64 stream
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uni() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double range(double a, double b) { return a + (b - a) * uni(); }
};

struct D3 {
    double x, y, z;
    D3 operator+(const D3& o) const { return {x + o.x, y + o.y, z + o.z}; }
    D3 operator-(const D3& o) const { return {x - o.x, y - o.y, z - o.z}; }
    D3 operator*(double s) const { return {x * s, y * s, z * s}; }
};
D3 dnorm(const D3& a) { double l = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); return {a.x / l, a.y / l, a.z / l}; }
D3 dcross(const D3& a, const D3& b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

struct Builder {
    pg_host_scene* s;
    void tri(const D3& a, const D3& b, const D3& c, const D3& na, const D3& nb, const D3& nc, uint32_t m) {
        const D3* p[3] = {&a, &b, &c};
        const D3* n[3] = {&na, &nb, &nc};
        for (int k = 0; k < 3; ++k) {
            s->pos.push_back((float)p[k]->x); s->pos.push_back((float)p[k]->y); s->pos.push_back((float)p[k]->z);
            s->nrm.push_back((float)n[k]->x); s->nrm.push_back((float)n[k]->y); s->nrm.push_back((float)n[k]->z);
        }
        s->mat.push_back(m);
    }
    void flat_tri(const D3& a, const D3& b, const D3& c, uint32_t m) {
        D3 n = dnorm(dcross(b - a, c - a));
        tri(a, b, c, n, n, n, m);
    }
    // quad a-b-c-d (counter-clockwise seen from the side the normal points to)
    void quad(const D3& a, const D3& b, const D3& c, const D3& d, uint32_t m) {
        flat_tri(a, b, c, m);
        flat_tri(a, c, d, m);
    }
    uint32_t material(int type, D3 albedo, D3 emission = {0, 0, 0}, double rough = 0.0,
                      D3 eta = {0, 0, 0}, D3 k = {0, 0, 0}, double ior = 1.5) {
        pg_material m;
        std::memset(&m, 0, sizeof(m));
        m.type = (float)type;
        m.albedo[0] = (float)albedo.x; m.albedo[1] = (float)albedo.y; m.albedo[2] = (float)albedo.z;
        m.emission[0] = (float)emission.x; m.emission[1] = (float)emission.y; m.emission[2] = (float)emission.z;
        m.eta[0] = (float)eta.x; m.eta[1] = (float)eta.y; m.eta[2] = (float)eta.z;
        m.k[0] = (float)k.x; m.k[1] = (float)k.y; m.k[2] = (float)k.z;
        m.int_ior = (float)ior; m.ext_ior = 1.f; m.roughness = (float)rough;
        s->mats.push_back(m);
        return (uint32_t)s->mats.size() - 1;
    }
    // box with half extents h, rotation rows R, centre c; 12 triangles, flat normals
    void box(const D3& c, const D3& h, const double R[9], uint32_t m) {
        D3 v[8];
        for (int i = 0; i < 8; ++i) {
            D3 l{(i & 1) ? h.x : -h.x, (i & 2) ? h.y : -h.y, (i & 4) ? h.z : -h.z};
            v[i] = D3{R[0] * l.x + R[1] * l.y + R[2] * l.z, R[3] * l.x + R[4] * l.y + R[5] * l.z,
                      R[6] * l.x + R[7] * l.y + R[8] * l.z} + c;
        }
        quad(v[0], v[2], v[3], v[1], m);  // -z
        quad(v[4], v[5], v[7], v[6], m);  // +z
        quad(v[0], v[4], v[6], v[2], m);  // -x
        quad(v[1], v[3], v[7], v[5], m);  // +x
        quad(v[0], v[1], v[5], v[4], m);  // -y
        quad(v[2], v[6], v[7], v[3], m);  // +y
    }
    // frequency-4 geodesic icosphere (320 faces), smooth normals
    void icosphere(const D3& c, double r, uint32_t m) {
        const double t = (1.0 + std::sqrt(5.0)) / 2.0;
        const D3 V[12] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                          {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
        const int F[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                              {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
                              {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
        const int n = 4;
        for (int f = 0; f < 20; ++f) {
            int ia = F[f][0], ib = F[f][1], ic = F[f][2];
            auto P = [&](int i, int j, D3& p, D3& nn) {
                int k = n - i - j;
                // weights on the corners; zero-weight corners add exactly 0, and the
                // two-corner sums on shared edges are commutative, so edge points match
                D3 s = V[ia] * (double)k + V[ib] * (double)i;
                s = s + V[ic] * (double)j;
                nn = dnorm(s);
                p = c + nn * r;
            };
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n - i; ++j) {
                    D3 p0, p1, p2, n0, n1, n2;
                    P(i, j, p0, n0); P(i + 1, j, p1, n1); P(i, j + 1, p2, n2);
                    tri(p0, p1, p2, n0, n1, n2, m);
                    if (i + j < n - 1) {
                        D3 p3, n3;
                        P(i + 1, j + 1, p3, n3);
                        tri(p1, p3, p2, n1, n3, n2, m);
                    }
                }
        }
    }
};

void random_rotation(Rng& rng, double R[9]) {
    double u1 = rng.uni(), u2 = rng.uni(), u3 = rng.uni();
    double a = std::sqrt(1.0 - u1), b = std::sqrt(u1);
    double x = a * dsin(2 * kPi * u2), y = a * dcos(2 * kPi * u2);
    double z = b * dsin(2 * kPi * u3), w = b * dcos(2 * kPi * u3);
    R[0] = 1 - 2 * (y * y + z * z); R[1] = 2 * (x * y - z * w); R[2] = 2 * (x * z + y * w);
    R[3] = 2 * (x * y + z * w); R[4] = 1 - 2 * (x * x + z * z); R[5] = 2 * (y * z - x * w);
    R[6] = 2 * (x * z - y * w); R[7] = 2 * (y * z + x * w); R[8] = 1 - 2 * (x * x + y * y);
}

void rot_y(double deg, double R[9]) {
    double s, c;
    dsincos(deg * kPi / 180.0, s, c);
    R[0] = c; R[1] = 0; R[2] = s; R[3] = 0; R[4] = 1; R[5] = 0; R[6] = -s; R[7] = 0; R[8] = c;
}

// Config 1: Cornell box. Room [-1,1] x [0,2] x [-1,1] open at z = +1; 5 walls
// (10 tris), two boxes rotated -18 / +18 degrees (24 tris), emissive ceiling quad
// (2 tris, Le = (17,12,4)); 256x256 pinhole camera with a 40 degree FOV.
void cornell(pg_host_scene* s) {
    Builder b{s};
    uint32_t white = b.material(PG_BSDF_DIFFUSE, {0.73, 0.73, 0.73});
    uint32_t red = b.material(PG_BSDF_DIFFUSE, {0.65, 0.05, 0.05});
    uint32_t green = b.material(PG_BSDF_DIFFUSE, {0.12, 0.45, 0.15});
    uint32_t light = b.material(PG_BSDF_DIFFUSE, {0.0, 0.0, 0.0}, {17.0, 12.0, 4.0});
    b.quad({-1, 0, -1}, {-1, 0, 1}, {1, 0, 1}, {1, 0, -1}, white);   // floor
    b.quad({-1, 2, -1}, {1, 2, -1}, {1, 2, 1}, {-1, 2, 1}, white);   // ceiling
    b.quad({-1, 0, -1}, {1, 0, -1}, {1, 2, -1}, {-1, 2, -1}, white); // back
    b.quad({-1, 0, -1}, {-1, 2, -1}, {-1, 2, 1}, {-1, 0, 1}, red);   // left
    b.quad({1, 0, -1}, {1, 0, 1}, {1, 2, 1}, {1, 2, -1}, green);     // right
    double R[9];
    rot_y(-18.0, R);
    b.box({0.35, 0.3, 0.3}, {0.3, 0.3, 0.3}, R, white);
    rot_y(18.0, R);
    b.box({-0.35, 0.6, -0.35}, {0.3, 0.6, 0.3}, R, white);
    // light quad just below the ceiling, normals facing down
    D3 dn{0, -1, 0};
    D3 a{-0.25, 1.98, -0.2}, c1{0.25, 1.98, -0.2}, c2{0.25, 1.98, 0.2}, d{-0.25, 1.98, 0.2};
    b.tri(a, c1, c2, dn, dn, dn, light);
    b.tri(a, c2, d, dn, dn, dn, light);
    s->cam = pg_camera_desc{{0.f, 1.f, 3.4f}, {0.f, 1.f, 0.f}, {0.f, 1.f, 0.f}, 40.f, 256.f, 256.f};
}

// Synthetic 1024x512 lat-long sky (y-up; theta = pi v, phi = 2 pi u as Lights.h:158-164):
// horizon-to-zenith gradient, dark ground, Gaussian sun lobe (sigma 1 degree, peak 5e4).
void sky(pg_host_scene* s, const D3& sun_dir) {
    const int W = 1024, H = 512;
    s->env_w = W; s->env_h = H;
    s->env.assign((size_t)W * H * 3, 0.f);
    const double sigma = 1.0 * kPi / 180.0;
    D3 sd = dnorm(sun_dir);
    for (int v = 0; v < H; ++v) {
        double st, ct;
        dsincos(kPi * (v + 0.5) / H, st, ct);
        for (int u = 0; u < W; ++u) {
            double sp, cp;
            dsincos(2 * kPi * (u + 0.5) / W, sp, cp);
            D3 d{cp * st, ct, sp * st};
            double r, g, bl;
            if (d.y > 0) {
                double t = std::sqrt(d.y);
                r = 1.0 * (1 - t) + 0.25 * t; g = 0.95 * (1 - t) + 0.45 * t; bl = 0.9 * (1 - t) + 1.0 * t;
            } else {
                r = 0.15; g = 0.13; bl = 0.10;
            }
            double c = d.x * sd.x + d.y * sd.y + d.z * sd.z;
            double a2 = 2.0 * (1.0 - c);  // chord^2 ~ angle^2 near the sun
            double sun = 5e4 * dexp(-a2 / (2 * sigma * sigma));
            size_t i = ((size_t)v * W + u) * 3;
            s->env[i] = (float)(r + sun); s->env[i + 1] = (float)(g + 0.95 * sun); s->env[i + 2] = (float)(bl + 0.88 * sun);
        }
    }
}

// Configs 2/3: seeded interior of a 20 x 10 x 20 room: 707 x 707 height-displaced
// floor grid (999,698 tris), 2,000 boxes with random scale/rotation (24,000 tris),
// 200 frequency-4 icospheres (64,000 tris), walls. Materials per object: 50%
// Lambertian albedo U[0.1,0.9], 40% GGX conductor roughness U[0.05,0.8], 10% smooth
// dielectric IOR 1.5. Config 2 is open to the sky; config 3 adds a ceiling and
// closes the walls except one 2 x 2 window in the +x wall.
void interior(pg_host_scene* s, uint32_t seed, bool closed, int grid) {
    Builder b{s};
    Rng rng(seed + 1000003ull * (closed ? 3 : 2));
    uint32_t floor_m = b.material(PG_BSDF_DIFFUSE, {0.5, 0.5, 0.5});
    uint32_t wall_m = b.material(PG_BSDF_DIFFUSE, {0.7, 0.7, 0.7});
    const double L = 10.0;
    auto height = [](double x, double z) {
        double sx, cx, sz, cz, s2, c2;
        dsincos(0.9 * x, sx, cx); dsincos(1.1 * z, sz, cz); dsincos(3.1 * x + 1.7 * z, s2, c2);
        double h = 0.08 * sx * cz + 0.03 * s2;
        double dhdx = 0.08 * 0.9 * cx * cz + 0.03 * 3.1 * c2;
        double dhdz = -0.08 * 1.1 * sx * sz + 0.03 * 1.7 * c2;
        return D3{h, dhdx, dhdz};
    };
    const int G = grid;
    std::vector<D3> gp((size_t)(G + 1) * (G + 1)), gn((size_t)(G + 1) * (G + 1));
    for (int j = 0; j <= G; ++j)
        for (int i = 0; i <= G; ++i) {
            double x = -L + 2 * L * i / G, z = -L + 2 * L * j / G;
            D3 h = height(x, z);
            gp[(size_t)j * (G + 1) + i] = D3{x, h.x, z};
            gn[(size_t)j * (G + 1) + i] = dnorm(D3{-h.y, 1.0, -h.z});
        }
    for (int j = 0; j < G; ++j)
        for (int i = 0; i < G; ++i) {
            size_t a = (size_t)j * (G + 1) + i, bq = a + 1, c = a + (G + 1), d = c + 1;
            b.tri(gp[a], gp[c], gp[bq], gn[a], gn[c], gn[bq], floor_m);
            b.tri(gp[bq], gp[c], gp[d], gn[bq], gn[c], gn[d], floor_m);
        }
    const double H = 10.0;
    // walls (normals inward)
    b.quad({-L, 0, -L}, {L, 0, -L}, {L, H, -L}, {-L, H, -L}, wall_m);   // z = -L
    b.quad({-L, 0, L}, {-L, H, L}, {L, H, L}, {L, 0, L}, wall_m);       // z = +L
    b.quad({-L, 0, -L}, {-L, H, -L}, {-L, H, L}, {-L, 0, L}, wall_m);   // x = -L
    if (!closed) {
        b.quad({L, 0, -L}, {L, 0, L}, {L, H, L}, {L, H, -L}, wall_m);   // x = +L
    } else {
        // x = +L with a 2 x 2 window at y in [4,6], z in [-1,1]
        b.quad({L, 0, -L}, {L, 0, L}, {L, 4, L}, {L, 4, -L}, wall_m);
        b.quad({L, 6, -L}, {L, 6, L}, {L, H, L}, {L, H, -L}, wall_m);
        b.quad({L, 4, -L}, {L, 4, -1}, {L, 6, -1}, {L, 6, -L}, wall_m);
        b.quad({L, 4, 1}, {L, 4, L}, {L, 6, L}, {L, 6, 1}, wall_m);
        b.quad({-L, H, -L}, {L, H, -L}, {L, H, L}, {-L, H, L}, wall_m); // ceiling
    }
    static const double metals[4][6] = {
        {0.143, 0.374, 1.442, 3.983, 2.385, 1.603},   // gold
        {0.200, 0.924, 1.102, 3.912, 2.452, 2.142},   // copper
        {1.657, 0.880, 0.521, 9.224, 6.270, 4.837},   // aluminium
        {0.155, 0.117, 0.138, 4.828, 3.122, 2.147}};  // silver
    auto object_material = [&](Rng& r) -> uint32_t {
        double p = r.uni();
        if (p < 0.5) return b.material(PG_BSDF_DIFFUSE, {r.range(0.1, 0.9), r.range(0.1, 0.9), r.range(0.1, 0.9)});
        if (p < 0.9) {
            double rough = r.range(0.05, 0.8);
            int mi = (int)(r.uni() * 4) & 3;
            return b.material(PG_BSDF_CONDUCTOR, {1, 1, 1}, {0, 0, 0}, rough,
                              {metals[mi][0], metals[mi][1], metals[mi][2]}, {metals[mi][3], metals[mi][4], metals[mi][5]});
        }
        return b.material(PG_BSDF_GLASS, {1, 1, 1}, {0, 0, 0}, 0.0, {0, 0, 0}, {0, 0, 0}, 1.5);
    };
    // objects keep a 3-unit clearance around the camera position
    const D3 eye = closed ? D3{-9.0, 6.0, -6.0} : D3{-8.5, 4.0, 8.5};
    auto clear_of_eye = [&](const D3& c) {
        D3 e = c - eye;
        return e.x * e.x + e.y * e.y + e.z * e.z > 9.0;
    };
    // objects stand on or just above the floor (centre height ~ their largest half
    // extent), so the space above them stays open for light entering the room
    for (int i = 0; i < 2000; ++i) {
        D3 h{rng.range(0.1, 0.6), rng.range(0.1, 0.6), rng.range(0.1, 0.6)};
        double hmax = h.x > h.y ? (h.x > h.z ? h.x : h.z) : (h.y > h.z ? h.y : h.z);
        D3 c{rng.range(-9.3, 9.3), hmax * rng.range(0.6, 1.0) + rng.range(0.0, 0.4), rng.range(-9.3, 9.3)};
        while (!clear_of_eye(c)) c = D3{rng.range(-9.3, 9.3), c.y, rng.range(-9.3, 9.3)};
        double R[9];
        random_rotation(rng, R);
        uint32_t m = object_material(rng);
        b.box(c, h, R, m);
    }
    for (int i = 0; i < 200; ++i) {
        double r = rng.range(0.2, 0.8);
        D3 c{rng.range(-9.0, 9.0), r + rng.range(0.0, 0.3), rng.range(-9.0, 9.0)};
        while (!clear_of_eye(c)) c = D3{rng.range(-9.0, 9.0), c.y, rng.range(-9.0, 9.0)};
        uint32_t m = object_material(rng);
        b.icosphere(c, r, m);
    }
    // sun 40 degrees above the +x horizon: enters the config-3 window
    sky(s, D3{0.766044443118978, 0.642787609686539, 0.0});
    if (!closed)
        s->cam = pg_camera_desc{{-8.5f, 4.0f, 8.5f}, {2.f, 1.f, -2.f}, {0.f, 1.f, 0.f}, 60.f, 1280.f, 720.f};
    else
        s->cam = pg_camera_desc{{-9.f, 6.f, -6.f}, {4.f, 1.f, 0.f}, {0.f, 1.f, 0.f}, 60.f, 720.f, 1280.f};
}

}  // namespace

extern "C" {

pg_status pg_scene_generate(int32_t config, uint32_t seed, pg_host_scene** out) {
    if (!out) return PG_ERR_INVALID;
    *out = nullptr;
    pg_host_scene* s = new (std::nothrow) pg_host_scene();
    if (!s) return PG_ERR_NOMEM;
    switch (config) {
    case 1: cornell(s); break;
    case 2: interior(s, seed, false, 707); break;
    case 3: interior(s, seed, true, 707); break;
    case 5: {
        // config-2 family: 100k..1M triangles from the seed
        Rng r(seed * 7919ull + 17ull);
        double target = 100000.0 + 900000.0 * r.uni();
        int grid = (int)std::sqrt((target - 88008.0) / 2.0);
        if (grid < 8) grid = 8;
        interior(s, seed, false, grid);
        break;
    }
    case 100: {
        // small test scene: open interior with a 64 x 64 floor (tests / smoke)
        interior(s, seed, false, 64);
        break;
    }
    case 101: {
        interior(s, seed, true, 64);
        break;
    }
    default: delete s; return PG_ERR_INVALID;
    }
    *out = s;
    return PG_OK;
}

pg_status pg_scene_sizes(const pg_host_scene* s, uint32_t* n_tris, uint32_t* n_materials, int32_t* env_width,
                         int32_t* env_height) {
    if (!s) return PG_ERR_INVALID;
    if (n_tris) *n_tris = (uint32_t)s->mat.size();
    if (n_materials) *n_materials = (uint32_t)s->mats.size();
    if (env_width) *env_width = s->env_w;
    if (env_height) *env_height = s->env_h;
    return PG_OK;
}

pg_status pg_scene_copy(const pg_host_scene* s, float* positions, float* normals, uint32_t* material_index,
                        pg_material* materials, float* env_rgb, pg_camera_desc* camera) {
    if (!s) return PG_ERR_INVALID;
    if (positions) std::memcpy(positions, s->pos.data(), s->pos.size() * sizeof(float));
    if (normals) std::memcpy(normals, s->nrm.data(), s->nrm.size() * sizeof(float));
    if (material_index) std::memcpy(material_index, s->mat.data(), s->mat.size() * sizeof(uint32_t));
    if (materials) std::memcpy(materials, s->mats.data(), s->mats.size() * sizeof(pg_material));
    if (env_rgb && !s->env.empty()) std::memcpy(env_rgb, s->env.data(), s->env.size() * sizeof(float));
    if (camera) *camera = s->cam;
    return PG_OK;
}

pg_status pg_scene_free(pg_host_scene* s) {
    delete s;
    return PG_OK;
}

}  // extern "C"
