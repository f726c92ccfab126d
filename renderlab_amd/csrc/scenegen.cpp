// Seeded synthetic scenes and cameras for the benchmark configurations (the
// reference's own fixture generator is src/bench/city.rs: boxes and rocks
// from a seed). Everything here is deterministic from the seed: splitmix64
// for the scene, Box-Muller for normal deviates, Shoemake's uniform
// quaternions for rotations.
#include <cmath>
#include <vector>

#include "internal.h"

namespace rq {
namespace {

struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double range(double a, double b) { return a + (b - a) * uniform(); }
    double normal() {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

struct Mat {
    double m[16];  // column-major
};
Mat identity() {
    Mat r{};
    r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0;
    return r;
}
Mat mul(const Mat& a, const Mat& b) {
    Mat r{};
    for (int c = 0; c < 4; ++c)
        for (int row = 0; row < 4; ++row) {
            double s = 0;
            for (int k = 0; k < 4; ++k) s += a.m[k * 4 + row] * b.m[c * 4 + k];
            r.m[c * 4 + row] = s;
        }
    return r;
}
Mat rotation(Rng& g) {
    double u1 = g.uniform(), u2 = g.uniform(), u3 = g.uniform();
    double q0 = std::sqrt(1 - u1) * std::sin(6.283185307179586 * u2);
    double q1 = std::sqrt(1 - u1) * std::cos(6.283185307179586 * u2);
    double q2 = std::sqrt(u1) * std::sin(6.283185307179586 * u3);
    double q3 = std::sqrt(u1) * std::cos(6.283185307179586 * u3);
    double x = q0, y = q1, z = q2, w = q3;
    Mat r = identity();
    r.m[0] = 1 - 2 * (y * y + z * z);
    r.m[1] = 2 * (x * y + z * w);
    r.m[2] = 2 * (x * z - y * w);
    r.m[4] = 2 * (x * y - z * w);
    r.m[5] = 1 - 2 * (x * x + z * z);
    r.m[6] = 2 * (y * z + x * w);
    r.m[8] = 2 * (x * z + y * w);
    r.m[9] = 2 * (y * z - x * w);
    r.m[10] = 1 - 2 * (x * x + y * y);
    return r;
}
Mat scale(double x, double y, double z) {
    Mat r = identity();
    r.m[0] = x;
    r.m[5] = y;
    r.m[10] = z;
    return r;
}
Mat translate(double x, double y, double z) {
    Mat r = identity();
    r.m[12] = x;
    r.m[13] = y;
    r.m[14] = z;
    return r;
}
void to_f32(const Mat& a, float out[16]) {
    for (int i = 0; i < 16; ++i) out[i] = (float)a.m[i];
}

struct MeshData {
    std::vector<float> p;
    std::vector<uint32_t> i;
};

MeshData unit_box() {
    MeshData m;
    for (int c = 0; c < 8; ++c) {
        m.p.push_back((c & 1) ? 1.0f : -1.0f);
        m.p.push_back((c & 2) ? 1.0f : -1.0f);
        m.p.push_back((c & 4) ? 1.0f : -1.0f);
    }
    const uint32_t f[12][3] = {{0, 2, 1}, {1, 2, 3}, {4, 5, 6}, {5, 7, 6}, {0, 1, 4}, {1, 5, 4},
                               {2, 6, 3}, {3, 6, 7}, {0, 4, 2}, {2, 4, 6}, {1, 3, 5}, {3, 7, 5}};
    for (auto& t : f) m.i.insert(m.i.end(), t, t + 3);
    return m;
}

// UV sphere, `seg` segments by `rings` rings: seg*(rings-2)*2 + 2*seg triangles.
MeshData uv_sphere(uint32_t seg, uint32_t rings) {
    MeshData m;
    m.p.insert(m.p.end(), {0.0f, 1.0f, 0.0f});
    for (uint32_t r = 1; r < rings; ++r) {
        double th = M_PI * r / rings;
        for (uint32_t s = 0; s < seg; ++s) {
            double ph = 2 * M_PI * s / seg;
            m.p.push_back((float)(std::sin(th) * std::cos(ph)));
            m.p.push_back((float)std::cos(th));
            m.p.push_back((float)(std::sin(th) * std::sin(ph)));
        }
    }
    m.p.insert(m.p.end(), {0.0f, -1.0f, 0.0f});
    uint32_t south = 1 + (rings - 1) * seg;
    auto at = [&](uint32_t r, uint32_t s) { return 1 + (r - 1) * seg + (s % seg); };
    for (uint32_t s = 0; s < seg; ++s) m.i.insert(m.i.end(), {0u, at(1, s + 1), at(1, s)});
    for (uint32_t r = 1; r + 1 < rings; ++r)
        for (uint32_t s = 0; s < seg; ++s) {
            m.i.insert(m.i.end(), {at(r, s), at(r, s + 1), at(r + 1, s)});
            m.i.insert(m.i.end(), {at(r, s + 1), at(r + 1, s + 1), at(r + 1, s)});
        }
    for (uint32_t s = 0; s < seg; ++s) m.i.insert(m.i.end(), {at(rings - 1, s), at(rings - 1, s + 1), south});
    return m;
}

MeshData quad(float half) {
    MeshData m;
    m.p = {-half, 0.0f, -half, half, 0.0f, -half, -half, 0.0f, half, half, 0.0f, half};
    m.i = {0, 2, 1, 1, 2, 3};
    return m;
}

// A cluster of `n` triangles with lognormal edge lengths (median `median`,
// sigma `sigma`), centres uniform in a ball sized to the cluster.
MeshData cluster(Rng& g, uint32_t n, double median, double sigma) {
    MeshData m;
    double radius = 0.5 * median * std::sqrt((double)n) + 0.5;
    for (uint32_t t = 0; t < n; ++t) {
        double cx, cy, cz;
        do {
            cx = g.range(-1, 1);
            cy = g.range(-1, 1);
            cz = g.range(-1, 1);
        } while (cx * cx + cy * cy + cz * cz > 1.0);
        double len = median * std::exp(sigma * g.normal());
        if (len > 40.0 * median) len = 40.0 * median;
        Mat r = rotation(g);
        double rad = len / std::sqrt(3.0);
        for (int k = 0; k < 3; ++k) {
            double a = 2.0943951023931953 * k + g.range(-0.3, 0.3);
            double lx = rad * std::cos(a) * g.range(0.8, 1.2), lz = rad * std::sin(a) * g.range(0.8, 1.2);
            double x = r.m[0] * lx + r.m[8] * lz, y = r.m[1] * lx + r.m[9] * lz, z = r.m[2] * lx + r.m[10] * lz;
            m.p.push_back((float)(cx * radius + x));
            m.p.push_back((float)(cy * radius + y));
            m.p.push_back((float)(cz * radius + z));
            m.i.push_back(3 * t + k);
        }
    }
    return m;
}

}  // namespace
}  // namespace rq

using namespace rq;

static rq_status gen_scene(uint32_t kind, uint64_t seed, uint32_t meshes, uint64_t triangles, bool reflection,
                           rq_scene** out, bool gpu = false, void* stream = nullptr) {
    if (!out || kind > 1) {
        set_error("rq_gen_scene: bad argument");
        return RQ_ERR_INVALID_INPUT;
    }
    Rng g{seed * 0x2545F4914F6CDD1Dull + 1};
    std::vector<MeshData> data;
    std::vector<rq_instance> inst;
    auto add = [&](uint32_t mesh, const Mat& model, uint32_t placement) {
        rq_instance r;
        std::memset(&r, 0, sizeof r);
        to_f32(model, r.model);
        r.mesh = mesh;
        r.identity = mesh + 1;
        r.placement = placement;
        inst.push_back(r);
    };
    if (kind == 0) {
        // 2,000 boxes, 8 UV spheres (32 x 16) and a 200 m ground quad in a 100 m cube.
        data.push_back(unit_box());
        data.push_back(uv_sphere(32, 16));
        data.push_back(quad(100.0f));
        for (uint32_t b = 0; b < 2000; ++b) {
            Mat s = scale(g.range(0.25, 2), g.range(0.25, 2), g.range(0.25, 2));
            Mat r = rotation(g);
            Mat t = translate(g.range(-50, 50), g.range(-50, 50), g.range(-50, 50));
            add(0, mul(t, mul(r, s)), (uint32_t)inst.size());
        }
        for (uint32_t s = 0; s < 8; ++s) {
            double rad = g.range(2, 6);
            Mat t = translate(g.range(-40, 40), g.range(-40, 40), g.range(-40, 40));
            add(1, mul(t, mul(rotation(g), scale(rad, rad, rad))), (uint32_t)inst.size());
        }
        add(2, translate(0, -50, 0), (uint32_t)inst.size());
    } else {
        // `meshes` lognormal-edge clusters spread uniformly in a 200 m cube, plus ground.
        if (meshes == 0) meshes = 5000;
        if (triangles < meshes) triangles = meshes;
        uint64_t per = triangles / meshes, extra = triangles % meshes;
        for (uint32_t m = 0; m < meshes; ++m) {
            uint32_t n = (uint32_t)(per + (m < extra ? 1 : 0));
            data.push_back(cluster(g, n, 0.3, 1.0));
            Mat t = translate(g.range(-100, 100), g.range(-100, 100), g.range(-100, 100));
            add(m, mul(t, rotation(g)), m);
        }
        data.push_back(quad(300.0f));
        add(meshes, translate(0, -100, 0), meshes);
    }
    std::vector<rq_mesh> ms(data.size());
    for (size_t m = 0; m < data.size(); ++m) {
        ms[m].positions = data[m].p.data();
        ms[m].vertex_count = (uint32_t)(data[m].p.size() / 3);
        ms[m].indices = data[m].i.data();
        ms[m].index_count = (uint32_t)data[m].i.size();
        ms[m].keep_primitives = 1;
    }
    if (gpu)
        return reflection ? rq_scene_build_reflection_gpu(ms.data(), (uint32_t)ms.size(), inst.data(),
                                                          (uint32_t)inst.size(), -1, out, stream)
                          : rq_scene_build_gpu(ms.data(), (uint32_t)ms.size(), inst.data(), (uint32_t)inst.size(), -1,
                                               out, stream);
    return reflection
               ? rq_scene_build_reflection(ms.data(), (uint32_t)ms.size(), inst.data(), (uint32_t)inst.size(), -1, out)
               : rq_scene_build(ms.data(), (uint32_t)ms.size(), inst.data(), (uint32_t)inst.size(), -1, out);
}

extern "C" rq_status rq_gen_scene(uint32_t kind, uint64_t seed, uint32_t meshes, uint64_t triangles, rq_scene** out) {
    return gen_scene(kind, seed, meshes, triangles, false, out);
}

extern "C" rq_status rq_gen_scene_gpu(uint32_t kind, uint64_t seed, uint32_t meshes, uint64_t triangles,
                                      rq_scene** out, void* stream) {
    return gen_scene(kind, seed, meshes, triangles, false, out, true, stream);
}

extern "C" rq_status rq_gen_reflection_scene(uint32_t kind, uint64_t seed, uint32_t meshes, uint64_t triangles,
                                             rq_scene** out) {
    return gen_scene(kind, seed, meshes, triangles, true, out);
}

extern "C" rq_status rq_gen_reflection_scene_gpu(uint32_t kind, uint64_t seed, uint32_t meshes, uint64_t triangles,
                                                 rq_scene** out, void* stream) {
    return gen_scene(kind, seed, meshes, triangles, true, out, true, stream);
}

extern "C" rq_status rq_gen_camera(uint32_t kind, uint64_t seed, uint32_t width, uint32_t height, float inverse[16]) {
    if (!inverse || width == 0 || height == 0 || kind > 1) {
        set_error("rq_gen_camera: bad argument");
        return RQ_ERR_INVALID_INPUT;
    }
    Rng g{seed * 0x9E3779B97F4A7C15ull + 7};
    double radius = kind == 0 ? 120.0 : 240.0;
    double az = g.range(0, 2 * M_PI), el = g.range(0.25, 0.5);
    double eye[3] = {radius * std::cos(el) * std::cos(az), radius * std::sin(el), radius * std::cos(el) * std::sin(az)};
    double at[3] = {0.0, kind == 0 ? -10.0 : -30.0, 0.0};
    double f[3] = {at[0] - eye[0], at[1] - eye[1], at[2] - eye[2]};
    double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (double& v : f) v /= fl;
    double up[3] = {0, 1, 0};
    double s[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
    double sl = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    for (double& v : s) v /= sl;
    double u[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2], s[0] * f[1] - s[1] * f[0]};
    Mat view = identity();
    for (int k = 0; k < 3; ++k) {
        view.m[k * 4 + 0] = s[k];
        view.m[k * 4 + 1] = u[k];
        view.m[k * 4 + 2] = -f[k];
    }
    view.m[12] = -(s[0] * eye[0] + s[1] * eye[1] + s[2] * eye[2]);
    view.m[13] = -(u[0] * eye[0] + u[1] * eye[1] + u[2] * eye[2]);
    view.m[14] = (f[0] * eye[0] + f[1] * eye[1] + f[2] * eye[2]);
    double n = 0.1, fa = 10000.0, fy = 1.0 / std::tan(M_PI / 6.0), aspect = (double)width / height;
    Mat proj{};
    proj.m[0] = fy / aspect;
    proj.m[5] = fy;
    proj.m[10] = n / (fa - n);
    proj.m[11] = -1.0;
    proj.m[14] = n * fa / (fa - n);
    Mat vp = mul(proj, view);
    // Inverse in double (Gauss-Jordan), stored as f32.
    double a[4][8];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            a[r][c] = vp.m[c * 4 + r];
            a[r][4 + c] = r == c ? 1.0 : 0.0;
        }
    for (int c = 0; c < 4; ++c) {
        int piv = c;
        for (int r = c + 1; r < 4; ++r)
            if (std::fabs(a[r][c]) > std::fabs(a[piv][c])) piv = r;
        for (int k = 0; k < 8; ++k) std::swap(a[c][k], a[piv][k]);
        double d = a[c][c];
        for (int k = 0; k < 8; ++k) a[c][k] /= d;
        for (int r = 0; r < 4; ++r)
            if (r != c) {
                double x = a[r][c];
                for (int k = 0; k < 8; ++k) a[r][k] -= x * a[c][k];
            }
    }
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) inverse[c * 4 + r] = (float)a[r][4 + c];
    return RQ_OK;
}
