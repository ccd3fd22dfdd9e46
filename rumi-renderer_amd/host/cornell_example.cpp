// The Cornell box written the way the reference's Scene::Build writes a scene
// (src/Scene/scene.cpp:135-204), through the drop-in classes of
// rumi_renderer.hpp, then rendered with Scene::Rendering.
//
//   cornell_example --describe            print the scene description (no GPU)
//   cornell_example W H SPP SEED OUT      render, write the RGB floats to OUT
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "rumi_renderer.hpp"

using namespace rumi;

class CornellScene : public Scene {
  public:
    CornellScene(Device* dev, unsigned w, unsigned h, int spp, uint64_t seed) : Scene(dev, spp, seed), _w(w), _h(h) {}

    void Build() override {
        Material* white = Share(new Matte(1.0f, Color(0.73f, 0.73f, 0.73f)));
        Material* red = Share(new Matte(1.0f, Color(0.65f, 0.05f, 0.05f)));
        Material* green = Share(new Matte(1.0f, Color(0.12f, 0.45f, 0.15f)));
        const float x0 = -0.5f, x1 = 0.5f, y0 = 0.f, y1 = 1.f, z0 = -0.5f, z1 = 0.5f;
        auto quad = [&](Point3 a, Point3 b, Point3 c, Point3 d, Material* m, Vector3 n) {
            Triangle* t0 = new Triangle(a, b, c);
            Triangle* t1 = new Triangle(a, c, d);
            t0->Normal(n, n, n);
            t1->Normal(n, n, n);
            Bind(t0, m);
            Bind(t1, m);
            AddObject(t0);
            AddObject(t1);
        };
        quad({x0, y0, z1}, {x1, y0, z1}, {x1, y0, z0}, {x0, y0, z0}, white, {0, 1, 0});   // floor
        quad({x0, y1, z0}, {x1, y1, z0}, {x1, y1, z1}, {x0, y1, z1}, white, {0, -1, 0});  // ceiling
        quad({x0, y0, z0}, {x1, y0, z0}, {x1, y1, z0}, {x0, y1, z0}, white, {0, 0, 1});   // back
        quad({x0, y0, z1}, {x0, y0, z0}, {x0, y1, z0}, {x0, y1, z1}, red, {1, 0, 0});     // left
        quad({x1, y0, z0}, {x1, y0, z1}, {x1, y1, z1}, {x1, y1, z0}, green, {-1, 0, 0});  // right
        const float s = 0.13f;
        Rect* lamp = new Rect({-s / 2, 0.999f, -s / 2}, {s, 0, 0}, {0, 0, s});
        AddObject(lamp);
        AddLight(new AreaLight(lamp, new Emissive(1.0f, Color(17.f, 12.f, 4.f))));
        Sphere* ball = new Sphere({-0.2f, 0.165f, -0.15f}, 0.165f);
        Bind(ball, white);
        AddObject(ball);
        Sphere* glass = new Sphere({0.2f, 0.1f, 0.15f}, 0.1f);
        glass->AddingMaterial(new Dielectric(1.5f, 1.0f, Color(0.f)));
        AddObject(glass);
        _camera = new Pinhole({0.f, 0.5f, 1.9f}, RotationY(3.14159265358979f), (float)_h / (float)_w);
        _camera->_fov = 40.0f;
        _image = new Film(_w, _h);
        _tracer_ptr = new PathTracer(this, 5);
    }

  private:
    unsigned _w, _h;
};

static void describe() {
    CornellScene sc(nullptr, 64, 64, 1, 1);
    sc.Build();
    const rumi_scene_desc& d = sc.Describe();
    std::printf("shapes %u materials %u textures %u\n", d.n_shapes, d.n_materials, d.n_textures);
    for (uint32_t i = 0; i < d.n_shapes; ++i) {
        const rumi_shape& s = d.shapes[i];
        std::printf("shape %d %d %d %d", s.kind, s.material, s.is_light, s.texture);
        for (float v : s.v) std::printf(" %.9g", v);
        for (float v : s.n) std::printf(" %.9g", v);
        std::printf(" %.9g\n", s.radius);
    }
    for (uint32_t i = 0; i < d.n_materials; ++i) {
        const rumi_material& m = d.materials[i];
        std::printf("material %d %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g\n", m.kind, m.kd, m.ks,
                    m.exponent, m.color[0], m.color[1], m.color[2], m.eta_in, m.eta_out);
    }
    const rumi_camera& c = d.camera;
    std::printf("camera %.9g %.9g %.9g", c.position[0], c.position[1], c.position[2]);
    for (float v : c.basis) std::printf(" %.9g", v);
    std::printf(" %.9g %.9g\n", c.fov_deg, c.y_ratio);
}

int main(int argc, char** argv) {
    if (argc == 2 && std::strcmp(argv[1], "--describe") == 0) {
        describe();
        return 0;
    }
    if (argc == 4 && std::strcmp(argv[1], "--model") == 0) {
        // the model scene's transform: RotationX(0) * Translation(0, -0.99, 0) * Scaling(s, s, s)
        const float sc = (float)std::atof(argv[3]);
        const Mat4x4 t = RotationX(0.0f) * Translation(0.0f, -0.99f, 0.0f) * Scaling(sc, sc, sc);
        Model m(argv[2], t);
        std::printf("model %u\n", m.Triangles());
        for (size_t i = 0; i < m._xyz.size(); ++i) {
            uint32_t a, b;
            std::memcpy(&a, &m._xyz[i], 4);
            std::memcpy(&b, &m._normals[i], 4);
            std::printf("%08x %08x\n", a, b);
        }
        return m.Triangles() > 0 ? 0 : 1;
    }
    if (argc != 6) {
        std::fprintf(stderr, "usage: %s --describe | --model OBJ SCALE | W H SPP SEED OUT\n", argv[0]);
        return 2;
    }
    const unsigned w = (unsigned)std::atoi(argv[1]), h = (unsigned)std::atoi(argv[2]);
    const int spp = std::atoi(argv[3]);
    const uint64_t seed = std::strtoull(argv[4], nullptr, 10);
    try {
        Device dev(0);
        CornellScene sc(&dev, w, h, spp, seed);
        sc.Build();
        float seconds = 0.f;
        sc.Rendering(&seconds);
        if (sc.Status() != RUMI_OK) return 1;
        FILE* f = std::fopen(argv[5], "wb");
        if (!f) return 1;
        std::fwrite(sc._image->_rgb.data(), sizeof(float), sc._image->_rgb.size(), f);
        std::fclose(f);
        std::printf("rendered %ux%u at %d spp in %.3f ms\n", w, h, spp, seconds * 1e3);
    } catch (const Error& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
