// Code written against blazeRT's scene API (blazert/scene.h, blazert/ray.h) built on blazert_amd/blazert.hpp.
//   dropin_scene semantics        : add/commit return values (no GPU needed)
//   dropin_scene render W H STEP  : scene_primitives (examples/scene_primitives/main.cc:47-95) on a pixel lattice,
//                                   one intersect1 per pixel; prints "x y label" lines (0 miss, 1 prim 0, 2 other)
#include <blazert_amd/blazert.hpp>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>

#ifndef DROPIN_FT
#define DROPIN_FT double
#endif
using ft = DROPIN_FT;  // the scene type: double, or float with -DDROPIN_FT=float

static int semantics() {
  auto centers = std::make_unique<blazert::Vec3rList<ft>>();
  auto radii = std::make_unique<std::vector<ft>>();
  centers->emplace_back(blazert::Vec3r<ft>{0, 0, 0});
  radii->emplace_back(1.0);
  blazert::Scene<ft> scene;
  const unsigned a = scene.add_spheres(*centers, *radii);
  const unsigned b = scene.add_spheres(*centers, *radii);  // second add of one type
  auto verts = std::make_unique<blazert::Vec3rList<ft>>();
  auto tris = std::make_unique<blazert::Vec3iList>();
  verts->emplace_back(0, 0, 0);
  verts->emplace_back(1, 0, 0);
  verts->emplace_back(0, 1, 0);
  tris->emplace_back(0, 1, 2);
  const unsigned c = scene.add_mesh(*verts, *tris);
  scene.commit();
  const unsigned d = scene.add_planes(*centers, *radii, *radii, blazert::Mat3rList<ft>{blazert::Mat3r<ft>::identity()});
  std::printf("%u %u %u %u %d\n", a, b, c, d, scene.has_been_committed ? 1 : 0);
  return (a == 0 && b == static_cast<unsigned>(-1) && c == 1 && d == static_cast<unsigned>(-1)) ? 0 : 1;
}

static int render(unsigned width, unsigned height, unsigned step) {
  auto centers = std::make_unique<blazert::Vec3rList<ft>>();
  auto semi_axes_a = std::make_unique<std::vector<ft>>();
  auto semi_axes_b = std::make_unique<std::vector<ft>>();
  auto heights = std::make_unique<std::vector<ft>>();
  auto rotations = std::make_unique<blazert::Mat3rList<ft>>();
  auto sph_centers = std::make_unique<blazert::Vec3rList<ft>>();
  auto radii = std::make_unique<std::vector<ft>>();
  blazert::Mat3r<ft> rot{{0, 0, 1}, {0, 1, 0}, {-1, 0, 0}};
  centers->emplace_back(blazert::Vec3r<ft>{-3, 3, 0});
  semi_axes_a->emplace_back(1);
  semi_axes_b->emplace_back(1);
  heights->emplace_back(1);
  rotations->push_back(rot);
  centers->emplace_back(blazert::Vec3r<ft>{1, 4, 0});
  semi_axes_a->emplace_back(0.5);
  semi_axes_b->emplace_back(0.5);
  heights->emplace_back(2);
  rotations->push_back(rot);
  sph_centers->emplace_back(blazert::Vec3r<ft>{1, 1, 0});
  radii->emplace_back(0.5);
  sph_centers->emplace_back(blazert::Vec3r<ft>{-2, 10, 0});
  radii->emplace_back(1.5);
  blazert::Scene<ft> scene;
  scene.add_cylinders(*centers, *semi_axes_a, *semi_axes_b, *heights, *rotations);
  scene.add_spheres(*sph_centers, *radii);
  scene.commit();
  for (unsigned y = 0; y < height; y += step)
    for (unsigned x = 0; x < width; x += step) {
      const blazert::Ray<ft> ray{{0.0, 5.0, 20.0},
                                 {static_cast<ft>((x / ft(width)) - 0.5), static_cast<ft>((y / ft(height)) - 0.5), ft(-1.)}};
      blazert::RayHit<ft> rayhit;
      const bool hit = intersect1(scene, ray, rayhit);
      std::printf("%u %u %d\n", x, y, hit ? (rayhit.prim_id == 0 ? 1 : 2) : 0);
    }
  return 0;
}

int main(int argc, char **argv) {
  if (argc > 1 && std::strcmp(argv[1], "render") == 0)
    return render(argc > 2 ? std::atoi(argv[2]) : 1080, argc > 3 ? std::atoi(argv[3]) : 720,
                  argc > 4 ? std::atoi(argv[4]) : 8);
  return semantics();
}
