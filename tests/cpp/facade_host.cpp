// Host-only checks of the SPH::Domain facade (no device calls): particle
// generators, AssignZone and the errors raised for settings the device path
// does not implement.  Prints one "key value..." line per check; the pytest
// side (tests/test_capi_host.py) compares them with numpy and the oracle.
#include <cmath>
#include <cstdio>
#include <functional>
#include <string>

#include "weldform/Domain.h"

static void expect_fatal(const char *name, const std::function<void(SPH::Domain &)> &setup) {
  SPH::Domain dom;
  dom.Dimension = 3;
  dom.auto_ts = false;
  dom.AddBoxLength(0, Vec3_t(0., 0., 0.), 4e-3, 4e-3, 4e-3, 0.5e-3, 2700.0, 1.2e-3, 1, 0, false, false);
  for (size_t a = 0; a < dom.Particles.Size(); ++a) {
    dom.Particles[a]->Fail = 1;
    dom.Particles[a]->Cs = 5000.0;
    dom.Particles[a]->G = 26e9;
  }
  setup(dom);
  try {
    dom.SolveDiffUpdateLeapFrog(1e-6, 1e-7, 1e-6, nullptr, 10);
    std::printf("fatal %s NONE\n", name);
  } catch (Fatal *e) {
    std::printf("fatal %s %s\n", name, e->Msg().c_str());
    delete e;
  }
}

int main() {
  // AddBoxLength cubic and hexagonal packing
  for (int type = 0; type < 2; ++type) {
    SPH::Domain dom;
    dom.Dimension = 3;
    dom.AddBoxLength(1, Vec3_t(0.1, -0.2, 0.3), 3e-3, 2e-3, 5e-3, 0.25e-3, 7850.0, 0.6e-3, type, 0, false, false);
    std::printf("box %d %zu %.17g\n", type, dom.Particles.Size(), dom.Particles[0]->Mass);
    double sx = 0, sy = 0, sz = 0;
    for (size_t a = 0; a < dom.Particles.Size(); ++a) {
      sx += dom.Particles[a]->x(0);
      sy += dom.Particles[a]->x(1);
      sz += dom.Particles[a]->x(2);
    }
    std::printf("boxsum %d %.17g %.17g %.17g\n", type, sx, sy, sz);
  }
  // plane strain: the 2D packings (type, rotation) of AddBoxLength
  const int p2[3][2] = {{0, 0}, {0, 90}, {1, 0}};
  for (const auto &tr : p2) {
    SPH::Domain dom;
    dom.Dimension = 2;
    dom.AddBoxLength(1, Vec3_t(0.1, -0.2, 0.0), 3e-3, 2e-3, 0.0, 0.25e-3, 7850.0, 0.6e-3, tr[0], tr[1], false, false);
    double sx = 0, sy = 0, sz = 0;
    for (size_t a = 0; a < dom.Particles.Size(); ++a) {
      sx += dom.Particles[a]->x(0);
      sy += dom.Particles[a]->x(1);
      sz += std::fabs(dom.Particles[a]->x(2));
    }
    std::printf("box2d %d %d %zu %.17g %.17g %.17g %.17g\n", tr[0], tr[1], dom.Particles.Size(),
                dom.Particles[0]->Mass, sx, sy, sz);
  }
  // AddCylinderLength and AssignZone
  {
    SPH::Domain dom;
    dom.Dimension = 3;
    dom.AddCylinderLength(0, Vec3_t(0., 0., 0.), 7.5e-3, 30e-3, 0.5e-3, 2700.0, 1.2e-3, false);
    Vec3_t lo(-1, -1, -1), hi(1, 1, 0.75e-3);
    int id = 2;
    const int nb = dom.AssignZone(lo, hi, id);
    std::printf("cyl %zu %.17g %d\n", dom.Particles.Size(), dom.Particles[0]->Mass, nb);
  }
  // rigid plane + AddTrimeshParticles (Source/Mesh.cpp:82-220, Source/Contact.cpp:10-37)
  {
    SPH::Domain dom;
    dom.Dimension = 3;
    dom.AddBoxLength(0, Vec3_t(0., 0., 0.), 4e-3, 4e-3, 4e-3, 0.5e-3, 2700.0, 1.2e-3, 1, 0, false, false);
    const size_t ns = dom.Particles.Size();
    SPH::TriMesh *mesh = new SPH::TriMesh;
    mesh->AxisPlaneMesh(2, false, Vec3_t(-1e-3, -1e-3, 5e-3), Vec3_t(5e-3, 5e-3, 5e-3), 6);
    dom.AddTrimeshParticles(mesh, 1.1, 10);
    const SPH::Particle *R = dom.Particles[ns];
    std::printf("mesh %zu %zu %.17g %.17g %.17g %d %d\n", dom.Particles.Size() - ns, mesh->element.Size(), R->h,
                mesh->element[0]->radius, R->normal(2), R->ID, dom.first_fem_particle_idx[0]);
    delete mesh;
  }
  // the contact time-step criterion and contact heat are accepted; contact
  // without a rigid surface is not
  expect_fatal("auto_ts_cont", [](SPH::Domain &d) {
    d.contact = true;
    d.auto_ts_cont = true;
    d.cont_heat_cond = d.cont_heat_fric = true;
  });
  expect_fatal("kernel", [](SPH::Domain &d) { d.Kernel_Set(SPH::Quintic_Spline); });
  expect_fatal("gradient", [](SPH::Domain &d) { d.Gradient_Approach_Set(SPH::Randles_Libersky); });
  // host hooks are accepted (the slow path): without a device the solve
  // fails where every solve does, at wf_create
  expect_fatal("callback", [](SPH::Domain &d) { d.GeneralAfter = [](SPH::Domain &) {}; });
  expect_fatal("contact", [](SPH::Domain &d) { d.contact = true; });
  expect_fatal("fail", [](SPH::Domain &d) { d.Particles[3]->Fail = 0; });
  expect_fatal("mixed_params", [](SPH::Domain &d) { d.Particles[5]->Alpha = 0.5; });
  expect_fatal("hollomon_mat", [](SPH::Domain &d) { d.Particles[0]->Material_model = HOLLOMON; });
  expect_fatal("no_device", [](SPH::Domain &) {});
  try {
    Hollomon bad(Elastic_(200e9, 0.3), 300e6, 3e9, 0.2);
    std::printf("fatal hollomon_def NONE\n");
  } catch (Fatal *e) {
    std::printf("fatal hollomon_def %s\n", e->Msg().c_str());
    delete e;
  }
  return 0;
}
