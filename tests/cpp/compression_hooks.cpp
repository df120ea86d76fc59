// A driver in the form of the reference's Compression.cpp (Compression.cpp:
// 39-76, 133): its boundary conditions are a host function in
// Domain::GeneralAfter, not bConds.  The hook zeroes a and imposes v on the
// bottom row (ID 2, fixed) and on the top row (ID 3, pressed down at
// VMAX t / TAU until TAU, VMAX after), and on the first of its three calls
// per step records the top row's reaction sum(m a) before zeroing it (the
// reaction observer of the reference's drivers).  GeneralBefore counts the
// steps.  Geometry and material as tests/cpp/upsetting.cpp (4,524
// particles, bilinear aluminium, dt = 0.4 h / Cs).
//
//   compression_hooks STEPS OUTKEY [leapfrog|kickdrift]
//
// Writes OUTKEY_reaction.txt (one "Rx Ry Rz" line per step) and
// OUTKEY_final.txt (one "ux uy uz sigma_eq pl_strain" line per particle).
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "weldform/Domain.h"

static double VMAX = 1.0, TAU = 1.0;
static long before_calls = 0, after_calls = 0;
static std::vector<Vec3_t> reaction;

static void CountSteps(SPH::Domain &) { ++before_calls; }

static void UserAcc(SPH::Domain &domi) {
  const double t = domi.getTime();
  const double vcompress = t < TAU ? VMAX / TAU * t : VMAX;
  if (after_calls++ % 3 == 0) {
    Vec3_t R(0., 0., 0.);
    for (size_t i = 0; i < domi.Particles.Size(); ++i)
      if (domi.Particles[i]->ID == 3) R = R + domi.Particles[i]->Mass * domi.Particles[i]->a;
    reaction.push_back(R);
  }
  for (size_t i = 0; i < domi.Particles.Size(); ++i) {
    SPH::Particle *P = domi.Particles[i];
    if (P->ID == 3) {
      P->v = Vec3_t(0.0, 0.0, -vcompress);
      P->a = Vec3_t(0.0, 0.0, 0.0);
    }
    if (P->ID == 2) {
      P->a = Vec3_t(0.0, 0.0, 0.0);
      P->v = Vec3_t(0.0, 0.0, 0.0);
    }
  }
}

int main(int argc, char **argv) try {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s STEPS OUTKEY [leapfrog|kickdrift]\n", argv[0]);
    return 2;
  }
  const long steps = std::atol(argv[1]);
  const std::string key = argv[2];
  const bool kickdrift = argc > 3 && std::strcmp(argv[3], "kickdrift") == 0;

  SPH::Domain dom;
  dom.Dimension = 3;
  dom.Nproc = 1;
  dom.Kernel_Set(SPH::Qubic_Spline);
  dom.Gradient_Approach_Set(SPH::Squared_density);
  dom.auto_ts = false;
  dom.GeneralAfter = &UserAcc;
  dom.GeneralBefore = &CountSteps;

  const double R = 7.5e-3, L = 30e-3, dx = 1.0e-3, rho = 2700.0;
  const double E = 68.9e9, nu = 0.3, Fy = 300e6, Et = 1.5e9;
  const double r = dx / 2.;
  const double h = 2. * r * 1.2;
  const double K = E / (3. * (1. - 2 * nu));
  const double G = E / (2. * (1. + nu));
  const double Cs = std::sqrt(K / rho);
  const double dt = 0.4 * h / Cs;
  TAU = 50 * dt;

  dom.DomMax(0) = L;
  dom.DomMin(0) = -L;
  dom.AddCylinderLength(0, Vec3_t(0., 0., 0.), R, L, r, rho, h, false);
  Elastic_ el(E, nu);
  Material_ *mat = new Material_(el);
  double zlo = 1e300, zhi = -1e300;
  for (size_t a = 0; a < dom.Particles.Size(); ++a) {
    SPH::Particle *P = dom.Particles[a];
    P->G = G;
    P->Cs = Cs;
    P->Material = 0;
    P->Material_model = BILINEAR;
    P->Fail = 1;
    P->Sigmay = Fy;
    P->Ep = E * Et / (E - Et);
    P->Alpha = 1.0;
    P->Beta = 0.0;
    P->TI = 0.0;
    P->mat = mat;
    zlo = std::min(zlo, P->x(2));
    zhi = std::max(zhi, P->x(2));
  }
  // the reference's AssignZone takes its arguments by reference
  Vec3_t lo(-1.0, -1.0, -1.0), hi(1.0, 1.0, zlo + 0.5 * r);
  int bottom = 2, top = 3;
  dom.AssignZone(lo, hi, bottom);
  lo = Vec3_t(-1.0, -1.0, zhi - 0.5 * r);
  hi = Vec3_t(1.0, 1.0, 1.0);
  dom.AssignZone(lo, hi, top);

  const double tf = (steps - 0.5) * dt;
  if (kickdrift)
    dom.SolveDiffUpdateKickDrift(tf, dt, tf, nullptr, 1000);
  else
    dom.SolveDiffUpdateLeapFrog(tf, dt, tf, nullptr, 1000);

  FILE *f = std::fopen((key + "_reaction.txt").c_str(), "w");
  for (const Vec3_t &v : reaction) std::fprintf(f, "%.17g %.17g %.17g\n", v(0), v(1), v(2));
  std::fclose(f);
  f = std::fopen((key + "_final.txt").c_str(), "w");
  for (size_t a = 0; a < dom.Particles.Size(); ++a) {
    const SPH::Particle *P = dom.Particles[a];
    std::fprintf(f, "%.17g %.17g %.17g %.17g %.17g\n", P->Displacement(0), P->Displacement(1),
                 P->Displacement(2), P->Sigma_eq, P->pl_strain);
  }
  std::fclose(f);
  std::printf("compression_hooks: %zu particles, %ld GeneralBefore calls, %ld GeneralAfter calls, time %.17g s\n",
              dom.Particles.Size(), before_calls, after_calls, dom.getTime());
  delete mat;
  return 0;
} catch (Fatal *e) {
  std::fprintf(stderr, "Fatal: %s\n", e->Msg().c_str());
  delete e;
  return 1;
}
