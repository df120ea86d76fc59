// Cylinder upsetting driven through the SPH::Domain facade, in the form of
// the reference's drivers (Compression.cpp:95-207, Source/WeldForm.cpp:
// 224-369): R = 7.5 mm, L = 30 mm, dx = 1 mm, aluminium (rho 2700, E 68.9 GPa,
// nu 0.3), bilinear hardening Sy = 300 MPa, Et = 1.5 GPa, Monaghan alpha = 1,
// bottom row fixed (zone 2), top row v_z = -1 m/s (zone 3), dt = 0.4 h / Cs.
//
//   upsetting STEPS OUTKEY [leapfrog|kickdrift] [fp64|mixed] [fixed|auto]
//
// "auto" turns on the velocity time-step criterion (Domain::auto_ts) and
// starts from dt = h / Cs, above its limit 0.7 dx / Cs; the end time stays
// that of STEPS fixed steps.
// Writes OUTKEY_reaction.txt (one line "Rx Ry Rz" per step for zone 3),
// OUTKEY_final.txt (one line "ux uy uz sigma_eq pl_strain" per particle,
// %.17g) and the solver's CSV outputs OUTKEY_NNNN.csv.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "weldform/Domain.h"

int main(int argc, char **argv) try {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s STEPS OUTKEY [leapfrog|kickdrift] [fp64|mixed] [fixed|auto]\n", argv[0]);
    return 2;
  }
  const long steps = std::atol(argv[1]);
  const std::string key = argv[2];
  const bool kickdrift = argc > 3 && std::strcmp(argv[3], "kickdrift") == 0;
  const bool mixed = argc > 4 && std::strcmp(argv[4], "mixed") == 0;
  const bool adaptive = argc > 5 && std::strcmp(argv[5], "auto") == 0;

  SPH::Domain dom;
  dom.Dimension = 3;
  dom.Nproc = 1;
  dom.Kernel_Set(SPH::Qubic_Spline);
  dom.Gradient_Approach_Set(SPH::Squared_density);
  dom.auto_ts = adaptive;
  dom.ts_nb_inc = 5;
  dom.precision = mixed ? SPH::Precision::Mixed : SPH::Precision::FP64;

  const double R = 7.5e-3, L = 30e-3, dx = 1.0e-3, rho = 2700.0;
  const double E = 68.9e9, nu = 0.3, Fy = 300e6, Et = 1.5e9;
  const double r = dx / 2.;
  const double h = 2. * r * 1.2;
  const double K = E / (3. * (1. - 2 * nu));
  const double G = E / (2. * (1. + nu));
  const double Cs = std::sqrt(K / rho);
  const double dt = 0.4 * h / Cs;
  const double dt0 = adaptive ? h / Cs : dt;

  // Source/WeldForm.cpp:408-411 leaves DomMax(0) = L, DomMin(0) = -L
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
  Vec3_t lo(-1.0, -1.0, -1.0), hi(1.0, 1.0, zlo + 0.5 * r);
  int bottom = 2, top = 3;
  dom.AssignZone(lo, hi, bottom);
  lo = Vec3_t(-1.0, -1.0, zhi - 0.5 * r);
  hi = Vec3_t(1.0, 1.0, 1.0);
  dom.AssignZone(lo, hi, top);

  SPH::boundaryCondition bc;
  bc.type = SPH::Velocity_BC;
  bc.zoneId = 2;
  bc.value = Vec3_t(0., 0., 0.);
  dom.bConds.push_back(bc);
  bc.zoneId = 3;
  bc.value = Vec3_t(0., 0., -1.0);
  dom.bConds.push_back(bc);

  // the loop runs while Time <= tf: half a step of margin gives exactly STEPS
  const double tf = (steps - 0.5) * dt;
  const double dtOut = (steps / 2) * dt;
  if (kickdrift)
    dom.SolveDiffUpdateKickDrift(tf, dt0, dtOut, key.c_str(), 1000);
  else
    dom.SolveDiffUpdateLeapFrog(tf, dt0, dtOut, key.c_str(), 1000);

  const std::vector<Vec3_t> &hist = dom.ZoneReactionHistory(3);
  FILE *f = std::fopen((key + "_reaction.txt").c_str(), "w");
  for (const Vec3_t &v : hist) std::fprintf(f, "%.17g %.17g %.17g\n", v(0), v(1), v(2));
  std::fclose(f);
  f = std::fopen((key + "_final.txt").c_str(), "w");
  for (size_t a = 0; a < dom.Particles.Size(); ++a) {
    const SPH::Particle *P = dom.Particles[a];
    std::fprintf(f, "%.17g %.17g %.17g %.17g %.17g\n", P->Displacement(0), P->Displacement(1),
                 P->Displacement(2), P->Sigma_eq, P->pl_strain);
  }
  std::fclose(f);
  std::printf("upsetting: %zu particles, %zu steps, time %.17g s, step %.17g s\n", dom.Particles.Size(),
              hist.size(), dom.getTime(), dom.getStepSize());
  delete mat;
  return 0;
} catch (Fatal *e) {
  std::fprintf(stderr, "Fatal: %s\n", e->Msg().c_str());
  delete e;
  return 1;
}
