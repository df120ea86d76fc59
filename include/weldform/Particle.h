// weldform/Particle.h -- SPH::Particle with the public members the explicit
// solid step reads and writes (Source/Particle.h:50-265) and the defaults of
// its constructor (Source/Particle.cpp:28-157).  Host-owned: the device keeps
// a structure-of-arrays mirror that SPH::Domain fills from and copies back
// into these objects.
#pragma once
#include "Material.h"
#include "matvec.h"

#define BILINEAR 0
#define HOLLOMON 1
#define JOHNSON_COOK 2
#define TH_BC_NONE 0
#define TH_BC_CONVECTION 1

namespace SPH {

class Particle {
 public:
  Particle(int Tag, Vec3_t const &x0, Vec3_t const &v0, double Mass0, double Density0, double h0,
           bool Fixed = false)
      : x(x0), v(v0), Density(Density0), RefDensity(Density0), Mass(Mass0), h(h0), ID(Tag), IsFree(!Fixed) {}

  // kinematics
  Vec3_t x, v, a, Displacement, VXSPH, normal;
  Vec3_t va, vb;  // Verlet / leapfrog half-step copies of older drivers; unused
  Vec3_t contforce;          // contact force of the last step (rigid contact)
  int element = 0, mesh = 0; // rigid-surface particles: element and mesh index
  // state
  double Density, RefDensity, Mass, h;
  double dDensity = 0.0, Pressure = 0.0, P0 = 0.0, Cs = 0.0;
  size_t PresEq = 0;
  Mat3_t Sigma, ShearStress, StrainRate, RotationRate, Strain, Strain_pl, Strain_pl_incr, TIR;
  double pl_strain = 0.0, delta_pl_strain = 0.0, Sigmay = 0.0, eff_strain_rate = 0.0, Sigma_eq = 0.0;
  double dam_D = 0.0;  // damage (Domain::model_damage)
  double G = 0.0, K = 0.0, Ep = 0.0, Et = 0.0;
  double Alpha = 0.0, Beta = 0.0, TI = 0.0, TIn = 4.0, TIInitDist = 0.0;
  int ID, Material = 0, Material_model = BILINEAR;
  bool IsFree;
  size_t Fail = 0;
  Material_ *mat = nullptr;
  // thermal
  double T = 0.0, k_T = 0.0, cp_T = 0.0, dTdt = 0.0, th_exp = 0.0;
  double q_source = 0.0, q_conv = 0.0, T_inf = 0.0, h_conv = 0.0, q_plheat = 0.0, ps_energy = 0.0;
  int Thermal_BC = TH_BC_NONE;
  int Nb = 0;
  int ContNb = 0;             // contact pairs of the last contact search
  Mat3_t gradCorrM;           // kernel-gradient correction (not formed here: 0)
  double friction_hfl = 0.0, cshearabs = 0.0;  // friction heating outputs (0 here)
  bool print_history = false;
  bool Shepard = false;  // density filter flag; unused on the device path

  // Particle::CalculateEquivalentStress (Source/Particle.cpp:623-633)
  void CalculateEquivalentStress() {
    const Mat3_t &S = ShearStress;
    const double J2 = 0.5 * (S(0, 0) * S(0, 0) + 2.0 * S(0, 1) * S(1, 0) + 2.0 * S(0, 2) * S(2, 0) +
                             S(1, 1) * S(1, 1) + 2.0 * S(1, 2) * S(2, 1) + S(2, 2) * S(2, 2));
    Sigma_eq = std::sqrt(3.0 * J2);
  }
};

}  // namespace SPH
