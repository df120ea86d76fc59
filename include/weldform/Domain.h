// weldform/Domain.h -- drop-in SPH::Domain for WeldForm drivers, running the
// explicit solid step on an MI355X through the C-ABI of weldform_hip.h.
//
// Kept from the reference (Source/Domain.h:113-537): class and method names,
// argument meanings and the public members drivers touch -- Dimension,
// Nproc, Particles, Kernel_Set, Gradient_Approach_Set, AddCylinderLength,
// AddBoxLength, AssignZone, DomMax/DomMin, bConds, amps, auto_ts,
// ts_nb_inc, nonlock_sum, gradKernelCorr, thermal_solver, pl_heating,
// pl_work_heat_frac, energy sums, getTime(), getStepSize(), WriteCSV and the
// solver entry points SolveDiffUpdateLeapFrog / SolveDiffUpdateLeapfrog /
// SolveDiffUpdateKickDrift (tf, dt, dtOut, FileKey, maxidx).
//
// Boundary conditions: the velocity and temperature BCs in bConds (with
// amps) are applied on the device where the reference calls GeneralAfter =
// UserAcc (Source/WeldForm.cpp:51-120), inside captured device steps.  A
// driver's own host function in GeneralBefore / GeneralAfter (e.g.
// Compression.cpp:39-76,133) runs on a slow path: each step is cut at the
// reference's four call sites (wf_step_part), and around each call the
// fields a hook reads (x, v, a, Displacement, Density, T, ID, and the
// Domain's time and sums) are copied into the Particle objects and the ones
// it writes (v, a, and T with the thermal solver) copied back.  Errors are
// thrown as Fatal*, as in the reference.
//
// Link with -lweldform_hip -lweldform_io (weldform_amd/lib); compiles with
// any C++17 compiler, no HIP headers needed.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../weldform_hip.h"
#include "../weldform_io.h"
#include "Material.h"
#include "Mesh.h"
#include "Particle.h"
#include "matvec.h"

namespace SPH {

enum Kernels_Type { Qubic_Spline = 0, Quintic = 1, Quintic_Spline = 2, Hyperbolic_Spline = 3 };
enum Gradient_Type { Squared_density = 0, Multiplied_density = 1, Randles_Libersky = 2 };
enum BC_TYPE { Velocity_BC = 0, Force_BC, Temperature_BC, Convection_BC, Symmetry_BC };
// Source/Domain.h:61-64
enum Function_Type { Constant = 0, Linear = 1, Multilinear = 2 };
enum Friction_Type { Fr_Sta = 0, Fr_Dyn, Fr_StaDyn, Fr_Bound };
enum Contact_Alg { Fraser = 0, Wang, Seo, Zhan, LSDyna };
enum class Precision { FP64 = WF_FP64, Mixed = WF_MIXED };

// amplitude table (Source/Domain.h:69-89)
struct amplitude {
  int id = 0;
  std::vector<double> time, value;
};

// boundary condition (Source/Domain.h:98-111)
struct boundaryCondition {
  int zoneId = 0;
  int type = Velocity_BC;
  bool free = true;
  int valueType = 0;  // 0 constant, 1 amplitude
  Vec3_t value;
  Vec3_t value_ang;
  int ampId = 0;
  double ampFactor = 1.0;
  double cv_coeff = 0.0, T_inf = 0.0, T = 0.0;
};

class Domain;
inline void General(Domain &) {}

class Domain {
 public:
  typedef void (*PtDom)(Domain &dom);

  Domain() {
    DomMax = Vec3_t(-100000000000.0);
    DomMin = Vec3_t(100000000000.0);
  }
  ~Domain() {
    close();
    for (size_t i = 0; i < Particles.Size(); ++i) delete Particles[i];
  }
  Domain(const Domain &) = delete;
  Domain &operator=(const Domain &) = delete;

  // ------------------------------------------------------------ settings
  int Dimension = 2;                 // reference default; drivers set 3
  size_t Nproc = 1;                  // accepted, unused (one GPU)
  Array<Particle *> Particles;
  Vec3_t DomMax, DomMin;
  PtDom GeneralBefore = &General;
  PtDom GeneralAfter = &General;
  std::vector<boundaryCondition> bConds;
  std::vector<amplitude> amps;
  bool auto_ts = true, auto_ts_acc = false, auto_ts_cont = false;
  int ts_nb_inc = 5;
  bool nonlock_sum = true;           // false: the Locking sums (float m/rho weights in the rates)
  bool model_damage = false;         // damage of the materials' DamageModel (Johnson-Cook)
  bool gradKernelCorr = false;       // Bonet kernel-gradient correction
  bool thermal_solver = false;
  bool pl_heating = false;
  bool contact = false;
  // rigid contact (Source/Domain.h:260-330): planes, their first particle
  // index and zone ids, Wang contact parameters and outputs
  std::vector<TriMesh *> trimesh;
  std::vector<int> first_fem_particle_idx, contact_surf_id;
  int meshcount = 0;
  // PFAC: the reference's constructor leaves it unset; 0.5 is its JSON
  // driver's default (Source/WeldForm.cpp:685)
  double friction_sta = 0.0, friction_dyn = 0.0, DFAC = 0.0, PFAC = 0.5;
  // friction_function Linear: friction_m T + friction_b (Wang's and
  // LS-Dyna's laws); fric_type Fr_Bound (Fraser's law)
  Function_Type friction_function = Constant;
  double friction_m = 0.0, friction_b = 0.0;
  Friction_Type fric_type = Fr_Dyn;
  bool cont_heat_cond = false, cont_heat_fric = false;
  double contact_hc = 0.0;  // contact heat conductance
  // min time step of the contact forces (auto_ts_cont) and the contact heat
  // totals (Source/Domain.h:327, 346-347; Source/Thermal.cpp:194-210)
  double min_force_ts = 0.0, contact_friction_work = 0.0, accum_cont_heat_cond = 0.0;
  Contact_Alg contact_alg = Wang;  // Source/Domain.cpp:169
  std::vector<Vec3_t> m_contact_force;
  double contact_force_sum = 0.0, contact_reaction_sum = 0.0, ext_forces_work = 0.0;
  double pl_work_heat_frac = 0.9;
  int solid_part_count = -1;
  double kin_energy_sum = 0.0, int_energy_sum = 0.0, plastic_work = 0.0;
  double m_maxT = 0.0, m_minT = 0.0;
  // Domain::max_disp (Source/Domain.h:470): the Leapfrog and KickDrift loops
  // compute a local of the same name (Source/SolverLeapfrog.cpp:140-148), so
  // the member a hook can read stays (0, 0, 0) on these paths
  Vec3_t max_disp = Vec3_t(0., 0., 0.);
  // device-path extensions
  Precision precision = Precision::FP64;
  int device = 0;
  int max_nb = 128;
  size_t Scheme = 0;   // accepted, unused: the solver entry point picks the scheme
  // adaptive time step (Source/Domain.cpp:109,153,170): velocity criterion
  // CFL, acceleration criterion sqrt_h_a, and the floor manual_min_ts, an int
  // in the reference (Source/Domain.h:328), so its 1.e-10 default stores 0
  double CFL = 0.7;
  double sqrt_h_a = 0.0025;
  int manual_min_ts = 0;

  void Kernel_Set(Kernels_Type const &KT) {
    KernelType = KT;
    Cellfac = (KT == Quintic_Spline) ? 3.0 : 2.0;
  }
  void Gradient_Approach_Set(Gradient_Type const &GT) { GradientType = GT; }
  double &getTime() { return Time; }
  const double &getStepSize() const { return deltat; }
  double getCellfac() const { return Cellfac; }

  // ------------------------------------------------------------ geometry
  // AddCylinderLength (Source/Domain.cpp:708-860): cubic lattice, mass
  // pi R^2 L rho / N over the particles in the domain
  void AddCylinderLength(int tag, Vec3_t const &V, double Rxy, double Lz, double r, double Density, double h,
                         bool Fixed, bool ghost = false) {
    if (Dimension != 3) throw new Fatal("AddCylinderLength: only the 3D generator is available");
    if (ghost) throw new Fatal("AddCylinderLength: ghost symmetry rows are not supported on the device path");
    const double Vv[3] = {V(0), V(1), V(2)};
    const int64_t n = wf_gen_cylinder(Vv, Rxy, Lz, r, nullptr, 0);
    if (n <= 0) throw new Fatal("AddCylinderLength: no particles generated");
    std::vector<double> xs(3 * (size_t)n);
    wf_gen_cylinder(Vv, Rxy, Lz, r, xs.data(), n);
    for (int64_t i = 0; i < n; ++i)
      Particles.Push(new Particle(tag, Vec3_t(xs[3 * i], xs[3 * i + 1], xs[3 * i + 2]), Vec3_t(0, 0, 0), 0.0,
                                  Density, h, Fixed));
    solid_part_count = (int)Particles.Size();
    const double Vol = M_PI * Rxy * Rxy * Lz;
    const double Mass = Vol * Density / Particles.Size();
    for (size_t i = 0; i < Particles.Size(); ++i) Particles[i]->Mass = Mass;
    R = r;
  }

  // AddBoxLength, 3D cubic (type 1) and hexagonal close packing (type 0),
  // Source/Domain.cpp:326-460; mass = box volume * rho / N.  In 2D the
  // plane z = 0 (Source/Domain.cpp:494-571): hexagonal rows (rotation 0) or
  // columns (rotation 90) with mass sqrt(3) r^2 rho each, or a square
  // lattice with mass 4 r^2 rho
  void AddBoxLength(int tag, Vec3_t const &V, double Lx, double Ly, double Lz, double r, double Density, double h,
                    int type, int rotation, bool random, bool Fixed) {
    if (Dimension != 3 && Dimension != 2) throw new Fatal("AddBoxLength: Dimension must be 2 or 3");
    if (!(type == 0 || type == 1) || !(rotation == 0 || rotation == 90))
      throw new Fatal("AddBoxLength: packing type must be 0 or 1 and rotation 0 or 90");
    if (random) throw new Fatal("AddBoxLength: randomised packing is not supported on the device path");
    if (Dimension == 2) {
      add_box_2d(tag, V, Lx, Ly, r, Density, h, type, rotation, Fixed);
      return;
    }
    const size_t pre = Particles.Size();
    size_t k = 0;
    double zp = V(2);
    while (zp <= V(2) + Lz - r) {
      size_t j = 0;
      double yp = V(1);
      while (yp <= V(1) + Ly - r) {
        size_t i = 0;
        double xp = V(0);
        while (xp <= V(0) + Lx - r) {
          double x, y, z;
          if (type == 0) {
            const bool odd = (k % 2 != 0) && (j % 2 != 0);
            x = V(0) + (2 * i + (j % 2) + (k % 2) + (odd ? -1 : 1)) * r;
            y = V(1) + (std::sqrt(3.0) * (j + (1.0 / 3.0) * (k % 2)) + 1) * r;
            z = V(2) + ((2 * std::sqrt(6.0) / 3) * k + 1) * r;
          } else {
            x = V(0) + (2.0 * i + 1) * r;
            y = V(1) + (2.0 * j + 1) * r;
            z = V(2) + (2.0 * k + 1) * r;
          }
          Particles.Push(new Particle(tag, Vec3_t(x, y, z), Vec3_t(0, 0, 0), 0.0, Density, h, Fixed));
          i++;
          if (type == 0) {
            const bool odd = (k % 2 != 0) && (j % 2 != 0);
            xp = V(0) + (2 * i + (j % 2) + (k % 2) + (odd ? -1 : 1)) * r;
          } else {
            xp = V(0) + (2 * i + 1) * r;
          }
        }
        j++;
        yp = (type == 0) ? V(1) + (std::sqrt(3.0) * (j + (1.0 / 3.0) * (k % 2)) + 1) * r : V(1) + (2.0 * j + 1) * r;
      }
      k++;
      zp = (type == 0) ? V(2) + ((2 * std::sqrt(6.0) / 3) * k + 1) * r : V(2) + (2.0 * k + 1) * r;
    }
    Vec3_t Max = V;
    for (size_t i = pre; i < Particles.Size(); ++i)
      for (int q = 0; q < 3; ++q)
        if (Particles[i]->x(q) > Max(q)) Max(q) = Particles[i]->x(q);
    Max += Vec3_t(r);
    const Vec3_t ext = Max - V;
    const double Mass = ext(0) * ext(1) * ext(2) * Density / (double)(Particles.Size() - pre);
    for (size_t i = pre; i < Particles.Size(); ++i) Particles[i]->Mass = Mass;
    solid_part_count = (int)Particles.Size();
    R = r;
  }

  // the 2D branches of AddBoxLength (Source/Domain.cpp:494-571)
  void add_box_2d(int tag, Vec3_t const &V, double Lx, double Ly, double r, double Density, double h, int type,
                  int rotation, bool Fixed) {
    const double s3 = std::sqrt(3.0);
    auto push = [&](double x, double y, double m) {
      Particles.Push(new Particle(tag, Vec3_t(x, y, 0.0), Vec3_t(0, 0, 0), m, Density, h, Fixed));
    };
    if (type == 1) {
      // square lattice of pitch 2r
      for (size_t j = 0;; ++j) {
        const double yp = j == 0 ? V(1) : V(1) + (2.0 * j + 1) * r;
        if (!(yp <= V(1) + Ly - r)) break;
        for (size_t i = 0;; ++i) {
          const double xp = i == 0 ? V(0) : V(0) + (2.0 * i + 1) * r;
          if (!(xp <= V(0) + Lx - r)) break;
          push(V(0) + (2.0 * i + 1) * r, V(1) + (2.0 * j + 1) * r, 2.0 * r * 2.0 * r * Density);
        }
      }
    } else if (rotation == 0) {
      // hexagonal rows: odd rows shifted by r
      for (size_t j = 0;; ++j) {
        const double yp = j == 0 ? V(1) : V(1) + (s3 * j + 1) * r;
        if (!(yp <= V(1) + Ly - r)) break;
        for (size_t i = 0;; ++i) {
          const double xp = i == 0 ? V(0) : V(0) + (2 * i + (j % 2) + 1) * r;
          if (!(xp <= V(0) + Lx - r)) break;
          push(V(0) + (2 * i + (j % 2) + 1) * r, V(1) + (s3 * j + 1) * r, s3 * r * r * Density);
        }
      }
    } else {
      // hexagonal columns: odd columns shifted by r
      for (size_t i = 0;; ++i) {
        const double xp = i == 0 ? V(0) : V(0) + (s3 * i + 1) * r;
        if (!(xp <= V(0) + Lx - r)) break;
        for (size_t j = 0;; ++j) {
          const double yp = j == 0 ? V(1) : V(1) + (2 * j + (i % 2) + 1) * r;
          if (!(yp <= V(1) + Ly - r)) break;
          push(V(0) + (s3 * i + 1) * r, V(1) + (2 * j + (i % 2) + 1) * r, s3 * r * r * Density);
        }
      }
    }
    solid_part_count = (int)Particles.Size();
    R = r;
  }

  // AssignZone (Source/Domain.cpp:691-706)
  int AssignZone(Vec3_t &start, Vec3_t &end, int &id) {
    int cnt = 0;
    for (size_t a = 0; a < Particles.Size(); ++a) {
      bool in = true;
      for (int i = 0; i < 3; ++i)
        if (Particles[a]->x(i) < start[i] || Particles[a]->x(i) > end[i]) in = false;
      if (in) {
        Particles[a]->ID = id;
        cnt++;
      }
    }
    return cnt;
  }

  // ------------------------------------------------------------ solvers
  void SolveDiffUpdateLeapFrog(double tf, double dt, double dtOut, char const *TheFileKey, size_t maxidx) {
    solve(tf, dt, dtOut, TheFileKey, maxidx, WF_LEAPFROG);
  }
  // declared (lower-case f) by Source/Domain.h:198 and called by older drivers
  void SolveDiffUpdateLeapfrog(double tf, double dt, double dtOut, char const *TheFileKey, size_t maxidx) {
    solve(tf, dt, dtOut, TheFileKey, maxidx, WF_LEAPFROG);
  }
  void SolveDiffUpdateKickDrift(double tf, double dt, double dtOut, char const *TheFileKey, size_t maxidx) {
    solve(tf, dt, dtOut, TheFileKey, maxidx, WF_KICKDRIFT);
  }

  // Domain::AddTrimeshParticles (Source/Contact.cpp:10-37): one particle per
  // element at its centroid, h = hfac * element radius, zone id `id`
  void AddTrimeshParticles(TriMesh *mesh, const float &hfac, const int &id) {
    if (!mesh || mesh->element.Size() == 0) throw new Fatal("AddTrimeshParticles: empty mesh");
    first_fem_particle_idx.push_back((int)Particles.Size());
    contact_surf_id.push_back(id);
    trimesh.push_back(mesh);
    mesh->id = id;
    mesh_hfac_.push_back(hfac);
    for (size_t e = 0; e < mesh->element.Size(); ++e) {
      Particle *P = new Particle(id, mesh->element[e]->centroid, Vec3_t(0, 0, 0), 0.0, 0.0,
                                 (double)hfac * mesh->element[e]->radius, false);
      P->normal = mesh->element[e]->normal;
      P->element = (int)e;
      P->mesh = meshcount;
      Particles.Push(P);
    }
    m_contact_force.push_back(Vec3_t());
    meshcount++;
    contact = true;
  }

  // contact force on plane m (Domain::m_contact_force[m]) at every step of the solve
  const std::vector<Vec3_t> &ContactForceHistory(int m) { return cforce_hist_[m]; }

  // reaction sum(m a) of a velocity-BC zone at the last step (taken before the
  // BC zeroes a), and its history over the solve
  Vec3_t ZoneReaction(int zone) const {
    auto it = reaction_.find(zone);
    if (it == reaction_.end() || it->second.empty()) return Vec3_t();
    return it->second.back();
  }
  const std::vector<Vec3_t> &ZoneReactionHistory(int zone) { return reaction_[zone]; }

  // Domain::OutputName (Source/Domain.h:449): names of the three UserOutput
  // datasets (their values are the default UserOutput's 0)
  std::string OutputName[3] = {"Property1", "Property2", "Property3"};

  // WriteXDMF (Source/Output.cpp:75-518): FileKey.hdf5 + FileKey.xmf of every
  // particle (solid, then rigid-surface), from the Particle objects
  void WriteXDMF(char const *FileKey) {
    const size_t N = Particles.Size();
    std::vector<float> f((size_t)WF_XDMF_NF * N);
    std::vector<int32_t> iv((size_t)WF_XDMF_NI * N);
    host_xdmf(0, N, N, f.data(), iv.data());
    write_xdmf(FileKey, N, f, iv);
  }

  // WriteCSV (Source/Output.cpp:521-554), from the Particle objects
  void WriteCSV(char const *FileKey) {
    std::vector<double> rows((size_t)WF_CSV_NCOL * Particles.Size());
    host_csv(0, Particles.Size(), rows.data());
    write_csv(FileKey, Particles.Size(), rows);
  }

  // the same two files of the device state during a solve: the solid
  // particles' tables are packed on the device (wf_pack_xdmf / wf_pack_csv),
  // the rigid-surface particles' rows come from their Particle objects
  // after their positions and velocities are pulled
  void WriteDeviceOutput(char const *FileKey, bool xdmf, bool csv) {
    if (!ctx_) throw new Fatal("WriteDeviceOutput: no device state (call inside or after a solve)");
    const size_t n = nsolid(), N = Particles.Size();
    pull_rigid();
    if (xdmf) {
      std::vector<float> fd((size_t)WF_XDMF_NF * n), f((size_t)WF_XDMF_NF * N);
      std::vector<int32_t> ivd((size_t)WF_XDMF_NI * n), iv((size_t)WF_XDMF_NI * N);
      check(wf_pack_xdmf(ctx_, fd.data(), ivd.data()));
      for (int d = 0; d < WF_XDMF_NF;) {
        const int w = xdmf_width(d);
        std::memcpy(&f[(size_t)d * N], &fd[(size_t)d * n], sizeof(float) * w * n);
        d += w;
      }
      for (int d = 0; d < WF_XDMF_NI; ++d) std::memcpy(&iv[(size_t)d * N], &ivd[(size_t)d * n], sizeof(int32_t) * n);
      host_xdmf(n, N, N, f.data(), iv.data());
      write_xdmf(FileKey, N, f, iv);
    }
    if (csv) {
      std::vector<double> rows((size_t)WF_CSV_NCOL * N);
      check(wf_pack_csv(ctx_, rows.data()));
      host_csv(n, N, rows.data() + (size_t)WF_CSV_NCOL * n);
      write_csv(FileKey, N, rows);
    }
  }

 private:
  double Time = 0.0, deltat = 0.0, Cellfac = 2.0, R = 0.0;
  int KernelType = 0, GradientType = 0;
  void *ctx_ = nullptr;
  std::map<int, std::vector<Vec3_t>> reaction_;
  std::map<int, std::vector<Vec3_t>> cforce_hist_;
  std::vector<float> mesh_hfac_;

  // solid particles come first; rigid-surface particles follow them
  size_t nsolid() const { return contact && !first_fem_particle_idx.empty() ? (size_t)first_fem_particle_idx[0] : Particles.Size(); }
  // steps per wf_step call: the device keeps the last 65536 reaction entries
  static constexpr int64_t kMaxChunk = 1 << 16;

  void close() {
    if (ctx_) wf_destroy(ctx_);
    ctx_ = nullptr;
  }
  void check(int rc) {
    if (rc != WF_OK) throw new Fatal(std::string("weldform_hip: ") + wf_last_error(ctx_));
  }

  void check_supported() {
    const size_t n = Particles.Size();
    if (n == 0) throw new Fatal("Particle Count is Null. Please Check Radius and Domain Dimensions.");
    if (Dimension != 3 && Dimension != 2) throw new Fatal("Dimension must be 2 (plane strain) or 3");
    if (KernelType != Qubic_Spline) throw new Fatal("the device path implements the cubic spline kernel only");
    if (GradientType != Squared_density) throw new Fatal("the device path implements GradientType 0 only");
    if (contact) {
      if (trimesh.empty()) throw new Fatal("contact is set but no rigid mesh was added (AddTrimeshParticles)");
      // SolveDiffUpdateLeapFrog dispatches Fraser, Wang, Seo and LSDyna
      // (Source/SolverLeapfrog.cpp:234-239); Zhan's law is never called
      if (contact_alg == Zhan) throw new Fatal("contact_alg Zhan is not dispatched by the solvers");
      if (friction_function == Multilinear) throw new Fatal("friction_function Multilinear is not implemented");
      // a mesh's m_w comes from the constant velocity BC of its zone (UserAcc
      // sets it every call); one set directly on a mesh without such a BC is
      // not carried to the device
      for (TriMesh *m : trimesh) {
        if (norm(m->m_w) == 0.0) continue;
        bool bc = false;
        for (const auto &b : bConds) bc = bc || (b.type == Velocity_BC && b.valueType == 0 && b.zoneId == m->id);
        if (!bc) throw new Fatal("a rotating contact mesh needs a constant velocity BC on its zone (valueAng)");
      }
      if (trimesh.size() > 4) throw new Fatal("at most 4 contact meshes");
    }
  }

  bool host_hooks() const { return GeneralBefore != &General || GeneralAfter != &General; }

  // the slow path around a host hook: what the hook may read, from the
  // device into the Particle objects ...
  void hook_pull() {
    const size_t n = nsolid();
    std::vector<double> v3(3 * n), v1(n);
    auto get3 = [&](int f, Vec3_t Particle::*m) {
      check(wf_get_field(ctx_, f, v3.data()));
      for (size_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) (Particles[i]->*m)(k) = v3[3 * i + k];
    };
    auto get1 = [&](int f, double Particle::*m) {
      check(wf_get_field(ctx_, f, v1.data()));
      for (size_t i = 0; i < n; ++i) Particles[i]->*m = v1[i];
    };
    get3(WF_F_X, &Particle::x);
    get3(WF_F_V, &Particle::v);
    get3(WF_F_A, &Particle::a);
    get3(WF_F_DISP, &Particle::Displacement);
    get1(WF_F_DENSITY, &Particle::Density);
    get1(WF_F_T, &Particle::T);
    std::vector<int32_t> id(n);
    check(wf_get_int_field(ctx_, WF_I_ID, id.data()));
    for (size_t i = 0; i < n; ++i) Particles[i]->ID = id[i];
    pull_scalars();
  }
  // ... and what it may write, back to the device
  void hook_push() {
    const size_t n = nsolid();
    std::vector<double> v3(3 * n), v1(n);
    auto put3 = [&](int f, Vec3_t Particle::*m) {
      for (size_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) v3[3 * i + k] = (Particles[i]->*m)(k);
      check(wf_set_field(ctx_, f, v3.data()));
    };
    put3(WF_F_V, &Particle::v);
    put3(WF_F_A, &Particle::a);
    if (thermal_solver) {
      for (size_t i = 0; i < n; ++i) v1[i] = Particles[i]->T;
      check(wf_set_field(ctx_, WF_F_T, v1.data()));
    }
  }
  void run_hook(PtDom f) {
    if (f == &General) return;
    hook_pull();
    f(*this);
    hook_push();
  }
  // one step with the host hooks at the reference's call sites
  // (Source/SolverLeapfrog.cpp:215,228,254,273; Source/SolverKickDrift.cpp:
  // 196,208,231,267)
  void step_with_hooks(int scheme) {
    check(wf_step_part(ctx_, scheme, WF_PART_SEARCH));
    run_hook(GeneralBefore);
    check(wf_step_part(ctx_, scheme, WF_PART_ACCEL));
    run_hook(GeneralAfter);
    check(wf_step_part(ctx_, scheme, WF_PART_KICK1));
    run_hook(GeneralAfter);
    check(wf_step_part(ctx_, scheme, WF_PART_KICK2));
    run_hook(GeneralAfter);
    check(wf_step_part(ctx_, scheme, WF_PART_REST));
  }

  // one wf_material per Particle::Material id; particles sharing an id must
  // share their parameters, since pairs form within one id
  wf_material material_of(const Particle *P) {
    wf_material m;
    std::memset(&m, 0, sizeof m);
    m.model = P->Material_model;
    m.G = P->G;
    m.Cs = P->Cs;
    m.ref_density = P->RefDensity;
    m.P0 = P->P0;
    m.alpha = P->Alpha;
    m.beta = P->Beta;
    m.Ep = P->Ep;
    m.k_T = P->k_T;
    m.cp_T = P->cp_T;
    m.th_exp = P->th_exp;
    if (P->TI > 0.0) {
      // tensile instability: TI, TIn and TIInitDist travel with the material
      m.ti = P->TI;
      m.ti_n = P->TIn;
      m.ti_dist = P->TIInitDist;
    }
    if (P->mat) {
      m.E = P->mat->Elastic().E();
      m.nu = P->mat->Elastic().nu();
    }
    if (P->Material_model == HOLLOMON) {
      const Hollomon *H = dynamic_cast<const Hollomon *>(P->mat);
      if (!H) throw new Fatal("Hollomon particle without a Hollomon material");
      m.hol_K = H->K;
      m.hol_m = H->m;
      m.hol_sy0 = H->sy0;
    } else if (P->Material_model == JOHNSON_COOK) {
      JohnsonCook *J = dynamic_cast<JohnsonCook *>(P->mat);
      if (!J) throw new Fatal("Johnson-Cook particle without a JohnsonCook material");
      m.jc_A = J->A;
      m.jc_B = J->B;
      m.jc_n = J->n;
      m.jc_C = J->C;
      m.jc_eps0 = J->eps_0;
      m.jc_m = J->m;
      m.jc_Tm = J->getT_m();
      m.jc_Tt = J->getT_t();
    } else if (P->Material_model != BILINEAR) {
      throw new Fatal("material model not implemented on the device path");
    }
    if (P->Fail != 1) throw new Fatal("only the Fail = 1 (plastic) stress update is implemented");
    return m;
  }

  std::vector<int> build_materials(std::vector<wf_material> &table) {
    const size_t n = nsolid();
    std::map<int, int> idx;
    table.clear();
    std::vector<int> mi(n);
    for (size_t i = 0; i < n; ++i) {
      const Particle *P = Particles[i];
      const wf_material m = material_of(P);
      auto it = idx.find(P->Material);
      if (it == idx.end()) {
        if (table.size() == 256) throw new Fatal("more than 256 material ids");
        it = idx.emplace(P->Material, (int)table.size()).first;
        table.push_back(m);
      } else if (std::memcmp(&table[it->second], &m, sizeof m) != 0) {
        throw new Fatal("particles of one Material id carry different parameters");
      }
      mi[i] = it->second;
    }
    return mi;
  }

  void upload() {
    const size_t n = nsolid();
    std::vector<double> v3(3 * n), v1(n), v9(9 * n);
    auto put3 = [&](int f, Vec3_t Particle::*m) {
      for (size_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) v3[3 * i + k] = (Particles[i]->*m)(k);
      check(wf_set_field(ctx_, f, v3.data()));
    };
    auto put1 = [&](int f, double Particle::*m) {
      for (size_t i = 0; i < n; ++i) v1[i] = Particles[i]->*m;
      check(wf_set_field(ctx_, f, v1.data()));
    };
    auto put9 = [&](int f, Mat3_t Particle::*m) {
      for (size_t i = 0; i < n; ++i)
        for (int k = 0; k < 9; ++k) v9[9 * i + k] = (Particles[i]->*m).m[k];
      check(wf_set_field(ctx_, f, v9.data()));
    };
    put3(WF_F_X, &Particle::x);
    put3(WF_F_V, &Particle::v);
    put3(WF_F_A, &Particle::a);
    put3(WF_F_DISP, &Particle::Displacement);
    put1(WF_F_DENSITY, &Particle::Density);
    put1(WF_F_MASS, &Particle::Mass);
    put1(WF_F_H, &Particle::h);
    put9(WF_F_SIGMA, &Particle::Sigma);
    put9(WF_F_SHEAR, &Particle::ShearStress);
    put9(WF_F_STRAIN, &Particle::Strain);
    put9(WF_F_STRAIN_PL, &Particle::Strain_pl);
    put9(WF_F_STRAIN_PL_INCR, &Particle::Strain_pl_incr);
    put1(WF_F_PL_STRAIN, &Particle::pl_strain);
    if (model_damage) put1(WF_F_DAMAGE, &Particle::dam_D);
    put1(WF_F_DPL_STRAIN, &Particle::delta_pl_strain);
    put1(WF_F_SIGMAY, &Particle::Sigmay);
    put1(WF_F_T, &Particle::T);
    put1(WF_F_DTDT, &Particle::dTdt);
    put1(WF_F_Q_SOURCE, &Particle::q_source);
    put1(WF_F_H_CONV, &Particle::h_conv);
    put1(WF_F_T_INF, &Particle::T_inf);
    put1(WF_F_Q_PLHEAT, &Particle::q_plheat);
    put1(WF_F_PS_ENERGY, &Particle::ps_energy);
    std::vector<int32_t> iv(n);
    for (size_t i = 0; i < n; ++i) iv[i] = Particles[i]->ID;
    check(wf_set_int_field(ctx_, WF_I_ID, iv.data()));
    for (size_t i = 0; i < n; ++i) iv[i] = Particles[i]->IsFree ? 1 : 0;
    check(wf_set_int_field(ctx_, WF_I_FREE, iv.data()));
    for (size_t i = 0; i < n; ++i) iv[i] = Particles[i]->Thermal_BC;
    check(wf_set_int_field(ctx_, WF_I_THERMAL_BC, iv.data()));
  }

  void download() {
    const size_t n = nsolid();
    std::vector<double> v3(3 * n), v1(n), v9(9 * n);
    auto get3 = [&](int f, Vec3_t Particle::*m) {
      check(wf_get_field(ctx_, f, v3.data()));
      for (size_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) (Particles[i]->*m)(k) = v3[3 * i + k];
    };
    auto get1 = [&](int f, double Particle::*m) {
      check(wf_get_field(ctx_, f, v1.data()));
      for (size_t i = 0; i < n; ++i) Particles[i]->*m = v1[i];
    };
    auto get9 = [&](int f, Mat3_t Particle::*m) {
      check(wf_get_field(ctx_, f, v9.data()));
      for (size_t i = 0; i < n; ++i)
        for (int k = 0; k < 9; ++k) (Particles[i]->*m).m[k] = v9[9 * i + k];
    };
    get3(WF_F_X, &Particle::x);
    get3(WF_F_V, &Particle::v);
    get3(WF_F_A, &Particle::a);
    get3(WF_F_DISP, &Particle::Displacement);
    get1(WF_F_DENSITY, &Particle::Density);
    get1(WF_F_DDENSITY, &Particle::dDensity);
    get9(WF_F_SIGMA, &Particle::Sigma);
    get9(WF_F_SHEAR, &Particle::ShearStress);
    get9(WF_F_STRAIN_RATE, &Particle::StrainRate);
    get9(WF_F_ROT_RATE, &Particle::RotationRate);
    get9(WF_F_STRAIN, &Particle::Strain);
    get9(WF_F_STRAIN_PL, &Particle::Strain_pl);
    get9(WF_F_STRAIN_PL_INCR, &Particle::Strain_pl_incr);
    get1(WF_F_PL_STRAIN, &Particle::pl_strain);
    if (model_damage) get1(WF_F_DAMAGE, &Particle::dam_D);
    get1(WF_F_DPL_STRAIN, &Particle::delta_pl_strain);
    get1(WF_F_SIGMAY, &Particle::Sigmay);
    get1(WF_F_EFF_SR, &Particle::eff_strain_rate);
    get1(WF_F_PRESSURE, &Particle::Pressure);
    get1(WF_F_T, &Particle::T);
    get1(WF_F_DTDT, &Particle::dTdt);
    get1(WF_F_Q_PLHEAT, &Particle::q_plheat);
    get1(WF_F_Q_CONV, &Particle::q_conv);
    get1(WF_F_PS_ENERGY, &Particle::ps_energy);
    if (gradKernelCorr) get9(WF_F_GRADCORR, &Particle::gradCorrM);
    if (contact) {
      get3(WF_F_CONTFORCE, &Particle::contforce);
      std::vector<int32_t> id(n);
      check(wf_get_int_field(ctx_, WF_I_ID, id.data()));
      for (size_t i = 0; i < n; ++i) Particles[i]->ID = id[i];
      pull_rigid();
    }
    std::vector<int32_t> nb(n), cnb(n);
    check(wf_get_int_field(ctx_, WF_I_NB, nb.data()));
    check(wf_get_int_field(ctx_, WF_I_CONTNB, cnb.data()));
    for (size_t i = 0; i < n; ++i) {
      Particles[i]->Nb = nb[i];
      Particles[i]->ContNb = cnb[i];
      Particles[i]->CalculateEquivalentStress();
    }
    pull_scalars();
  }

  void pull_scalars() {
    wf_scalars s{};
    check(wf_get_scalars(ctx_, &s));
    kin_energy_sum = s.kin_energy_sum;
    contact_force_sum = s.contact_force_sum;
    contact_reaction_sum = s.contact_reaction_sum;
    ext_forces_work = s.ext_forces_work;
    contact_friction_work = s.contact_friction_work;
    accum_cont_heat_cond = s.accum_cont_heat_cond;
    min_force_ts = s.min_force_ts;
    int_energy_sum = s.int_energy_sum;
    plastic_work = s.plastic_work;
    m_maxT = s.max_T;
    m_minT = s.min_T;
  }

  // width of the float dataset starting at offset d of the XDMF table
  static int xdmf_width(int d) {
    switch (d) {
      case WF_XDMF_POSITION: case WF_XDMF_VELOCITY: case WF_XDMF_ACCELERATION:
      case WF_XDMF_DISPLACEMENT: case WF_XDMF_CONTACT_FORCE: return 3;
      case WF_XDMF_SIGMA: case WF_XDMF_SHEARS: case WF_XDMF_STRAIN: case WF_XDMF_STRAIN_RATE:
      case WF_XDMF_STRAIN_PL: case WF_XDMF_GRADCORRMAT: return 6;
      default: return 1;
    }
  }

  // rows [i0, i1) of the XDMF table of N particles from the Particle objects
  // (Source/Output.cpp:113-215)
  void host_xdmf(size_t i0, size_t i1, size_t N, float *f, int32_t *iv) {
    auto at = [&](int d, size_t i) { return f + (size_t)d * N + (size_t)xdmf_width(d) * i; };
    auto t6 = [](float *o, const Mat3_t &m) {
      o[0] = (float)m(0, 0), o[1] = (float)m(1, 1), o[2] = (float)m(2, 2);
      o[3] = (float)m(0, 1), o[4] = (float)m(1, 2), o[5] = (float)m(0, 2);
    };
    for (size_t i = i0; i < i1; ++i) {
      Particle *P = Particles[i];
      P->CalculateEquivalentStress();
      for (int k = 0; k < 3; ++k) {
        at(WF_XDMF_POSITION, i)[k] = (float)P->x(k);
        at(WF_XDMF_VELOCITY, i)[k] = (float)P->v(k);
        at(WF_XDMF_ACCELERATION, i)[k] = (float)P->a(k);
        at(WF_XDMF_DISPLACEMENT, i)[k] = (float)P->Displacement(k);
        at(WF_XDMF_CONTACT_FORCE, i)[k] = (float)P->contforce(k);
      }
      *at(WF_XDMF_PRESSURE, i) = (float)P->Pressure;
      *at(WF_XDMF_DENSITY, i) = (float)P->Density;
      *at(WF_XDMF_PROPERTY1, i) = *at(WF_XDMF_PROPERTY2, i) = *at(WF_XDMF_PROPERTY3, i) = 0.f;
      *at(WF_XDMF_MASS, i) = (float)P->Mass;
      *at(WF_XDMF_H, i) = (float)P->h;
      t6(at(WF_XDMF_SIGMA, i), P->Sigma);
      t6(at(WF_XDMF_SHEARS, i), P->Sigma);  // the reference writes Sigma under ShearS
      t6(at(WF_XDMF_STRAIN, i), P->Strain);
      t6(at(WF_XDMF_STRAIN_RATE, i), P->StrainRate);
      t6(at(WF_XDMF_STRAIN_PL, i), P->Strain_pl);
      t6(at(WF_XDMF_GRADCORRMAT, i), P->gradCorrM);
      *at(WF_XDMF_SIGMA_EQ, i) = (float)P->Sigma_eq;
      *at(WF_XDMF_TEMPERATURE, i) = (float)P->T;
      *at(WF_XDMF_PL_STRAIN, i) = (float)P->pl_strain;
      *at(WF_XDMF_Q_FRICTION, i) = (float)P->friction_hfl;
      *at(WF_XDMF_C_SHEARABS, i) = (float)P->cshearabs;
      *at(WF_XDMF_PS_EN, i) = (float)P->ps_energy;
      iv[(size_t)WF_XDMF_TAG * N + i] = P->ID;
      iv[(size_t)WF_XDMF_NEIGHBORS * N + i] = P->Nb;
      iv[(size_t)WF_XDMF_CONTNEIB * N + i] = P->ContNb;
    }
  }

  // rows [i0, i1) of the CSV table from the Particle objects
  void host_csv(size_t i0, size_t i1, double *rows) {
    for (size_t i = i0; i < i1; ++i) {
      Particle *P = Particles[i];
      P->CalculateEquivalentStress();
      double *r = rows + (size_t)WF_CSV_NCOL * (i - i0);
      const double v[WF_CSV_NCOL] = {P->x(0), P->x(1), P->x(2), (double)P->ID, P->Sigma_eq, P->pl_strain,
                                     P->v(0), P->v(1), P->v(2), P->a(0), P->a(1), P->a(2),
                                     P->contforce(0), P->contforce(1), P->contforce(2), P->Pressure, P->T};
      std::memcpy(r, v, sizeof v);
    }
  }

  void write_xdmf(char const *key, size_t N, const std::vector<float> &f, const std::vector<int32_t> &iv) {
    const char *names[3] = {OutputName[0].c_str(), OutputName[1].c_str(), OutputName[2].c_str()};
    if (wfio_write_xdmf(key, (int64_t)N, f.data(), iv.data(), names) != 0) throw new Fatal(wfio_last_error());
  }
  void write_csv(char const *key, size_t N, const std::vector<double> &rows) {
    const int threads = Nproc > 0 ? (int)Nproc : 1;
    if (wfio_write_csv(key, (int64_t)N, rows.data(), threads) != 0) throw new Fatal(wfio_last_error());
  }

  // positions and velocities of the rigid-surface particles
  void pull_rigid() {
    const size_t n = nsolid(), nr = Particles.Size() - n;
    if (!contact || nr == 0) return;
    std::vector<double> r3(3 * nr);
    auto rig = [&](int f, Vec3_t Particle::*m) {
      check(wf_get_rigid_field(ctx_, f, r3.data()));
      for (size_t r = 0; r < nr; ++r)
        for (int k = 0; k < 3; ++k) (Particles[n + r]->*m)(k) = r3[3 * r + k];
    };
    rig(WF_R_X, &Particle::x);
    rig(WF_R_V, &Particle::v);
    rig(WF_R_NORMAL, &Particle::normal);
  }

  void pull_reactions(int64_t nsteps) {
    std::vector<double> h(3 * (size_t)nsteps);
    for (int m = 0; m < meshcount; ++m) {
      int64_t got = 0;
      check(wf_get_contact_force_history(ctx_, m, h.data(), nsteps, &got));
      auto &dst = cforce_hist_[m];
      for (int64_t s = 0; s < got; ++s) dst.push_back(Vec3_t(h[3 * s], h[3 * s + 1], h[3 * s + 2]));
      if (got > 0) m_contact_force[m] = dst.back();
    }
    for (const auto &b : bConds) {
      if (b.type != Velocity_BC) continue;
      int64_t got = 0;
      check(wf_get_reaction_history(ctx_, b.zoneId, h.data(), nsteps, &got));
      auto &dst = reaction_[b.zoneId];
      for (int64_t s = 0; s < got; ++s) dst.push_back(Vec3_t(h[3 * s], h[3 * s + 1], h[3 * s + 2]));
    }
  }

  // SolveDiffUpdateLeapFrog / KickDrift (Source/SolverLeapfrog.cpp:24-451,
  // Source/SolverKickDrift.cpp:23-401): the loop runs while Time <= tf and
  // idx_out <= maxidx and writes output after the step that reaches tout.
  // With a fixed step, Time advances on the host with the same arithmetic,
  // so the device runs whole stretches of steps between outputs without
  // synchronising.  With the adaptive step the device holds Time; a step
  // never exceeds max(2 x the previous step, manual_min_ts) nor
  // max(dt, manual_min_ts) (AdaptiveTimeStep blends harmonically towards
  // the initial dt), so each stretch is as long as cannot reach tout or tf
  // before its last step, and Time is read back after it.
  void solve(double tf, double dt, double dtOut, char const *key, size_t maxidx, int scheme) {
    check_supported();
    std::vector<wf_material> table;
    const std::vector<int> mi = build_materials(table);
    close();
    const size_t n = nsolid();
    wf_config cfg{};
    cfg.precision = (int32_t)precision;
    cfg.ts_nb_inc = ts_nb_inc;
    cfg.max_nb = max_nb;
    cfg.thermal_solver = thermal_solver;
    cfg.pl_heating = pl_heating;
    cfg.device = device;
    cfg.cellfac = Cellfac;
    cfg.pl_work_heat_frac = pl_work_heat_frac;
    cfg.auto_ts = (auto_ts ? 1 : 0) | (auto_ts_acc ? 2 : 0) | (auto_ts_cont ? 4 : 0);
    cfg.cfl = CFL;
    cfg.sqrt_h_a = sqrt_h_a;
    cfg.min_ts = (double)manual_min_ts;
    if (nonlock_sum && gradKernelCorr) {
      // Source/SolverLeapfrog.cpp:114-117
      std::cout << "WARNING: Nishimura summation is not working with Gradient Kernel Correction. Summation "
                   "changed to Locking."
                << std::endl;
      nonlock_sum = false;
    }
    cfg.grad_corr = gradKernelCorr ? 1 : 0;
    cfg.locking = nonlock_sum ? 0 : 1;
    for (int k = 0; k < 3; ++k) {
      cfg.dom_max[k] = DomMax(k);
      cfg.dom_min[k] = DomMin(k);
    }
    if (wf_create(&cfg, (int64_t)n, &ctx_) != WF_OK) throw new Fatal(std::string("weldform_hip: ") + wf_last_error(nullptr));
    for (size_t k = 0; k < table.size(); ++k) check(wf_set_material(ctx_, (int32_t)k, &table[k]));
    if (model_damage) {
      // each material's JohnsonCookDamage (Source/Damage.cpp:9-46)
      std::vector<char> done(table.size(), 0);
      for (size_t i = 0; i < n; ++i) {
        const Particle *P = Particles[i];
        if (done[mi[i]] || !P->mat || !P->mat->damage) continue;
        const JohnsonCookDamage *J = dynamic_cast<const JohnsonCookDamage *>(P->mat->damage);
        if (!J) throw new Fatal("only the Johnson-Cook damage criterion is implemented on the device path");
        const double D[5] = {J->m_D1, J->m_D2, J->m_D3, J->m_D4, J->m_D5};
        check(wf_set_material_damage(ctx_, mi[i], 1, D));
        done[mi[i]] = 1;
      }
    }
    check(wf_set_int_field(ctx_, WF_I_MATERIAL, mi.data()));
    upload();
    for (const auto &a : amps)
      check(wf_add_amplitude(ctx_, a.id, (int32_t)a.time.size(), a.time.data(), a.value.data()));
    for (const auto &b : bConds) {
      if (b.type == Velocity_BC) {
        const double v[3] = {b.value(0), b.value(1), b.value(2)};
        check(wf_add_velocity_bc(ctx_, b.zoneId, b.valueType, v, b.ampId, b.ampFactor));
        if (norm(b.value_ang) != 0.0) {
          const double w[3] = {b.value_ang(0), b.value_ang(1), b.value_ang(2)};
          check(wf_set_bc_angular_velocity(ctx_, b.zoneId, w));
        }
      } else if (b.type == Temperature_BC) {
        check(wf_add_temperature_bc(ctx_, b.zoneId, b.T));
      }
    }
    for (size_t m = 0; m < trimesh.size(); ++m) {
      const TriMesh *M = trimesh[m];
      int rc;
      if (M->dens > 0) {
        const double a[3] = {M->p1(0), M->p1(1), M->p1(2)}, b[3] = {M->p2(0), M->p2(1), M->p2(2)};
        rc = wf_add_rigid_plane(ctx_, M->id, M->axis, M->flip ? 1 : 0, a, b, M->dens, (double)mesh_hfac_[m]);
      } else {
        // a mesh read from a file: its nodes, triangles and current normals
        std::vector<double> nd(3 * M->node.Size()), nr(3 * M->element.Size());
        std::vector<int32_t> el(3 * M->element.Size());
        for (size_t n = 0; n < M->node.Size(); ++n)
          for (int k = 0; k < 3; ++k) nd[3 * n + k] = (*M->node[n])(k);
        for (size_t e = 0; e < M->element.Size(); ++e)
          for (int k = 0; k < 3; ++k) {
            el[3 * e + k] = M->element[e]->node[k];
            nr[3 * e + k] = M->element[e]->normal(k);
          }
        rc = wf_add_rigid_mesh(ctx_, M->id, (int64_t)M->node.Size(), nd.data(), (int64_t)M->element.Size(), el.data(),
                               nr.data(), (double)mesh_hfac_[m]);
      }
      if (rc < 0) check(rc);
    }
    if (Dimension == 2) check(wf_set_dimension(ctx_, 2));
    if (contact) {
      check(wf_set_contact(ctx_, friction_sta, friction_dyn, DFAC));
      check(wf_set_contact_law(ctx_, (int32_t)contact_alg, PFAC, fric_type == Fr_Bound ? 1 : 0));
      if (friction_function == Linear) check(wf_set_friction_function(ctx_, 1, friction_m, friction_b));
      if (cont_heat_cond || cont_heat_fric) check(wf_set_contact_heat(ctx_, cont_heat_cond, cont_heat_fric, contact_hc));
      // the rigid particles' temperatures as the initial conditions and
      // temperature BCs of their zones left them
      const size_t ns = nsolid();
      std::vector<double> rt(Particles.Size() - ns);
      for (size_t r = 0; r < rt.size(); ++r) rt[r] = Particles[ns + r]->T;
      check(wf_set_rigid_temperature(ctx_, (int64_t)rt.size(), rt.data()));
    }
    deltat = dt;
    check(wf_init(ctx_, dt));
    if (key && wfio_has_hdf5()) {
      // the initial model (Source/SolverLeapfrog.cpp:57-63)
      WriteDeviceOutput((std::string(key) + "_Initial").c_str(), true, false);
    }
    reaction_.clear();
    cforce_hist_.clear();
    size_t idx_out = 1;
    double tout = Time;
    const bool adaptive = auto_ts || auto_ts_acc;
    while (Time <= tf && idx_out <= maxidx) {
      int64_t k = 0;
      double t = Time;
      bool out = false;
      // host hooks: one step per stretch, cut at the hooks' call sites
      const int64_t chunk = host_hooks() ? 1 : kMaxChunk;
      if (!adaptive) {
        while (t <= tf && k < chunk) {
          t += deltat;
          ++k;
          if (t >= tout) {
            out = true;
            break;
          }
        }
      } else {
        // largest k whose first k-1 steps surely end before min(tout, tf)
        const double rem = std::min(tout, tf) - Time, floor_ts = (double)manual_min_ts;
        double bound = 0.0, step = deltat;
        k = 1;
        while (k < chunk) {
          step = std::min(std::max(2.0 * step, floor_ts), std::max(dt, floor_ts));
          if ((bound + step) * (1.0 + 1e-9) >= rem) break;
          bound += step;
          ++k;
        }
      }
      if (host_hooks())
        step_with_hooks(scheme);
      else
        check(wf_step(ctx_, scheme, k));
      check(wf_sync(ctx_));
      pull_reactions(k);
      if (adaptive) {
        wf_scalars s;
        check(wf_get_scalars(ctx_, &s));
        Time = s.time;
        deltat = s.dt;
        out = Time >= tout;
      } else {
        Time = t;
      }
      if (out) {
        if (key) {
          // SolverLeapfrog.cpp:364-373: FileKey_%04d.{hdf5,xmf,csv}
          char fn[4096];
          std::snprintf(fn, sizeof fn, "%s_%04zu", key, idx_out);
          WriteDeviceOutput(fn, wfio_has_hdf5() != 0, true);
        }
        idx_out++;
        tout += dtOut;
      }
    }
    download();
  }
};

}  // namespace SPH
