// WeldForm command-line driver on the device path: reads a scene file in the
// format of the reference's driver (Source/WeldForm.cpp:147-944, keys and
// defaults as there) and runs SolveDiffUpdateLeapFrog / KickDrift through
// the SPH::Domain facade on one GPU.
//
//   WeldForm input.json [--dry-run] [--mixed] [--out KEY] [--dump FILE]
//
// --dry-run builds the domain and prints the set-up without touching the
// device; --mixed selects the mixed-precision step; --out sets the output
// file key (the reference writes "test06_NNNN"); --dump writes the final
// displacement and plastic strain of every particle at full precision.  Scene features off the
// device path (2D / axisymmetric blocks, LS-Dyna particle files, Rankine
// damage, GMT, the Fraser and thermal-only solvers, smoothing-length update)
// stop with a Fatal naming the key.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "weldform/Domain.h"
#include "weldform/Json.h"

using namespace SPH;

static bool in(const std::string &s, std::initializer_list<const char *> v) {
  for (const char *x : v)
    if (s == x) return true;
  return false;
}

int main(int argc, char **argv) try {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s input.json [--dry-run] [--mixed] [--out KEY] [--dump FILE]\n", argv[0]);
    return 2;
  }
  bool dry = false, mixed = false;
  std::string key = "test06", dump;
  for (int a = 2; a < argc; ++a) {
    if (!std::strcmp(argv[a], "--dry-run")) dry = true;
    else if (!std::strcmp(argv[a], "--mixed")) mixed = true;
    else if (!std::strcmp(argv[a], "--out") && a + 1 < argc) key = argv[++a];
    else if (!std::strcmp(argv[a], "--dump") && a + 1 < argc) dump = argv[++a];
    else throw new Fatal(std::string("unknown option ") + argv[a]);
  }
  const json::Value j = json::parse_file(argv[1]);
  const json::Value &config = j["Configuration"], &material = j["Materials"], &domblock = j["DomainBlocks"],
                    &domzones = j["DomainZones"], &amplitudes = j["Amplitudes"], &rigbodies = j["RigidBodies"],
                    &contact_ = j["Contact"], &bcs = j["BoundaryConditions"], &ics = j["InitialConditions"];

  Domain dom;
  dom.Dimension = 3;
  dom.precision = mixed ? Precision::Mixed : Precision::FP64;
  int np = 4;
  readValue(config["Nproc"], np);
  dom.Nproc = np;
  std::string sumType = "Nishimura";
  readValue(config["sumType"], sumType);
  if (sumType == "Locking") dom.nonlock_sum = false;
  else if (sumType == "Nishimura") dom.nonlock_sum = true;
  double ts = 0.0;
  readValue(config["timeStepSize"], ts);
  dom.Kernel_Set(Qubic_Spline);
  std::string solver = "Mech";
  readValue(config["solver"], solver);
  if (in(solver, {"Thermal", "Mech-Thermal", "Mech-Thermal-KickDrift", "Mech-Thermal-Fraser", "Mech-Thermal-LeapFrog"}))
    dom.thermal_solver = true;
  readValue(config["integrationMethod"], dom.Scheme);
  double r = 0.0, hfactor = 0.0;
  readValue(config["particleRadius"], r);
  readValue(config["hFactor"], hfactor);
  if (!(r > 0.0) || !(hfactor > 0.0)) throw new Fatal("Configuration needs particleRadius and hFactor");

  // ------------------------------------------------------------ material
  double rho = 0.0, E = 0.0, nu = 0.0, Fy = -1.0, Ep = 0.0;
  std::vector<double> c(10, 0.0);
  std::string mattype = "Bilinear";
  bool plastic_heat = false;
  const json::Value &m0 = material[0];
  readValue(m0["type"], mattype);
  readValue(m0["density0"], rho);
  readValue(m0["youngsModulus"], E);
  readValue(m0["poissonsRatio"], nu);
  readValue(m0["yieldStress0"], Fy);
  readArray(m0["const"], c);
  dom.pl_heating = (readValue(m0["plasticHeat"], plastic_heat), plastic_heat);
  Elastic_ el(E, nu);
  Material_ *mat = nullptr;
  int model = BILINEAR;
  if (mattype == "Bilinear") {
    if (c.empty()) throw new Fatal("Bilinear material needs const [Et]");
    Ep = E * c[0] / (E - c[0]);
    mat = new Material_(el);
  } else if (mattype == "Hollomon") {
    if (c.size() < 2) throw new Fatal("Hollomon material needs const [K, m]");
    mat = new Hollomon(el, Fy, c[0], c[1]);
    model = HOLLOMON;
  } else if (mattype == "JohnsonCook") {
    if (c.size() < 8) throw new Fatal("JohnsonCook material needs const [A, B, n, C, eps_0, m, T_m, T_t]");
    mat = new JohnsonCook(el, c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7]);
    model = JOHNSON_COOK;
  } else if (mattype == "GMT") {
    throw new Fatal("the GMT material is not implemented on the device path");
  } else {
    throw new Fatal("Invalid material type.");
  }
  // Source/WeldForm.cpp:293-321: the damage model switches the Locking sums on
  std::string damage_mod;
  readValue(m0["damageModel"], damage_mod);
  if (damage_mod == "JohnsonCook") {
    std::vector<double> D(5, 0.0);
    readArray(m0["damageParams"], D);
    if (D.size() < 5) throw new Fatal("JohnsonCook damage needs damageParams [D1, D2, D3, D4, D5]");
    if (model != JOHNSON_COOK) throw new Fatal("JohnsonCook damage needs a JohnsonCook material on the device path");
    mat->damage = new JohnsonCookDamage(D[0], D[1], D[2], D[3], D[4], ((JohnsonCook *)mat)->getRefStrainRate());
    mat->damage->mat = mat;
    std::cout << "Assigned Johnson Cook Damage Model" << std::endl;
    dom.model_damage = true;
    dom.nonlock_sum = false;
  } else if (damage_mod == "Rankine") {
    throw new Fatal("the Rankine damage criterion is not implemented on the device path");
  } else if (!damage_mod.empty()) {
    throw new Fatal("unknown damageModel " + damage_mod);
  }
  double k_T = 0.0, cp_T = 0.0, th_ex = 0.0;
  readValue(m0["thermalCond"], k_T);
  readValue(m0["thermalHeatCap"], cp_T);
  readValue(m0["thermalExp"], th_ex);
  const double K = E / (3. * (1. - 2 * nu));
  const double G = E / (2. * (1. + nu));
  const double dx = 2. * r;
  const double h = dx * hfactor;
  const double Cs = std::sqrt(K / rho);

  // ------------------------------------------------------------ time step, switches
  double timestep = 0.0, cflFactor = 0.0;
  int cflMethod = 0;
  double output_time = 0.0, sim_time = 0.0;
  std::string cont_alg = "Wang";
  bool auto_ts[] = {true, false, false};
  readValue(config["cflMethod"], cflMethod);
  if (cflMethod == 0) {
    readValue(config["timeStepSize"], timestep);
  } else {
    readValue(config["cflFactor"], cflFactor);
    timestep = (cflFactor * h / (Cs));
  }
  if (!(timestep > 0.0)) throw new Fatal("no time step: set timeStepSize (cflMethod 0) or cflFactor");
  double min_ts_manual = timestep / 1.e-5;
  readValue(config["minTS"], min_ts_manual);
  dom.manual_min_ts = (int)min_ts_manual;  // an int member, as in the reference
  readValue(config["outTime"], output_time);
  readValue(config["simTime"], sim_time);
  readBoolVector(config["autoTS"], auto_ts);
  double alpha = 1., beta = 0.;
  bool h_upd = false;
  double tensins = 0.3;
  int nb_upd_freq = 5;
  bool kernel_grad_corr = false;
  int gradType = 0;
  readValue(config["artifViscAlpha"], alpha);
  readValue(config["artifViscBeta"], beta);
  readValue(config["contAlgorithm"], cont_alg);
  readValue(config["kernelGradCorr"], kernel_grad_corr);
  readValue(config["stressGradType"], gradType);
  readValue(config["smoothlenUpdate"], h_upd);
  readValue(config["nbsearchFreq"], nb_upd_freq);
  dom.auto_ts = auto_ts[0];
  dom.auto_ts_acc = auto_ts[1];
  dom.auto_ts_cont = auto_ts[2];
  dom.Gradient_Approach_Set((Gradient_Type)gradType);
  readValue(config["tensileInstability"], tensins);
  if (h_upd) throw new Fatal("smoothing-length update (smoothlenUpdate) is not implemented on the device path");

  // ------------------------------------------------------------ domain block
  Vec3_t start(0, 0, 0), L(0, 0, 0);
  int id = 0;
  std::string domtype = "Box";
  int matID = 0;
  std::string gridCS = "Cartesian";
  double slice_ang = 2.0000001 * M_PI;
  bool sym[] = {false, false, false};
  const json::Value &b0 = domblock[0];
  readValue(b0["id"], id);
  readVector(b0["start"], start);
  readVector(b0["dim"], L);
  readValue(b0["type"], domtype);
  readValue(b0["matID"], matID);
  readValue(b0["gridCoordSys"], gridCS);
  readValue(b0["sliceAngle"], slice_ang);
  readBoolVector(b0["sym"], sym);
  for (int i = 0; i < 3; i++) {  // only DomMax(0) / DomMin(0), as the reference writes them
    dom.DomMax(0) = L[i];
    dom.DomMin(0) = -L[i];
  }
  // a block thinner than h is plane strain (Source/WeldForm.cpp:420-424)
  if (std::fabs(L[2]) < h) {
    dom.Dimension = 2;
    std::cout << "Z Value is less than h. Dimension is set to 2." << std::endl;
  }
  if (domtype == "Box") {
    if (gridCS == "AxiSymmetric") throw new Fatal("axisymmetric blocks are not implemented on the device path");
    dom.AddBoxLength(id, start, L[0], L[1], L[2], r, rho, h, 1, 0, false, false);
  } else if (domtype == "Cylinder") {
    if (sym[0] && sym[1]) throw new Fatal("symmetric cylinder blocks (sym) are not implemented on the device path");
    if (gridCS != "Cartesian")
      throw new Fatal("cylinder grid '" + gridCS + "' is not implemented on the device path (use Cartesian)");
    dom.AddCylinderLength(0, start, L[0] / 2., L[2], r, rho, h, false, sym[2]);
  } else if (domtype == "File") {
    throw new Fatal("particle files (DomainBlocks type File) are not implemented on the device path");
  } else {
    throw new Fatal("unknown DomainBlocks type " + domtype);
  }

  // ------------------------------------------------------------ zones
  for (size_t z = 0; z < domzones.size(); ++z) {
    const json::Value &zone = domzones[z];
    int zoneid = 0;
    Vec3_t zs(0, 0, 0), ze(0, 0, 0);
    readValue(zone["id"], zoneid);
    readVector(zone["start"], zs);
    readVector(zone["end"], ze);
    const int cnt = dom.AssignZone(zs, ze, zoneid);
    std::printf("Zone %d, particle count: %d\n", zoneid, cnt);
  }

  // ------------------------------------------------------------ rigid bodies
  std::vector<TriMesh *> mesh;
  for (size_t rb = 0; rb < rigbodies.size(); ++rb) {
    const json::Value &body = rigbodies[rb];
    std::string rigbody_type;
    if (!readValue(body["type"], rigbody_type)) continue;
    Vec3_t bs(0, 0, 0), dim(0, 0, 0), translation(0, 0, 0);
    readVector(body["start"], bs);
    readVector(body["dim"], dim);
    bool flipnormals = false;
    readValue(body["flipNormals"], flipnormals);
    if (rigbody_type != "Plane" && rigbody_type != "File")
      throw new Fatal("rigid body type " + rigbody_type + " is not implemented on the device path (Plane and File only)");
    if (rigbody_type == "Plane" && dim(0) != 0. && dim(1) != 0. && dim(2) != 0.)
      throw new Fatal("ERROR: Contact Plane Surface should have one null dimension");
    dom.contact = true;
    TriMesh *M;
    if (rigbody_type == "Plane") {
      int dens = 10;
      readValue(body["partSide"], dens);
      M = new TriMesh;
      M->AxisPlaneMesh(2, flipnormals, bs, Vec3_t(bs(0) + dim(0), bs(1) + dim(1), bs(2)), dens);
    } else {
      std::string filename;
      readValue(body["fileName"], filename);
      std::printf("Reading Mesh input file %s\n", filename.c_str());
      NastranReader reader(filename.c_str());
      M = new TriMesh(reader, flipnormals);
    }
    mesh.push_back(M);
    // scale and translation as the reference applies them (its TriMesh
    // helpers recompute centroids and normals from the winding); a plane
    // changed this way goes to the device as a general mesh
    double scalefactor = 1.0;
    readValue(body["scaleFactor"], scalefactor);
    if (scalefactor != 1.0) {
      M->Scale(scalefactor);
      M->dens = 0;
    }
    readVector(body["translation"], translation);
    if (norm(translation) > 1.0e-6) {
      for (size_t n = 0; n < M->node.Size(); ++n) *M->node[n] = *M->node[n] + translation;
      M->CalcCentroids();
      M->CalcNormals();
      M->dens = 0;
    }
    M->CalcSpheres();
    int zid = 0;
    readValue(body["zoneId"], zid);
    dom.AddTrimeshParticles(M, 1.1, zid);
    double penaltyfac = 0.5;
    double fric_sta = 0.0, fric_dyn = 0.0, heat_cond = 0.0;
    readValue(contact_[0]["fricCoeffStatic"], fric_sta);
    readValue(contact_[0]["fricCoeffDynamic"], fric_dyn);
    readValue(contact_[0]["heatCondCoeff"], heat_cond);
    readValue(contact_[0]["penaltyFactor"], penaltyfac);
    // Source/WeldForm.cpp:695-705: "Fraser" (disabled there) and any other
    // value keep Wang's law
    if (cont_alg == "Seo") dom.contact_alg = Seo;
    else if (cont_alg == "LSDyna") dom.contact_alg = LSDyna;
    bool heat_cond_ = false, heat_fric_ = false;
    readValue(contact_[0]["heatConductance"], heat_cond_);
    readValue(contact_[0]["heatFriction"], heat_fric_);
    if (heat_cond_) {
      dom.cont_heat_cond = true;
      dom.contact_hc = heat_cond;
    }
    if (heat_fric_) dom.cont_heat_fric = true;
    dom.friction_dyn = fric_dyn;
    dom.friction_sta = fric_sta;
    dom.PFAC = penaltyfac;
    dom.DFAC = 0.0;
  }

  // ------------------------------------------------------------ amplitudes
  for (size_t a = 0; a < amplitudes.size(); ++a) {
    const json::Value &ampl = amplitudes[a];
    amplitude amp;
    std::vector<double> time, value;
    readValue(ampl["id"], amp.id);
    readArray(ampl["time"], time);
    readValue(ampl["value"], value);
    if (value.size() < time.size()) throw new Fatal("amplitude with fewer values than times");
    for (size_t i = 0; i < time.size(); ++i) {
      amp.time.push_back(time[i]);
      amp.value.push_back(value[i]);
    }
    dom.amps.push_back(amp);
  }

  // ------------------------------------------------------------ boundary conditions
  for (size_t q = 0; q < bcs.size(); ++q) {
    const json::Value &bc = bcs[q];
    boundaryCondition bcon;
    bcon.type = 0;
    bcon.valueType = 0;
    bcon.value_ang = Vec3_t(0, 0, 0);
    readValue(bc["zoneId"], bcon.zoneId);
    readValue(bc["type"], bcon.type);
    readValue(bc["valueType"], bcon.valueType);
    readVector(bc["value"], bcon.value);
    readVector(bc["valueAng"], bcon.value_ang);
    if (bcon.valueType == 1) {
      double ampfactor = 1.0;
      readValue(bc["amplitudeId"], bcon.ampId);
      readValue(bc["amplitudeFactor"], ampfactor);
      bcon.ampFactor = ampfactor;
    }
    if (bcon.type == Convection_BC) {
      readValue(bc["convCoeff"], bcon.cv_coeff);
      readValue(bc["infTemp"], bcon.T_inf);
      for (size_t p = 0; p < dom.Particles.Size(); ++p)
        if (dom.Particles[p]->ID == bcon.zoneId) {
          dom.Particles[p]->Thermal_BC = TH_BC_CONVECTION;
          dom.Particles[p]->T_inf = bcon.T_inf;
          dom.Particles[p]->h_conv = bcon.cv_coeff;
        }
    } else if (bcon.type == Temperature_BC) {
      readValue(bc["Temp"], bcon.T);
      for (size_t p = 0; p < dom.Particles.Size(); ++p)
        if (dom.Particles[p]->ID == bcon.zoneId) dom.Particles[p]->T = bcon.T;
    } else if (bcon.type != Velocity_BC) {
      throw new Fatal("boundary condition type " + std::to_string(bcon.type) + " is not implemented on the device path");
    }
    readValue(bc["free"], bcon.free);
    dom.bConds.push_back(bcon);
  }

  // ------------------------------------------------------------ initial conditions
  double IniTemp = 0.;
  for (size_t q = 0; q < ics.size(); ++q) {
    const json::Value &ic = ics[q];
    int icid = 0;
    readValue(ic["Temp"], IniTemp);
    readValue(ic["id"], icid);
    std::vector<double> value(3, 0.0);
    readValue(ic["value"], value);
    value.resize(3, 0.0);
    int zoneid = -1;
    readValue(ic["zoneId"], zoneid);
    if (zoneid == -1) {
      for (int p = 0; p < dom.solid_part_count; ++p) {
        dom.Particles[p]->a = Vec3_t(0., 0., 0.);
        dom.Particles[p]->v = Vec3_t(value[0], value[1], value[2]);
        if (dom.thermal_solver) dom.Particles[p]->T = IniTemp;
      }
    } else {
      for (size_t p = 0; p < dom.Particles.Size(); ++p)
        if (dom.Particles[p]->ID == zoneid && dom.thermal_solver) dom.Particles[p]->T = IniTemp;
    }
  }

  dom.gradKernelCorr = kernel_grad_corr;
  dom.ts_nb_inc = nb_upd_freq;
  if (dom.Particles.Size() == 0) throw new Fatal("Particle Count is Null. Please Check Radius and Domain Dimensions.");
  for (size_t a = 0; a < dom.Particles.Size(); ++a) {
    Particle *P = dom.Particles[a];
    P->G = G;
    P->Cs = Cs;
    P->Shepard = false;
    if (mattype == "Bilinear") P->Ep = Ep;
    P->Material_model = model;
    P->mat = mat;
    P->Fail = 1;
    P->Alpha = alpha;
    P->Beta = beta;
    P->TI = tensins;
    P->TIInitDist = dx;
    P->k_T = k_T;
    P->cp_T = cp_T;
    P->th_exp = th_ex;
    if (mattype == "Hollomon" || mattype == "JohnsonCook") {
      P->Sigmay = mat->CalcYieldStress(0.0, 0.0, P->T);
    } else {
      if (Fy > 0.0) P->Sigmay = Fy;
      else throw new Fatal("Invalid Initial Yield Stress.");
    }
  }

  if (in(solver, {"Mech-Fraser", "Mech-Thermal-Fraser"}))
    throw new Fatal("the Fraser solver is not on the device path (use Mech or Mech-LeapFrog)");
  if (solver == "Thermal") throw new Fatal("the thermal-only solver is not on the device path");
  const bool kickdrift = in(solver, {"Mech-KickDrift", "Mech-Thermal-KickDrift"});
  if (!kickdrift && !in(solver, {"Mech", "Mech-Thermal", "Mech-LeapFrog", "Mech-Thermal-LeapFrog"}))
    throw new Fatal("Invalid solver.");
  std::printf("WeldForm (device path): %zu particles (%d solid), dx %.17g, h %.17g, Cs %.17g, dt %.17g, simTime %.17g, "
              "outTime %.17g, solver %s, material %s, TI %.3g, autoTS [%d %d %d], contact %d, thermal %d\n",
              dom.Particles.Size(), dom.solid_part_count, dx, h, Cs, timestep, sim_time, output_time, solver.c_str(),
              mattype.c_str(), tensins, (int)auto_ts[0], (int)auto_ts[1], (int)auto_ts[2], (int)dom.contact,
              (int)dom.thermal_solver);
  std::printf("Sigmay0 %.17g, Ep %.17g, BCs %zu, amplitudes %zu, meshes %zu\n", dom.Particles[0]->Sigmay,
              dom.Particles[0]->Ep, dom.bConds.size(), dom.amps.size(), mesh.size());
  if (dom.contact) {
    // contact heat and the rigid particles' temperatures (first and last)
    const size_t ns = (size_t)dom.solid_part_count, nall = dom.Particles.Size();
    std::printf("contact heat conduction %d (hc %.17g), friction %d, rigid T %.17g .. %.17g\n", (int)dom.cont_heat_cond,
                dom.contact_hc, (int)dom.cont_heat_fric, nall > ns ? dom.Particles[ns]->T : 0.0,
                nall > ns ? dom.Particles[nall - 1]->T : 0.0);
  }
  if (dry) {
    for (TriMesh *M : mesh) delete M;
    delete mat;
    return 0;
  }
  if (kickdrift) dom.SolveDiffUpdateKickDrift(sim_time, timestep, output_time, key.c_str(), 1000000);
  else dom.SolveDiffUpdateLeapFrog(sim_time, timestep, output_time, key.c_str(), 1000000);
  // the state after the solve goes to maz.hdf5 / maz.xmf (Source/WeldForm.cpp:941);
  // without the HDF5 library the solve's CSV files are the output
  if (wfio_has_hdf5()) dom.WriteXDMF("maz");
  if (!dump.empty()) {
    FILE *f = std::fopen(dump.c_str(), "w");
    if (!f) throw new Fatal("cannot write " + dump);
    for (int a = 0; a < dom.solid_part_count; ++a) {
      const Particle *P = dom.Particles[a];
      std::fprintf(f, "%.17g %.17g %.17g %.17g\n", P->Displacement(0), P->Displacement(1), P->Displacement(2),
                   P->pl_strain);
    }
    std::fclose(f);
  }
  std::printf("done: time %.9g, step %.6g, kinetic %.9g, internal %.9g\n", dom.getTime(), dom.getStepSize(),
              dom.kin_energy_sum, dom.int_energy_sum);
  // the totals the reference prints with its outputs (Source/SolverLeapfrog.cpp:404-407)
  if (dom.cont_heat_cond) std::printf("Total contact heat flux%.9g\n", dom.accum_cont_heat_cond);
  if (dom.cont_heat_fric) std::printf("Contact friction work %.9g\n", dom.contact_friction_work);
  for (TriMesh *M : mesh) delete M;
  delete mat;
  return 0;
} catch (Fatal *e) {
  std::fprintf(stderr, "Fatal: %s\n", e->Msg().c_str());
  delete e;
  return 1;
}
