/*
 * weldform_hip.h -- C-ABI of the MI355X explicit SPH step (libweldform_hip.so).
 *
 * Plain C types only: pointers, sizes, POD structs.  Every call returns 0 on
 * success or a negative WF_E* code; wf_last_error() gives the message.  No
 * C++ exception crosses this boundary.  One host thread per context; each
 * context owns its device memory and one HIP stream (default priority).
 *
 * Particle data crosses the boundary in the reference's layout and in the
 * caller's particle order: Vec3_t fields as 3 doubles per particle (x,y,z),
 * Mat3_t fields as 9 doubles per particle (row-major 3x3).  Internally the
 * state is kept in a Morton-sorted structure-of-arrays; the library maps
 * indices both ways.
 *
 * Reference interfaces replaced (luchete80/WeldForm):
 *   wf_create/wf_set_*       <- SPH::Domain / SPH::Particle public members
 *                               (Source/Domain.h:113-537, Source/Particle.h:50-265)
 *   wf_add_velocity_bc etc.  <- Domain::bConds / amps and the UserAcc hook
 *                               (Source/Domain.h:69-111, Source/WeldForm.cpp:51-120)
 *   wf_init                  <- set-up part of SolveDiffUpdateLeapFrog
 *                               (Source/SolverLeapfrog.cpp:46-103)
 *   wf_step(WF_LEAPFROG)     <- one iteration of SolveDiffUpdateLeapFrog's loop
 *                               (Source/SolverLeapfrog.cpp:124-362)
 *   wf_step(WF_KICKDRIFT)    <- one iteration of SolveDiffUpdateKickDrift's loop
 *                               (Source/SolverKickDrift.cpp:117-330)
 *   wf_rebuild_neighbours    <- ClearNbData + MainNeighbourSearch + CalcPairPosList
 *                               (Source/Neighbour.cpp:346-359,479-497,617-694)
 *   wf_eval(WF_EVAL_ACCEL)   <- StartAcceleration's tensile-instability stress
 *                               (Source/Domain.cpp:1519-1597) + CalcAccel +
 *                               AccelReduction (Source/InteractionAlt.cpp:12-260)
 *   wf_eval(WF_EVAL_DENSITY) <- CalcDensInc + DensReduction (Source/InteractionAlt.cpp:483-640)
 *   wf_eval(WF_EVAL_RATES)   <- CalcRateTensors + RateTensorsReduction
 *                               (Source/InteractionAlt.cpp:263-479)
 *   wf_eval(WF_EVAL_STRESS)  <- Particle::CalcStressStrain for every particle
 *                               (Source/Particle.cpp:960-1067)
 *   wf_gen_cylinder          <- Domain::AddCylinderLength (Source/Domain.cpp:708-860)
 *   wf_pack_xdmf / wf_pack_csv <- the particle tables of Domain::WriteXDMF /
 *                               Domain::WriteCSV (Source/Output.cpp:75-554); the
 *                               files themselves: weldform_io.h
 *   wf_adj_*                 <- no counterpart in the reference: the reverse-mode
 *                               adjoint of the FP64 Leapfrog step for parameter
 *                               calibration (see the block above wf_adj_forward)
 */
#ifndef WELDFORM_HIP_H
#define WELDFORM_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WF_ABI_VERSION 8

/* error codes */
#define WF_OK 0
#define WF_EINVAL (-1)      /* bad argument                                  */
#define WF_EHIP (-2)        /* HIP runtime error                             */
#define WF_ESTATE (-3)      /* call out of order (e.g. step before init)     */
#define WF_ENBOVERFLOW (-4) /* a particle has more neighbours than max_nb    */
#define WF_ENAN (-5)        /* non-finite value detected in the state        */
#define WF_EUNSUPPORTED (-6)/* feature outside the implemented path          */
#define WF_ENOMEM (-7)

/* precision of the pair passes */
#define WF_FP64 0           /* double everywhere                              */
#define WF_MIXED 1          /* pair arithmetic in float; integration, return
                               map and global reductions in double           */

/* integration schemes */
#define WF_LEAPFROG 0
#define WF_KICKDRIFT 1

/* material models (Source/Particle.h:34-37) */
#define WF_BILINEAR 0
#define WF_HOLLOMON 1
#define WF_JOHNSON_COOK 2

/* double fields (wf_set_field / wf_get_field); width per particle in [] */
#define WF_F_X 0            /* [3] position            Particle::x          */
#define WF_F_V 1            /* [3] velocity            Particle::v          */
#define WF_F_A 2            /* [3] acceleration        Particle::a          */
#define WF_F_DISP 3         /* [3] Particle::Displacement                     */
#define WF_F_DENSITY 4      /* [1] Particle::Density                          */
#define WF_F_MASS 5         /* [1] Particle::Mass                             */
#define WF_F_H 6            /* [1] Particle::h                                */
#define WF_F_SIGMA 7        /* [9] Particle::Sigma                            */
#define WF_F_SHEAR 8        /* [9] Particle::ShearStress                      */
#define WF_F_STRAIN_RATE 9  /* [9] Particle::StrainRate                       */
#define WF_F_ROT_RATE 10    /* [9] Particle::RotationRate                     */
#define WF_F_STRAIN 11      /* [9] Particle::Strain                           */
#define WF_F_STRAIN_PL 12   /* [9] Particle::Strain_pl                        */
#define WF_F_PL_STRAIN 13   /* [1] Particle::pl_strain                        */
#define WF_F_DPL_STRAIN 14  /* [1] Particle::delta_pl_strain                  */
#define WF_F_SIGMAY 15      /* [1] Particle::Sigmay                           */
#define WF_F_EFF_SR 16      /* [1] Particle::eff_strain_rate                  */
#define WF_F_PRESSURE 17    /* [1] Particle::Pressure                         */
#define WF_F_T 18           /* [1] Particle::T                                */
#define WF_F_DTDT 19        /* [1] Particle::dTdt                             */
#define WF_F_DDENSITY 20    /* [1] Particle::dDensity                         */
#define WF_F_Q_PLHEAT 21    /* [1] Particle::q_plheat                         */
#define WF_F_STRAIN_PL_INCR 22 /* [9] Particle::Strain_pl_incr                */
#define WF_F_Q_CONV 23      /* [1] Particle::q_conv                           */
#define WF_F_H_CONV 24      /* [1] Particle::h_conv                           */
#define WF_F_T_INF 25       /* [1] Particle::T_inf                            */
#define WF_F_Q_SOURCE 26    /* [1] Particle::q_source                         */
#define WF_F_PS_ENERGY 27   /* [1] Particle::ps_energy                        */
#define WF_F_CONTFORCE 28   /* [3] read-only: Particle::contforce of the last
                               step's contact (Source/Contact_Wang.cpp)      */
#define WF_F_GRADCORR 29    /* [9] read-only: Particle::gradCorrM (identity
                               unless wf_config.grad_corr)                  */
#define WF_F_DAMAGE 30      /* [1] Particle::dam_D (damage, 0..1)             */
#define WF_F_ZWAB 31        /* [1] read-only: Particle::ZWab of the last
                               wf_calc_force2233 (0 before it)              */
#define WF_F_VXSPH 32       /* [3] read-only: Particle::VXSPH, likewise        */
#define WF_F_SUMDEN 33      /* [1] read-only: Particle::SumDen, likewise       */
#define WF_NFIELDS 34

/* rigid-surface particle fields (wf_get_rigid_field), 3 doubles each, in
 * the order the planes were added, element order within a plane */
#define WF_R_X 0            /* particle position (moves with the plane)      */
#define WF_R_V 1            /* particle velocity                             */
#define WF_R_NORMAL 2       /* element normal                                */

/* int fields */
#define WF_I_ID 0           /* Particle::ID (zone id used by BCs)            */
#define WF_I_MATERIAL 1     /* index into the wf_set_material table; pairs
                               form only between particles of one index,
                               as Particle::Material gates SMPairs          */
#define WF_I_FREE 2         /* Particle::IsFree (only free-free pairs)        */
#define WF_I_THERMAL_BC 3   /* Particle::Thermal_BC (1 = convection)          */
#define WF_I_NB 4           /* read-only: Particle::Nb (neighbour count)      */
#define WF_I_CONTNB 5       /* read-only: Particle::ContNb (contact pairs of
                               the last contact search; 0 without contact)  */
#define WF_NIFIELDS 6

/* single-pass evaluations (wf_eval) */
#define WF_EVAL_ACCEL 0     /* a   <- sum over pairs, from x, v, Density, Sigma */
#define WF_EVAL_DENSITY 1   /* dDensity <- from x, v, Density                  */
#define WF_EVAL_RATES 2     /* StrainRate, RotationRate <- from x, v, Density  */
#define WF_EVAL_STRESS 3    /* CalcStressStrain(dt) of every particle using the
                               current StrainRate / RotationRate               */

typedef struct wf_config {
  int32_t precision;        /* WF_FP64 or WF_MIXED                             */
  int32_t ts_nb_inc;        /* Domain::ts_nb_inc (neighbour search period, 5)  */
  int32_t max_nb;           /* neighbour capacity per particle (1..128)       */
  int32_t thermal_solver;   /* Domain::thermal_solver                          */
  int32_t pl_heating;       /* Domain::pl_heating                              */
  int32_t device;           /* HIP device ordinal                              */
  int32_t no_graphs;        /* 1: launch kernels one by one instead of replaying
                               a captured HIP graph per step                  */
  int32_t stage_max;        /* WF_MIXED: neighbours of each 128-particle tile are
                               staged in LDS when the tile's neighbourhood has
                               at most this many particles (0: 896, the most;
                               -1: never, gather from HBM).  Results do not
                               depend on it.                                  */
  int32_t lanes;            /* threads per particle in the pair passes: 1, 2
                               or 4; each takes every lanes-th batch of 8
                               neighbours and the partial sums are added in
                               lane order (0: the default for the precision).
                               Sums, and so results, depend on it at the
                               rounding level.                                */
  int32_t auto_ts;          /* adaptive time step (Domain::AdaptiveTimeStep):
                               bit 0 velocity criterion (auto_ts), bit 1
                               acceleration criterion (auto_ts_acc), bit 2
                               the contact criterion (auto_ts_cont: the step
                               drops to Domain::min_force_ts, the least
                               2 m |v| / |f_n| of the contact pairs, Source/
                               Domain.cpp:222-229, Source/Contact_Wang.cpp:
                               226-233; it acts with bit 0 or 1); 0: fixed
                               dt.  Leapfrog and KickDrift.                    */
  double cellfac;           /* Domain::Cellfac (2 for the cubic spline)       */
  double pl_work_heat_frac; /* Domain::pl_work_heat_frac (0.9)                */
  double dom_max[3];        /* Domain::DomMax (default -1e11)                  */
  double dom_min[3];        /* Domain::DomMin (default +1e11)                  */
  double cfl;               /* Domain::CFL (0.7) for the velocity criterion    */
  double sqrt_h_a;          /* Domain::sqrt_h_a (0.0025), acceleration one     */
  double min_ts;            /* Domain::manual_min_ts (1e-10)                   */
  int32_t grad_corr;        /* Domain::gradKernelCorr: Bonet kernel-gradient
                               correction (CalcGradCorrMatrix, Source/Domain.cpp:
                               1832-1940) in the three pair passes; all
                               particles must be free                         */
  int32_t locking;          /* !Domain::nonlock_sum: the Locking sums, whose
                               rate tensors weigh pairs by m_j/rho_j rounded
                               to float (Source/InteractionAlt.cpp:412,425).
                               The reference switches to them when
                               gradKernelCorr is set; so does the facade.     */
} wf_config;

typedef struct wf_material {
  int32_t model;            /* WF_BILINEAR / WF_HOLLOMON / WF_JOHNSON_COOK     */
  int32_t reserved;
  double E, nu;             /* Elastic_ (Source/Material.h:64-74)              */
  double G;                 /* Particle::G                                     */
  double Cs;                /* Particle::Cs                                    */
  double ref_density;       /* Particle::RefDensity                            */
  double P0;                /* Particle::P0                                    */
  double alpha, beta;       /* Particle::Alpha / Beta (artificial viscosity)   */
  double Ep;                /* Particle::Ep (bilinear plastic modulus)         */
  double hol_K, hol_m, hol_sy0;                 /* Hollomon                     */
  double jc_A, jc_B, jc_n, jc_C, jc_eps0, jc_m, jc_Tm, jc_Tt; /* Johnson-Cook  */
  double k_T, cp_T, th_exp; /* Particle::k_T, cp_T, th_exp                     */
  double ti, ti_n, ti_dist; /* Particle::TI, TIn (4), TIInitDist: tensile
                               instability (Source/Domain.cpp:1529-1563,
                               Source/InteractionAlt.cpp:104-110); ti = 0 off.
                               Set before wf_init.                            */
} wf_material;

typedef struct wf_scalars {
  double time;              /* Domain::Time                                    */
  double dt;                /* Domain::deltat                                  */
  int64_t steps;
  int64_t nb_searches;      /* neighbour searches performed                    */
  int64_t npairs;           /* unordered free-free pairs (SMPairs size)        */
  int32_t max_nb_count;     /* largest per-particle neighbour count            */
  int32_t reserved;
  double kin_energy_sum;    /* Domain::kin_energy_sum                          */
  double int_energy_sum;    /* Domain::int_energy_sum                          */
  double plastic_work;      /* Domain::plastic_work                            */
  double max_pl_strain;     /* max over particles of pl_strain                 */
  double max_T, min_T;      /* Domain::m_maxT / m_minT                         */
  double contact_force_sum; /* Domain::contact_force_sum (last step)          */
  double contact_reaction_sum; /* Domain::contact_reaction_sum (last step)    */
  double ext_forces_work;   /* Domain::ext_forces_work                         */
  int64_t contact_pairs;    /* contact pairs of the last search (ContPairs)   */
  double contact_friction_work; /* Domain::contact_friction_work (friction heat) */
  double accum_cont_heat_cond;  /* Domain::accum_cont_heat_cond (conducted heat) */
  double min_force_ts;      /* Domain::min_force_ts (contact step criterion, auto_ts bit 2) */
} wf_scalars;

int wf_abi_version(void);
const char *wf_last_error(const void *ctx); /* ctx may be NULL (last global) */
int wf_device_count(int32_t *count);

/* life cycle */
int wf_create(const wf_config *cfg, int64_t n_particles, void **ctx);
int wf_destroy(void *ctx);

/* state */
int wf_set_material(void *ctx, int32_t index, const wf_material *m);
int wf_set_field(void *ctx, int32_t field, const double *host);
int wf_get_field(void *ctx, int32_t field, double *host);
int wf_set_int_field(void *ctx, int32_t field, const int32_t *host);
int wf_get_int_field(void *ctx, int32_t field, int32_t *host);

/* boundary conditions, applied where the reference calls GeneralAfter */
int wf_add_amplitude(void *ctx, int32_t id, int32_t npts, const double *time, const double *value);
int wf_add_velocity_bc(void *ctx, int32_t zone, int32_t value_type, const double value[3],
                       int32_t amp_id, double amp_factor);
int wf_add_temperature_bc(void *ctx, int32_t zone, double T);
/* boundaryCondition::value_ang (Source/Domain.h:98-111) of the velocity BCs
 * already added on `zone`: a constant BC on a rigid mesh's zone sets the
 * mesh's angular velocity m_w (SetRotAxisVel, Source/WeldForm.cpp:97-101),
 * and TriMesh::Update moves its nodes with m_v + m_w x node (rotation about
 * the origin).  Amplitude BCs leave m_w as it was. */
int wf_set_bc_angular_velocity(void *ctx, int32_t zone, const double w[3]);

/* rigid contact surfaces, before wf_init.  wf_add_rigid_plane adds the
 * plane mesh of TriMesh::AxisPlaneMesh (Source/Mesh.cpp:82-202): normal to
 * `axis`, from p1 spanning p2, dens x dens squares of side
 * (p2 - p1)[first in-plane axis] / dens split into two triangles, winding
 * and normal set by `flip` (the reference's flipNormals); its particles
 * (Domain::AddTrimeshParticles, Source/Contact.cpp:10-37) carry zone id
 * `zone` and h = hfac * element radius.  A velocity BC on `zone` moves the
 * plane.  Returns the number of triangles, or a negative WF_E* code.
 * wf_set_contact sets Domain::friction_sta, friction_dyn and DFAC of the
 * Wang penalty contact (Source/Contact_Wang.cpp:15-348), before or after
 * wf_init (the next step uses the new values). */
int wf_add_rigid_plane(void *ctx, int32_t zone, int32_t axis, int32_t flip, const double p1[3], const double p2[3],
                       int32_t dens, double hfac);
/* wf_add_rigid_mesh adds a rigid surface of nelem triangles over nnode
 * nodes (3 doubles each; elems: 3 0-based node indices per triangle, the
 * winding the one the mesh update recomputes normals from), as
 * TriMesh(NastranReader &, flip) builds it from a mesh file
 * (Source/Mesh.cpp:17-64) after the caller has applied flipNormals (first
 * two nodes swapped), scaleFactor and translation: centroids, the initial
 * element normals (3 per triangle; NULL: from the winding -- the reference
 * takes them from the file's unflipped winding until they are recomputed),
 * CalcSpheres, and particles with h = hfac * element radius.  Planes and
 * meshes share the limit of 4 surfaces.  Returns the number of triangles. */
int wf_add_rigid_mesh(void *ctx, int32_t zone, int64_t nnode, const double *nodes, int64_t nelem,
                      const int32_t *elems, const double *normals, double hfac);
int wf_set_contact(void *ctx, double friction_sta, double friction_dyn, double damping);
/* Coulomb friction of material `index`'s particles against the rigid
 * surfaces, in place of wf_set_contact's for them (no counterpart in the
 * reference, whose friction_sta/dyn belong to the Domain: a batch of
 * specimens in one context gives each specimen its own, as its own Domain
 * would).  May be called before or after wf_init; the adjoint refuses
 * contexts that use it. */
int wf_set_material_friction(void *ctx, int32_t index, double friction_sta, double friction_dyn);
/* wf_set_contact_law <- Domain::contact_alg / PFAC / fric_type
 * (Source/Domain.h:64,298,369; dispatched by SolveDiffUpdateLeapFrog,
 * Source/SolverLeapfrog.cpp:234-239; scene key "contAlgorithm",
 * Source/WeldForm.cpp:695-705): law 0 Fraser (Domain::CalcContactForces,
 * Source/Contact.cpp:223-554), 1 Wang (the default), 2 Seo
 * (CalcContactForces2, Source/Contact.cpp:562-770), 4 LSDyna
 * (CalcContactForcesLS, Source/Contact_LS.cpp:15-292); pfac > 0 (default
 * 0.5) scales the penalty of the Fraser, Seo and LS-Dyna laws (Fraser and
 * Seo: PFAC cont_stiff, with Particle::cont_stiff formed at wf_init from
 * the initial state as SolveDiffUpdateKickDrift forms it, Source/
 * SolverKickDrift.cpp:79-84; LS-Dyna: PFAC 2 m / dt^2); fric_bound selects
 * fric_type Fr_Bound in Fraser's law.  Under Leapfrog these three laws
 * leave a unchanged (their contforce and sums are outputs) except Fraser's
 * Fr_Bound; KickDrift always uses Wang's.  Before wf_init.
 * wf_set_friction_function <- Domain::friction_function Linear with
 * friction_m, friction_b (Source/Domain.h:293-295): Wang's and LS-Dyna's
 * laws take friction_sta = friction_dyn = m T + b of the contacting
 * particle (Source/Contact_Wang.cpp:241-244, Source/Contact_LS.cpp:217-220);
 * linear 0 restores the constant coefficients.  Before wf_init. */
int wf_set_contact_law(void *ctx, int32_t law, double pfac, int32_t fric_bound);
/* wf_set_dimension <- Domain::Dimension (Source/Domain.h:331): 3, or 2 for
 * plane strain, every particle in one z-plane.  In 2D the cubic spline
 * takes its 2D constants (Source/Functions.cpp:32,81), StartAcceleration
 * zeroes each z velocity at the top of the step (Source/Domain.cpp:1592),
 * the pair forces lose their z part (Source/InteractionAlt.cpp:184,
 * Source/Interaction.cpp:237), the cell grid has one z layer
 * (Source/Neighbour.cpp:56-75), the free surface takes at most 12
 * neighbours (Source/Domain.cpp:1474), the kernel-gradient correction sets
 * the sum's zz entry to 1 (Source/Domain.cpp:1911-1915) and convection uses
 * (m/rho)^(1/3) (Source/Thermal.cpp:354-366).  Rigid contact, the
 * tensile-instability term, the adjoint and the axisymmetric branches are
 * refused in 2D (WF_EUNSUPPORTED).  Before wf_init. */
int wf_set_dimension(void *ctx, int32_t dim);
/* wf_set_material_damage <- Material_::damage = new JohnsonCookDamage(D1,
 * D2, D3, D4, D5, eps_0) and Domain::model_damage (Source/Material.h:47-62,
 * scene keys "damageModel": "JohnsonCook" and "damageParams",
 * Source/WeldForm.cpp:293-321): model 1 Johnson-Cook, D[5] = D1..D5; model
 * 0 none.  With damage on any material the Leapfrog step adds
 * Domain::CalcDamage after the stress update (Source/Damage.cpp:9-46):
 * dam_D += delta_pl_strain / eps_f(delta_pl_strain / dt, -P / Sigma_eq, T)
 * with eps_f = (D1 + D2 exp(D3 s)) (1 + D4 ln(rate / eps_0)) (1 + D5 T_h)
 * (Source/Material.cpp:76-92), capped at 1; the acceleration and density
 * passes weight each pair by 1 - (D_i + D_j) / 2 (Source/InteractionAlt.cpp:
 * 54-60,558-572); the free surface is refreshed every step; the neighbour
 * list of the first step is kept (Source/SolverLeapfrog.cpp:160-206).  The
 * material must be Johnson-Cook (its T_t, T_m, eps_0 enter eps_f), the
 * context FP64 with the Locking sums (wf_config.locking 1, which the scene
 * driver sets with damage), and only WF_LEAPFROG steps (KickDrift has no
 * damage).  Before wf_init. */
int wf_set_material_damage(void *ctx, int32_t index, int32_t model, const double D[5]);
int wf_set_friction_function(void *ctx, int32_t linear, double m, double b);
/* wf_set_contact_heat <- Domain::cont_heat_cond / contact_hc / cont_heat_fric
 * (Source/Domain.h:339-347, set from a scene's Contact "heatConductance",
 * "heatCondCoeff" and "heatFriction", Source/WeldForm.cpp:711-723): every
 * contact pair sets q_cont_conv = rho h_c dS2 (T_rigid - T) / m with
 * dS2 = (m / rho)^(2/3) (Source/Contact_Wang.cpp:320-324), and in the
 * sliding branch q_fric_work = |f_t . v_r| rho / m, friction_hfl and
 * cshearabs (:291-300); the thermal step adds both to dT/dt
 * (Source/Thermal.cpp:165-189).  Before wf_init.
 * wf_set_rigid_temperature <- the T of the rigid particles (set by an
 * initial condition or temperature BC on the mesh's zone, Source/WeldForm.cpp:
 * 80-86, 850-866; the thermal step leaves them unchanged): count = the rigid
 * particles of the meshes added, in mesh order; 0 unless set.  Before
 * wf_init. */
int wf_set_contact_heat(void *ctx, int32_t conduction, int32_t friction, double hc);
int wf_set_rigid_temperature(void *ctx, int64_t count, const double *T);
/* contact force on a plane (Domain::m_contact_force[mesh]) per step; last
 * min(cap, steps) entries */
int wf_get_contact_force_history(void *ctx, int32_t mesh, double *out3, int64_t cap, int64_t *count);
int wf_get_rigid_field(void *ctx, int32_t field, double *host);

/* solve */
int wf_init(void *ctx, double dt);
int wf_step(void *ctx, int32_t scheme, int64_t nsteps); /* enqueues; see wf_sync */
/* set-up of the step loop of `scheme` (WF_LEAPFROG / WF_KICKDRIFT) without
 * running a step: captures and instantiates the step's HIP graphs (one per
 * position-buffer parity), which wf_step otherwise does on its first two
 * calls.  No reference counterpart (the reference has no set-up of its
 * loop beyond SolveDiffUpdateLeapFrog's own first lines,
 * Source/SolverLeapfrog.cpp:24-123); optional. */
int wf_prepare(void *ctx, int32_t scheme);
int wf_sync(void *ctx);                                 /* waits, reports device errors */
/* one part of a step, for a caller that runs host hooks between the parts:
 * the reference's GeneralBefore / GeneralAfter (Source/Domain.h:398-399),
 * called at Source/SolverLeapfrog.cpp:215,228,254,273 and
 * Source/SolverKickDrift.cpp:196,208,231,267.  Parts 0..4 in order per step:
 *   WF_PART_SEARCH  neighbour schedule and search            -> GeneralBefore
 *   WF_PART_ACCEL   CalcAccel + AccelReduction, then the reaction observer
 *                   (before any hook changes a)              -> GeneralAfter
 *   WF_PART_KICK1   contact; Leapfrog: the first-step half kick;
 *                   KickDrift: the time step, v += a dt_prev/2 -> GeneralAfter
 *   WF_PART_KICK2   Leapfrog: the time step, v += a (dt_prev + dt)/2;
 *                   KickDrift: density, drift, v += a dt/2   -> GeneralAfter
 *   WF_PART_REST    the rest of the step
 * Each call waits for its work, so fields a hook changes with wf_set_field
 * (v, a, T, ...) enter the parts after it.  The velocity and temperature BCs
 * of wf_add_*_bc still apply in the kick parts, as UserAcc would.  wf_step
 * is refused while a step is part-way. */
#define WF_PART_SEARCH 0
#define WF_PART_ACCEL 1
#define WF_PART_KICK1 2
#define WF_PART_KICK2 3
#define WF_PART_REST 4
int wf_step_part(void *ctx, int32_t scheme, int32_t part);
int wf_get_scalars(void *ctx, wf_scalars *out);
/* reaction sum(m*a) of every particle of a velocity-BC zone, taken where
 * the reference first calls GeneralAfter in a step (after AccelReduction,
 * before the BC zeroes a).  Copies the last min(cap, steps) entries. */
int wf_get_reaction_history(void *ctx, int32_t zone, double *out3, int64_t cap, int64_t *count);
/* Per-group sums (no counterpart in the reference; for a batch of specimens
 * in one context, one group per material index): call before wf_init.
 * Every step then keeps, per group, the reaction of velocity-BC zone `zone`
 * (sum of m a over the group's particles carrying that id, observed where
 * wf_get_reaction_history's is; zone -1: none) and, with rigid meshes, the
 * contact force of each mesh on the group (the m_contact_force split by
 * group), for the last `cap` steps.  Each sum is taken in a fixed order.
 * wf_get_group_history: what 0 the reaction (mesh ignored), 1 the contact
 * force of `mesh`; out3 receives the last min(steps, cap) triples, oldest
 * first. */
int wf_set_group_sums(void *ctx, int32_t zone, int64_t cap);
int wf_get_group_history(void *ctx, int32_t what, int32_t group, int32_t mesh, double *out3, int64_t cap,
                         int64_t *count);

/* neighbours */
int wf_rebuild_neighbours(void *ctx);
int wf_get_neighbour_pairs(void *ctx, int32_t *pairs2, int64_t cap, int64_t *count);

/* single passes for parity checks (no time advance) */
int wf_eval(void *ctx, int32_t what, double dt);
/* wf_calc_force2233 <- Domain::LastComputeAcceleration's pair loop, i.e.
 * Domain::CalcForce2233 over every pair (Source/Domain.cpp:1700-1709,
 * Source/Interaction.cpp:25-333): one fused FP64 pass over each particle's
 * row at the current state (after wf_rebuild_neighbours), no time advance.
 * It sets a, dDensity, StrainRate and RotationRate (the Locking form, float
 * m_j/rho_j weights in the rates), and WF_F_ZWAB, WF_F_VXSPH (with the XSPH
 * coefficient `xsph`) and, when `shepard` != 0, WF_F_SUMDEN (the Shepard
 * sum due for every particle).  GradientType 0; refused with the
 * tensile-instability term.  Leapfrog and KickDrift do not call it. */
int wf_calc_force2233(void *ctx, double xsph, int32_t shepard);

/* kernel timing with HIP events on the context's stream.
 * kernel: 0 accel, 1 density+drift, 2 rates+stress, 3 neighbour build,
 *         4 whole step, 5 after-search set-up (kernel-gradient correction,
 *         free surface and contact partners).  avg_ms is the mean duration
 *         per launch. */
int wf_set_timing(void *ctx, int32_t enable);
int wf_get_kernel_time(void *ctx, int32_t kernel, double *avg_ms, int64_t *launches);
int wf_get_neighbour_stats(void *ctx, int64_t *directed_pairs, int32_t *max_count);
/* LDS staging of the mixed-precision pair passes after the last neighbour
 * build (no counterpart in the reference: a diagnostic of this library's
 * tiles of consecutive particles): the number of tiles, how many of them
 * gather their neighbours from LDS (their staged set fits; the others gather
 * from HBM) and the total size of the staged sets.  FP64 contexts report no
 * staged tiles. */
int wf_get_stage_stats(void *ctx, int64_t *tiles, int64_t *staged_tiles, int64_t *staged_values);
/* cells per axis of the reference grid (CellInitiate, Source/Neighbour.cpp:
 * 10-139: sized by h_max over all particles, rigid ones included; the
 * contact partners and the thermal pair orientation use it) and of the grid
 * the SPH neighbour search runs on (cells at least Cellfac x the largest SPH
 * h wide, finer than the reference's when rigid particles are larger; the
 * same pairs are found).  Diagnostic; no counterpart in the reference. */
int wf_get_cell_grids(void *ctx, int32_t reference[3], int32_t search[3]);

/* output tables, packed on the device in the caller's particle order and
 * copied to the host in one transfer (the first call allocates a device
 * staging buffer of WF_XDMF_NF * 4 + WF_XDMF_NI * 4 bytes per particle,
 * kept until wf_destroy).
 *
 * wf_pack_xdmf <- the dataset packing of Domain::WriteXDMF
 * (Source/Output.cpp:83-215): f receives WF_XDMF_NF * n floats, dataset
 * after dataset, dataset D of width w at f[D * n], particle i's values at
 * f[D * n + w * i ...]; iv receives WF_XDMF_NI * n ints the same way.
 * Tensors of width 6 are (00, 11, 22, 01, 12, 02).  As in the reference,
 * ShearS holds Sigma (Source/Output.cpp:250-251), Property1..3 hold the
 * default UserOutput's 0, Sigma_eq is CalculateEquivalentStress of
 * ShearStress.  gradcorrmat is gradCorrM with wf_config.grad_corr, else 0;
 * q_friction (friction_hfl) and c_shearabs are those of the last contact
 * evaluation, 0 without wf_set_contact_heat's friction heating. */
#define WF_XDMF_POSITION 0        /* width 3 */
#define WF_XDMF_VELOCITY 3        /* width 3 */
#define WF_XDMF_ACCELERATION 6    /* width 3 */
#define WF_XDMF_PRESSURE 9
#define WF_XDMF_DENSITY 10
#define WF_XDMF_PROPERTY1 11
#define WF_XDMF_PROPERTY2 12
#define WF_XDMF_PROPERTY3 13
#define WF_XDMF_MASS 14
#define WF_XDMF_H 15
#define WF_XDMF_SIGMA 16          /* width 6 */
#define WF_XDMF_SHEARS 22         /* width 6 */
#define WF_XDMF_STRAIN 28         /* width 6 */
#define WF_XDMF_STRAIN_RATE 34    /* width 6 */
#define WF_XDMF_STRAIN_PL 40      /* width 6 */
#define WF_XDMF_GRADCORRMAT 46    /* width 6 */
#define WF_XDMF_SIGMA_EQ 52
#define WF_XDMF_TEMPERATURE 53
#define WF_XDMF_PL_STRAIN 54
#define WF_XDMF_Q_FRICTION 55
#define WF_XDMF_C_SHEARABS 56
#define WF_XDMF_PS_EN 57
#define WF_XDMF_DISPLACEMENT 58   /* width 3 */
#define WF_XDMF_CONTACT_FORCE 61  /* width 3 */
#define WF_XDMF_NF 64
#define WF_XDMF_TAG 0             /* int datasets */
#define WF_XDMF_NEIGHBORS 1
#define WF_XDMF_CONTNEIB 2
#define WF_XDMF_NI 3
int wf_pack_xdmf(void *ctx, float *f, int32_t *iv);

/* wf_pack_xdmf_noalloc: no counterpart in the reference.  wf_pack_xdmf for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc): the same tables,
 * bit for bit, staged in a static device array of 64 MiB per GPU instead
 * of a staging buffer allocated on the first call.  The array is shared by
 * every call on that GPU and kept until the process ends, so calls that
 * overlap on one GPU (from several threads or contexts) race.  Contexts of
 * more than wf_pack_xdmf_noalloc_max() particles (250,406: 268 bytes each)
 * get WF_EINVAL.  No hipMalloc / hipFree is called. */
int wf_pack_xdmf_noalloc(void *ctx, float *f, int32_t *iv);
int64_t wf_pack_xdmf_noalloc_max(void);

/* wf_pack_csv <- the rows of Domain::WriteCSV (Source/Output.cpp:521-554):
 * rows receives WF_CSV_NCOL doubles per particle, particle after particle:
 * X, Y, Z, ID, Sigma_eq, Pl_Strain, vx, vy, vz, ax, ay, az, CFx, CFy, CFz,
 * p, Temp. */
#define WF_CSV_NCOL 17
int wf_pack_csv(void *ctx, double *rows);

/* adjoint (no counterpart in the reference).  A window of cfg->steps FP64
 * Leapfrog steps from the current state; the loss is
 *     L = sum_n (F_n - target_n)^2
 * with F_n the component cfg->loss_component of, per step n, the contact
 * force on rigid plane cfg->loss_index (WF_ADJ_PLANE, Domain::
 * m_contact_force) or the reaction of velocity-BC zone cfg->loss_index
 * (WF_ADJ_ZONE, as wf_get_reaction_history).  wf_adj_forward runs the window,
 * keeping a snapshot of the device state every cfg->checkpoint_every steps,
 * and returns L (and F_n in history[steps] when not NULL).  wf_adj_backward
 * replays each segment from its snapshot with a tape and runs the reverse
 * step kernels over it, returning dL/d(A, B, n, C, m) of the Johnson-Cook
 * material cfg->material and dL/d(friction_sta, friction_dyn) in grad[7]
 * (friction_sta enters only the static/dynamic switch: its derivative is 0).
 * Branches (yield, contact activity, friction regime, neighbour lists) are
 * those of the forward.  After it the state is the window's end state;
 * wf_adj_get_field gives dL/d(initial state) for WF_F_X, WF_F_V,
 * WF_F_DENSITY, WF_F_SHEAR (9 per particle), WF_F_SIGMAY, WF_F_PL_STRAIN,
 * and wf_adj_restore puts the whole device state back to the window's start
 * (to run the window again, e.g. with other material parameters; material
 * parameters set since the window started revert too, so set them after
 * the restore).  Needs precision
 * WF_FP64, a fixed step (auto_ts 0) and thermal_solver 0. */
#define WF_ADJ_PLANE 0
#define WF_ADJ_ZONE 1
#define WF_ADJ_NPARAM 7
typedef struct {
  int64_t steps;              /* window length                                 */
  int32_t checkpoint_every;   /* snapshot interval k (tape of k steps)         */
  int32_t loss_kind;          /* WF_ADJ_PLANE or WF_ADJ_ZONE                   */
  int32_t loss_index;         /* plane (order of wf_add_rigid_plane) or zone id */
  int32_t loss_component;     /* 0, 1, 2                                       */
  int32_t material;           /* material index of the Johnson-Cook gradients  */
  int32_t reserved;
} wf_adj_config;
int wf_adj_forward(void *ctx, const wf_adj_config *cfg, const double *target, double *loss, double *history);
int wf_adj_backward(void *ctx, double *grad);
int wf_adj_get_field(void *ctx, int32_t field, double *host);
int wf_adj_restore(void *ctx);

/* host-only geometry helper: AddCylinderLength lattice (3 doubles per
 * particle).  Returns the particle count; writes at most cap particles. */
int64_t wf_gen_cylinder(const double V[3], double Rxy, double Lz, double r, double *out, int64_t cap);

#ifdef __cplusplus
}
#endif
#endif
