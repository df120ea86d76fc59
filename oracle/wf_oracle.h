/*
 * wf_oracle.h -- CPU oracle for the explicit SPH step of WeldForm.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP path is
 * compared against, and the CPU baseline timed next to it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product library (weldform_amd) never links or calls it.
 *
 * It restates, in plain C + OpenMP, the reference's Leapfrog / KickDrift
 * path with Nproc=1 semantics ("Nishimura" pair order), following:
 *   Source/SolverLeapfrog.cpp:24-451, Source/SolverKickDrift.cpp:23-401,
 *   Source/Neighbour.cpp:10-694, Source/InteractionAlt.cpp:12-640,
 *   Source/Functions.cpp:25-124,206-256, Source/Particle.cpp:28-157,922-1067,
 *   Source/Material.cpp:6-72, Source/Domain.cpp:78-175,674-860,1519-1597,
 *   1752-1829,1832-1940 (kernel-gradient correction; with the Locking
 *   branches of Source/InteractionAlt.cpp), Source/Thermal.cpp:53-219,339-389,721-748,
 *   Source/ThermalStuct.cpp:7-39, Source/WeldForm.cpp:51-120,338-369,884-921.
 * Each function in wf_oracle.c cites the lines it follows.
 *
 * Parallel mode (omp threads > 1) keeps the Nproc=1 pair order and the
 * per-particle summation order, so results are bitwise those of 1 thread.
 */
#ifndef WF_ORACLE_H
#define WF_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define WFO_MAX_NB 100          /* MAX_NB_PER_PART, Source/Domain.h:49 */
#define WFO_MAX_ZONES 64
#define WFO_MAX_BCS 32
#define WFO_MAX_AMPS 16
#define WFO_MAX_AMP_PTS 64
#define WFO_MAX_MESH 8

enum { WFO_BILINEAR = 0, WFO_HOLLOMON = 1, WFO_JOHNSON_COOK = 2 };

/* double fields (per particle; n, 3n or 9n values, row-major 3x3) */
enum {
  WFO_X = 0, WFO_V, WFO_A, WFO_XPREV, WFO_DISP,            /* 3 */
  WFO_RHO, WFO_DRHO, WFO_MASS, WFO_H, WFO_REFRHO, WFO_P,   /* 1 */
  WFO_CS, WFO_P0, WFO_G, WFO_ALPHA, WFO_BETA,              /* 1 */
  WFO_SIGMA, WFO_SHEAR, WFO_SR, WFO_RR, WFO_STRAIN,        /* 9 */
  WFO_EPL, WFO_EPLINC,                                     /* 9 */
  WFO_PLSTRAIN, WFO_DPLSTRAIN, WFO_SIGMAY, WFO_EFFSR,      /* 1 */
  WFO_EP, WFO_ET, WFO_T, WFO_DTDT, WFO_KT, WFO_CPT,        /* 1 */
  WFO_THEXP, WFO_QCONV, WFO_QSRC, WFO_QPL, WFO_QCONT,      /* 1 */
  WFO_HCONV, WFO_TINF, WFO_PSENERGY,                       /* 1 */
  WFO_NORMAL,                                              /* 3 */
  WFO_CONTFORCE,                                           /* 3 */
  WFO_TI, WFO_TIN, WFO_TIDIST,                             /* 1: TI, TIn, TIInitDist */
  WFO_TIR,                                                 /* 9 */
  WFO_GCM,                                                 /* 9: gradCorrM */
  WFO_QFRIC, WFO_CSHEAR, WFO_FHFL,                         /* 1: q_fric_work, cshearabs, friction_hfl */
  WFO_DAMD,                                                /* 1: dam_D */
  WFO_ZWAB, WFO_VXSPH, WFO_SUMDEN,                         /* 1, 3, 1: ZWab, VXSPH, SumDen */
  WFO_NFIELDS
};
/* int fields */
enum { WFO_ID = 0, WFO_MATERIAL, WFO_MATIDX, WFO_MODEL, WFO_FREE, WFO_THBC,
       WFO_NB, WFO_ELEMENT, WFO_MESHIDX, WFO_IDORIG, WFO_NIFIELDS };

typedef struct wfo_domain wfo_domain;

typedef struct {
  int model;                    /* WFO_BILINEAR / HOLLOMON / JOHNSON_COOK */
  double E, nu;
  double K, m, sy0;             /* Hollomon (Source/Material.cpp:48-57)   */
  double A, B, n, C, eps_0, jm, Tm, Tt;   /* Johnson-Cook                */
} wfo_material;

wfo_domain *wfo_create(int n);
void wfo_destroy(wfo_domain *d);
int wfo_n(const wfo_domain *d);
double *wfo_f64(wfo_domain *d, int field);
int *wfo_i32(wfo_domain *d, int field);
int wfo_set_param(wfo_domain *d, const char *name, double v);
double wfo_get_param(const wfo_domain *d, const char *name);
int wfo_set_material(wfo_domain *d, int idx, const wfo_material *m);
int wfo_set_material_friction(wfo_domain *d, int idx, double sta, double dyn);
/* material idx's damage model: 0 none, 1 JohnsonCookDamage with D1..D5 */
int wfo_set_material_damage(wfo_domain *d, int idx, int model, const double D[5]);
void wfo_set_threads(int n);
const char *wfo_last_error(void);

/* boundary conditions (Source/Domain.h:69-111, Source/WeldForm.cpp:51-92) */
int wfo_add_velocity_bc(wfo_domain *d, int zone, int value_type,
                        const double value[3], int amp_id, double amp_factor);
int wfo_set_bc_angular(wfo_domain *d, int zone, const double w[3]);
int wfo_add_temperature_bc(wfo_domain *d, int zone, double T);
int wfo_add_amplitude(wfo_domain *d, int id, int npts, const double *t,
                      const double *v);
int wfo_assign_zone(wfo_domain *d, const double start[3], const double end[3],
                    int id);

/* geometry: AddCylinderLength (Source/Domain.cpp:708-860).  Writes up to
 * cap positions (3 per particle), returns the count (or the count needed
 * when out == NULL). */
int64_t wfo_gen_cylinder(const double V[3], double Rxy, double Lz, double r,
                         double *out, int64_t cap);

/* solver pieces */
int wfo_solve_init(wfo_domain *d, double dt);
int wfo_step_leapfrog(wfo_domain *d);
int wfo_step_kickdrift(wfo_domain *d);
int wfo_neighbour_search(wfo_domain *d);          /* Clear + Main + tables */
int64_t wfo_npairs(const wfo_domain *d);
const int *wfo_pairs(const wfo_domain *d);         /* (temp1,temp2) pairs  */
int wfo_start_acceleration(wfo_domain *d);
int wfo_calc_grad_corr(wfo_domain *d);   /* CalcGradCorrMatrix: gradCorrM of every particle */
int wfo_calc_accel(wfo_domain *d);                 /* CalcAccel+Reduction  */
int wfo_calc_dens(wfo_domain *d);                  /* CalcDensInc+Reduction*/
int wfo_calc_rates(wfo_domain *d);                 /* CalcRateTensors+Red. */
/* LastComputeAcceleration's pair loop: CalcForce2233 over the pairs */
int wfo_calc_force2233(wfo_domain *d, double xsph, int shepard);
int wfo_stress_all(wfo_domain *d, double dt);      /* CalcStressStrain     */
int wfo_thermal_calcs(wfo_domain *d, double dt);   /* ThermalCalcs         */
int wfo_general_after(wfo_domain *d);              /* UserAcc BC hook      */
int wfo_plastic_work_heat(wfo_domain *d, double dt); /* CalcPlasticWorkHeat */
int wfo_energies(wfo_domain *d);                   /* Kin + Int EnergyEqn  */
/* rigid plane meshes and contact (Source/Mesh.cpp, Source/Contact*.cpp):
 * the mesh's particles are written at [first, first + 2 dens^2) */
int wfo_add_mesh(wfo_domain *d, int first, int id, int nnodes, const double *nodes, int nelem, const int *el,
                 const double *normals, double hfac);    /* general triangle mesh (TriMesh from a mesh file) */
int wfo_add_plane_mesh(wfo_domain *d, int first, int id, int axis, int flip, const double p1[3],
                       const double p2[3], int dens, double hfac);
int wfo_contact_force(const wfo_domain *d, int mesh, double out[3]);
int wfo_group_contact_force(const wfo_domain *d, int group, int mesh, double out[3]);
int64_t wfo_ncontact_pairs(const wfo_domain *d);
const int *wfo_contact_pairs(const wfo_domain *d);          /* (SPH, rigid) */
/* reaction observed at the first BC call of the last step:
 * sum of m*a over particles with ID == zone (before BCs zero a) */
int wfo_reaction(const wfo_domain *d, int zone, double out[3]);
/* kernel functions, exported for closed-form pins (3D; the _dim forms take
 * Domain::Dimension, 2 or 3) */
double wfo_kernel(double q, double h);
double wfo_kernel_dim(double q, double h, int dim);
double wfo_grad_kernel_dim(double q, double h, int dim);
/* symmetric 3x3 eigen-decomposition (cyclic Jacobi): A = V diag(w) V^T */
void wfo_sym3_eig(const double *A, double *w, double *V);
double wfo_grad_kernel(double q, double h);

#ifdef __cplusplus
}
#endif
#endif
