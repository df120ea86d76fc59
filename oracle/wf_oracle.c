/*
 * wf_oracle.c -- CPU oracle (TEST INFRASTRUCTURE, see wf_oracle.h).
 *
 * Plain C restatement of the reference's explicit SPH step with Nproc=1
 * pair order.  Arithmetic expressions keep the reference's operand order
 * (the 3x3 helpers of External/matvec.h:549-850 are element-wise, so every
 * matrix expression below is written per element in the same order).
 * Build with -ffp-contract=off so each line rounds as written.
 */
#include "wf_oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MIN_PS_FOR_NBSEARCH 1.e-6 /* Source/Domain.cpp:28 */

static char g_err[512];
static int g_threads = 1;
static int fail(const char *msg) {
  snprintf(g_err, sizeof g_err, "%s", msg);
  return -1;
}
const char *wfo_last_error(void) { return g_err; }
void wfo_set_threads(int n) { g_threads = n > 0 ? n : 1; }

static const int k_fdim[WFO_NFIELDS] = {
    3, 3, 3, 3, 3,        /* X V A XPREV DISP */
    1, 1, 1, 1, 1, 1,     /* RHO DRHO MASS H REFRHO P */
    1, 1, 1, 1, 1,        /* CS P0 G ALPHA BETA */
    9, 9, 9, 9, 9, 9, 9,  /* SIGMA SHEAR SR RR STRAIN EPL EPLINC */
    1, 1, 1, 1,           /* PLSTRAIN DPLSTRAIN SIGMAY EFFSR */
    1, 1, 1, 1, 1, 1,     /* EP ET T DTDT KT CPT */
    1, 1, 1, 1, 1,        /* THEXP QCONV QSRC QPL QCONT */
    1, 1, 1,              /* HCONV TINF PSENERGY */
    3,                    /* NORMAL */
    3,                    /* CONTFORCE */
    1, 1, 1,              /* TI TIN TIDIST */
    9,                    /* TIR */
    9,                    /* GCM */
    1, 1, 1,              /* QFRIC CSHEAR FHFL */
    1,                    /* DAMD */
    1, 3, 1               /* ZWAB VXSPH SUMDEN */
};

typedef struct { int zone, type, value_type, amp_id; double value[3], amp_factor, T, value_ang[3]; } bc_t;

/* rigid triangle mesh (Source/Mesh.h, Source/Mesh.cpp) */
typedef struct {
  int id, nnodes, nelem, first;   /* first = index of its first rigid particle */
  double *node, *node_v;          /* 3 per node */
  int *el;                        /* 3 node indices per element */
  double *centroid, *normal, *pplane, *radius;
  int *nfar;
  double mv[3];                   /* TriMesh::m_v */
  double mw[3];                   /* TriMesh::m_w (rotation about the origin) */
} mesh_t;
typedef struct { int id, n; double t[WFO_MAX_AMP_PTS], v[WFO_MAX_AMP_PTS]; } amp_t;

struct wfo_domain {
  int n, solid_count;
  double *f[WFO_NFIELDS];
  int *fi[WFO_NIFIELDS];
  wfo_material mats[256];
  /* per-material Coulomb friction (static, dynamic); a negative static
   * value means the domain's friction_sta / friction_dyn */
  double mat_fr[256][2];
  /* Domain::model_damage and each material's DamageModel (0 none,
   * 1 JohnsonCookDamage D1..D5, Source/Material.h:47-62) */
  int model_damage;
  int dam_model[256];
  double dam_par[256][5];
  /* Domain::Dimension: 3, or 2 for plane strain (all particles in one
   * z-plane; Source/Domain.h:331) */
  int dim;
  /* domain flags / parameters (Source/Domain.cpp:78-175) */
  double Cellfac, DomMax[3], DomMin[3], Time, deltat, pl_work_heat_frac;
  int ts_nb_inc, thermal_solver, pl_heating;
  /* cell grid (Source/Neighbour.cpp:10-139) */
  double TRPR[3], BLPF[3], CellSize[3], hmax;
  int CellNo[3];
  int *HOC, *LL;
  /* pairs: SMPairs in Nproc=1 order, (temp1, temp2) */
  int *pairs;
  int64_t npairs, cap_pairs, n_other_pairs;
  /* Nishimura tables (Source/Neighbour.cpp:570-694) */
  int *ipair, *jpair, *Anei, *Aref;
  double *pair_force, *pair_densinc, *pair_sr, *pair_rr;
  /* solver state (Source/SolverLeapfrog.cpp:71-77) */
  int isfirst, isyielding, ts_i, nb_cleared, initialized;
  double max_pl_strain, kin_energy_sum, int_energy_sum, plastic_work;
  double m_maxT, m_minT;
  int64_t steps, nb_searches;
  /* BCs */
  bc_t bcs[WFO_MAX_BCS];
  int nbcs;
  amp_t amps[WFO_MAX_AMPS];
  int namps;
  double react[WFO_MAX_ZONES][3];
  int react_pending;
  double *temp_work, *dkin_work;
  /* contact (Source/Contact.cpp, Source/Contact_Wang.cpp) */
  int contact, nmesh, id_free_surf;
  mesh_t mesh[WFO_MAX_MESH];
  double friction_sta, friction_dyn, DFAC;
  /* Domain::contact_alg (Source/Domain.h:64: 0 Fraser, 1 Wang, 2 Seo,
   * 4 LSDyna), PFAC, fric_type == Fr_Bound, friction_function == Linear
   * with friction_m / friction_b, and Particle::cont_stiff (Fraser
   * 3-153, formed at wfo_solve_init for the Fraser and Seo laws) */
  int contact_alg, fric_bound, friction_linear;
  double PFAC, friction_m, friction_b;
  double *cont_stiff;
  /* contact heat (Source/Domain.h:339-347): conduction through the contact
   * with conductance contact_hc, friction work as heat; their totals */
  int cont_heat_cond, cont_heat_fric;
  double contact_hc, contact_friction_work, accum_cont_heat_cond;
  /* the contact time-step criterion (auto_ts bit 2, Domain::auto_ts_cont)
   * and its running minimum Domain::min_force_ts */
  double min_force_ts;
  int *rig, *cont;                /* RIGPairs and ContPairs, (temp1, temp2) */
  int64_t nrig, caprig, ncont, capcont;
  double m_contact_force[WFO_MAX_MESH][3], contact_force_sum, contact_reaction_sum, ext_forces_work;
  /* m_contact_force split by the particle's material index (one group per
   * specimen of a batch) */
  double g_contact_force[256][WFO_MAX_MESH][3];
  /* adaptive time step (Source/Domain.cpp:206-319): bit 0 velocity
   * criterion (auto_ts), bit 1 acceleration criterion (auto_ts_acc), bit 2
   * the contact criterion (auto_ts_cont), which acts only with bit 0 or 1 */
  int auto_ts;
  double CFL, sqrt_h_a, manual_min_ts, deltatint, deltatmin;
  /* Domain::nonlock_sum (1: Nishimura pair buffers and reductions, 0:
   * Locking, each pair adds into both particles in pair order) and
   * Domain::gradKernelCorr (Bonet kernel-gradient correction) */
  int nonlock_sum, grad_corr;
  double *pair_force2;            /* Locking: the second particle's pair term */
};

static int contact_nb_update(wfo_domain *d);

static void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s);
  if (!p) { fprintf(stderr, "wf_oracle: out of memory\n"); abort(); }
  return p;
}

wfo_domain *wfo_create(int n) {
  wfo_domain *d = (wfo_domain *)xcalloc(1, sizeof(wfo_domain));
  d->n = n;
  d->solid_count = n;
  for (int k = 0; k < WFO_NFIELDS; ++k) d->f[k] = (double *)xcalloc((size_t)n * k_fdim[k], sizeof(double));
  for (int k = 0; k < WFO_NIFIELDS; ++k) d->fi[k] = (int *)xcalloc((size_t)n, sizeof(int));
  /* Particle constructor defaults (Source/Particle.cpp:28-157) */
  for (int i = 0; i < n; ++i) {
    d->fi[WFO_FREE][i] = 1;
    d->f[WFO_H][i] = 0.0;
    d->f[WFO_TIN][i] = 4.0;
  }
  for (int k = 0; k < 256; ++k) d->mat_fr[k][0] = d->mat_fr[k][1] = -1.;
  d->contact_alg = 1;  /* Wang: the JSON driver's default */
  d->PFAC = 0.5;       /* the JSON driver's penaltyFactor default (Source/WeldForm.cpp:685) */
  /* Domain constructor defaults (Source/Domain.cpp:78-175) */
  d->Cellfac = 2.0;
  d->dim = 3;
  for (int k = 0; k < 3; ++k) { d->DomMax[k] = -100000000000.0; d->DomMin[k] = 100000000000.0; }
  d->ts_nb_inc = 5;
  d->pl_work_heat_frac = 0.9;
  d->CFL = 0.7;
  d->sqrt_h_a = 0.0025;
  /* Domain::manual_min_ts is declared int (Source/Domain.h:328), so the
   * 1.e-10 of Source/Domain.cpp:170 stores 0: no floor by default */
  d->manual_min_ts = 0.0;
  d->nonlock_sum = 1;
  d->LL = (int *)xcalloc((size_t)n, sizeof(int));
  d->ipair = (int *)xcalloc((size_t)n, sizeof(int));
  d->jpair = (int *)xcalloc((size_t)n, sizeof(int));
  d->Anei = (int *)xcalloc((size_t)n * WFO_MAX_NB, sizeof(int));
  d->Aref = (int *)xcalloc((size_t)n * WFO_MAX_NB, sizeof(int));
  d->temp_work = (double *)xcalloc((size_t)n, sizeof(double));
  d->dkin_work = (double *)xcalloc((size_t)n, sizeof(double));
  d->isfirst = 1;
  return d;
}

void wfo_destroy(wfo_domain *d) {
  if (!d) return;
  for (int k = 0; k < WFO_NFIELDS; ++k) free(d->f[k]);
  for (int k = 0; k < WFO_NIFIELDS; ++k) free(d->fi[k]);
  free(d->HOC); free(d->LL); free(d->pairs); free(d->ipair); free(d->jpair);
  free(d->Anei); free(d->Aref); free(d->pair_force); free(d->pair_densinc);
  free(d->pair_sr); free(d->pair_rr); free(d->temp_work); free(d->dkin_work); free(d->pair_force2);
  for (int m = 0; m < d->nmesh; m++) {
    mesh_t *M = &d->mesh[m];
    free(M->node); free(M->node_v); free(M->el); free(M->centroid); free(M->normal);
    free(M->pplane); free(M->radius); free(M->nfar);
  }
  free(d->rig); free(d->cont); free(d->cont_stiff);
  free(d);
}

int wfo_n(const wfo_domain *d) { return d->n; }
double *wfo_f64(wfo_domain *d, int f) { return (f >= 0 && f < WFO_NFIELDS) ? d->f[f] : NULL; }
int *wfo_i32(wfo_domain *d, int f) { return (f >= 0 && f < WFO_NIFIELDS) ? d->fi[f] : NULL; }

int wfo_set_material(wfo_domain *d, int idx, const wfo_material *m) {
  if (idx < 0 || idx >= 256) return fail("material index out of range");
  d->mats[idx] = *m;
  if (m->model == WFO_HOLLOMON) {
    /* Hollomon ctor checks eps0 <= eps1 (Source/Material.cpp:48-57) */
    double eps0 = m->sy0 / m->E, eps1 = pow(m->sy0 / m->K, 1. / m->m);
    if (eps0 > eps1) return fail("Hollomon material bad definition");
  }
  return 0;
}

/* Friction per material: the batched ensembles give every specimen its own
 * Coulomb coefficients; the reference has one pair per domain
 * (Domain::friction_sta / friction_dyn, Source/Contact_Wang.cpp:250-280). */
int wfo_set_material_friction(wfo_domain *d, int idx, double sta, double dyn) {
  if (idx < 0 || idx >= 256) return fail("material index out of range");
  d->mat_fr[idx][0] = sta;
  d->mat_fr[idx][1] = dyn;
  return 0;
}

int wfo_set_material_damage(wfo_domain *d, int idx, int model, const double D[5]) {
  if (idx < 0 || idx >= 256) return fail("material index out of range");
  if (model != 0 && model != 1) return fail("damage model: 0 none, 1 Johnson-Cook");
  d->dam_model[idx] = model;
  for (int k = 0; k < 5; k++) d->dam_par[idx][k] = model ? D[k] : 0.0;
  return 0;
}

int wfo_set_param(wfo_domain *d, const char *s, double v) {
  if (!strcmp(s, "Cellfac")) d->Cellfac = v;
  else if (!strcmp(s, "Dimension")) {
    if (v != 2.0 && v != 3.0) return fail("Dimension must be 2 or 3");
    d->dim = (int)v;
  }
  else if (!strcmp(s, "ts_nb_inc")) d->ts_nb_inc = (int)v;
  else if (!strcmp(s, "thermal_solver")) d->thermal_solver = (int)v;
  else if (!strcmp(s, "pl_heating")) d->pl_heating = (int)v;
  else if (!strcmp(s, "pl_work_heat_frac")) d->pl_work_heat_frac = v;
  else if (!strcmp(s, "solid_count")) d->solid_count = (int)v;
  else if (!strcmp(s, "contact")) d->contact = (int)v;
  else if (!strcmp(s, "friction_sta")) d->friction_sta = v;
  else if (!strcmp(s, "friction_dyn")) d->friction_dyn = v;
  else if (!strcmp(s, "contact_alg")) d->contact_alg = (int)v;
  else if (!strcmp(s, "model_damage")) d->model_damage = (int)v;
  else if (!strcmp(s, "PFAC")) d->PFAC = v;
  else if (!strcmp(s, "fric_bound")) d->fric_bound = (int)v;
  else if (!strcmp(s, "friction_linear")) d->friction_linear = (int)v;
  else if (!strcmp(s, "friction_m")) d->friction_m = v;
  else if (!strcmp(s, "friction_b")) d->friction_b = v;
  else if (!strcmp(s, "DFAC")) d->DFAC = v;
  else if (!strcmp(s, "cont_heat_cond")) d->cont_heat_cond = (int)v;
  else if (!strcmp(s, "cont_heat_fric")) d->cont_heat_fric = (int)v;
  else if (!strcmp(s, "contact_hc")) d->contact_hc = v;
  else if (!strcmp(s, "min_force_ts")) d->min_force_ts = v;
  else if (!strcmp(s, "auto_ts")) d->auto_ts = (int)v;
  else if (!strcmp(s, "CFL")) d->CFL = v;
  else if (!strcmp(s, "sqrt_h_a")) d->sqrt_h_a = v;
  else if (!strcmp(s, "manual_min_ts")) d->manual_min_ts = v;
  else if (!strcmp(s, "nonlock_sum")) d->nonlock_sum = (int)v;
  else if (!strcmp(s, "grad_corr")) d->grad_corr = (int)v;
  else if (!strcmp(s, "Time")) d->Time = v;
  else if (!strcmp(s, "deltat")) d->deltat = v;
  else if (!strncmp(s, "DomMax", 6) && s[6] >= '0' && s[6] <= '2') d->DomMax[s[6] - '0'] = v;
  else if (!strncmp(s, "DomMin", 6) && s[6] >= '0' && s[6] <= '2') d->DomMin[s[6] - '0'] = v;
  else return fail("unknown parameter");
  return 0;
}

double wfo_get_param(const wfo_domain *d, const char *s) {
  if (!strcmp(s, "Time")) return d->Time;
  if (!strcmp(s, "deltat")) return d->deltat;
  if (!strcmp(s, "kin_energy_sum")) return d->kin_energy_sum;
  if (!strcmp(s, "contact_force_sum")) return d->contact_force_sum;
  if (!strcmp(s, "contact_reaction_sum")) return d->contact_reaction_sum;
  if (!strcmp(s, "ext_forces_work")) return d->ext_forces_work;
  if (!strcmp(s, "min_force_ts")) return d->min_force_ts;
  if (!strcmp(s, "contact_friction_work")) return d->contact_friction_work;
  if (!strcmp(s, "accum_cont_heat_cond")) return d->accum_cont_heat_cond;
  if (!strcmp(s, "ncont")) return (double)d->ncont;
  if (!strcmp(s, "int_energy_sum")) return d->int_energy_sum;
  if (!strcmp(s, "plastic_work")) return d->plastic_work;
  if (!strcmp(s, "max_pl_strain")) return d->max_pl_strain;
  if (!strcmp(s, "steps")) return (double)d->steps;
  if (!strcmp(s, "nb_searches")) return (double)d->nb_searches;
  if (!strcmp(s, "hmax")) return d->hmax;
  if (!strcmp(s, "Dimension")) return (double)d->dim;
  if (!strcmp(s, "m_maxT")) return d->m_maxT;
  if (!strcmp(s, "m_minT")) return d->m_minT;
  if (!strcmp(s, "n_other_pairs")) return (double)d->n_other_pairs;
  if (!strcmp(s, "nonlock_sum")) return (double)d->nonlock_sum;
  if (!strcmp(s, "grad_corr")) return (double)d->grad_corr;
  if (!strncmp(s, "CellNo", 6)) return d->CellNo[s[6] - '0'];
  if (!strncmp(s, "CellSize", 8)) return d->CellSize[s[8] - '0'];
  if (!strncmp(s, "BLPF", 4)) return d->BLPF[s[4] - '0'];
  if (!strncmp(s, "TRPR", 4)) return d->TRPR[s[4] - '0'];
  return NAN;
}

/* ------------------------------------------------------------------ */
/* Cubic-spline kernel (Source/Functions.cpp:25-37, 74-87): C = 1/(pi h^3)
 * for W and 1/(pi h^4) for the gradient in 3D, 10/(7 pi h^2) and
 * 10/(7 pi h^3) in 2D */
double wfo_kernel_dim(double q, double h, int dim) {
  double C = dim == 2 ? 10.0 / (7.0 * h * h * M_PI) : 1.0 / (h * h * h * M_PI);
  if (q < 1.0) return C * (1.0 - (3.0 / 2.0) * q * q + (3.0 / 4.0) * q * q * q);
  else if (q < 2.0) return C * ((1.0 / 4.0) * (2.0 - q) * (2.0 - q) * (2.0 - q));
  return 0.0;
}
double wfo_grad_kernel_dim(double q, double h, int dim) {
  double C = dim == 2 ? 10.0 / (7.0 * h * h * h * M_PI) : 1.0 / (h * h * h * h * M_PI);
  if (q == 0.0) return C / h * (-3.0 + (9.0 / 2.0) * q);
  else if (q < 1.0) return C / (q * h) * (-3.0 * q + (9.0 / 4.0) * q * q);
  else if (q < 2.0) return C / (q * h) * ((-3.0 / 4.0) * (2.0 - q) * (2.0 - q));
  return 0.0;
}
double wfo_kernel(double q, double h) { return wfo_kernel_dim(q, h, 3); }
double wfo_grad_kernel(double q, double h) { return wfo_grad_kernel_dim(q, h, 3); }

static inline double norm3(const double *a) {
  return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
}
static inline double dot3(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* ------------------------------------------------------------------ */
/* Geometry: calcHalfPartCount + AddCylinderLength, 3D, no ghost rows  */
/* (Source/Domain.cpp:674-689, 708-860)                                 */
static int half_part_count(double r, double R, int xinc) {
  int ypartcount = -1;
  if (xinc > 0) {
    ypartcount = 1;
    double yp = r;
    double xp = r + (double)(xinc - 1) * 2. * r;
    double rad = sqrt(yp * yp + xp * xp);
    while (rad <= R - r) {
      yp += 2. * r;
      rad = sqrt(yp * yp + xp * xp);
      ypartcount++;
    }
    ypartcount -= 1;
  }
  return ypartcount;
}

int64_t wfo_gen_cylinder(const double V[3], double Rxy, double Lz, double r, double *out, int64_t cap) {
  int64_t cnt = 0;
  int numpartxy = half_part_count(r, Rxy, 1);
  double zp = V[2] + r;
  while (zp <= (V[2] + Lz - r)) {
    double yp = V[1] - r - (2. * r * (numpartxy - 1));
    int numypart = 2 * numpartxy;
    int yinc = numpartxy, yinc_sign = -1;
    for (int j = 0; j < numypart; j++) {
      int numxpart = half_part_count(r, Rxy, yinc);
      double xp = V[0] - r - (2. * r * (numxpart - 1));
      for (int i = 0; i < 2 * numxpart; i++) {
        if (out && cnt < cap) { out[3 * cnt] = xp; out[3 * cnt + 1] = yp; out[3 * cnt + 2] = zp; }
        cnt++;
        xp += 2. * r;
      }
      yp += 2. * r;
      yinc += yinc_sign;
      if (yinc < 1) { yinc = 1; yinc_sign = 1; }
    }
    zp += 2. * r;
  }
  return cnt;
}

/* AssignZone (Source/Domain.cpp:691-706) */
int wfo_assign_zone(wfo_domain *d, const double start[3], const double end[3], int id) {
  int cnt = 0;
  for (int a = 0; a < d->n; a++) {
    int inc = 1;
    for (int i = 0; i < 3; i++)
      if (d->f[WFO_X][3 * a + i] < start[i] || d->f[WFO_X][3 * a + i] > end[i]) inc = 0;
    if (inc) { d->fi[WFO_ID][a] = id; cnt++; }
  }
  return cnt;
}

/* ------------------------------------------------------------------ */
/* Boundary conditions: amplitude::getValAtTime (Source/Domain.h:75-87)*/
static double amp_value(const amp_t *a, double t) {
  int i = a->n - 1;
  int end = 0;
  while (!end) {
    i--;
    if (i == 0 || t > a->t[i]) end = 1;
  }
  return a->v[i] + (a->v[i + 1] - a->v[i]) / (a->t[i + 1] - a->t[i]) * (t - a->t[i]);
}

int wfo_add_velocity_bc(wfo_domain *d, int zone, int value_type, const double value[3], int amp_id, double amp_factor) {
  if (d->nbcs >= WFO_MAX_BCS) return fail("too many BCs");
  bc_t *b = &d->bcs[d->nbcs++];
  memset(b, 0, sizeof *b);
  b->zone = zone; b->type = 0; b->value_type = value_type; b->amp_id = amp_id; b->amp_factor = amp_factor;
  for (int k = 0; k < 3; k++) b->value[k] = value[k];
  return 0;
}
/* boundaryCondition::value_ang of the velocity BCs of a zone: the rotation
 * a constant BC gives its rigid mesh (Source/WeldForm.cpp:97-101) */
int wfo_set_bc_angular(wfo_domain *d, int zone, const double w[3]) {
  int found = 0;
  for (int b = 0; b < d->nbcs; b++)
    if (d->bcs[b].type == 0 && d->bcs[b].zone == zone) {
      for (int k = 0; k < 3; k++) d->bcs[b].value_ang[k] = w[k];
      found = 1;
    }
  return found ? 0 : fail("no velocity BC on that zone");
}
int wfo_add_temperature_bc(wfo_domain *d, int zone, double T) {
  if (d->nbcs >= WFO_MAX_BCS) return fail("too many BCs");
  bc_t *b = &d->bcs[d->nbcs++];
  memset(b, 0, sizeof *b);
  b->zone = zone; b->type = 2; b->T = T;
  return 0;
}
int wfo_add_amplitude(wfo_domain *d, int id, int npts, const double *t, const double *v) {
  if (d->namps >= WFO_MAX_AMPS || npts < 2 || npts > WFO_MAX_AMP_PTS) return fail("bad amplitude");
  amp_t *a = &d->amps[d->namps++];
  a->id = id; a->n = npts;
  for (int k = 0; k < npts; k++) { a->t[k] = t[k]; a->v[k] = v[k]; }
  return 0;
}

/* UserAcc (Source/WeldForm.cpp:51-92): velocity BC zeroes a and imposes v;
 * temperature BC imposes T.  The first call of each step also records the
 * reaction sum(m*a) of every velocity-BC zone before a is zeroed. */
int wfo_general_after(wfo_domain *d) {
  double *x_a = d->f[WFO_A], *v = d->f[WFO_V], *m = d->f[WFO_MASS], *T = d->f[WFO_T];
  int *id = d->fi[WFO_ID];
  if (d->react_pending) {
    memset(d->react, 0, sizeof d->react);
    for (int b = 0; b < d->nbcs; b++) {
      if (d->bcs[b].type != 0) continue;
      int z = d->bcs[b].zone;
      if (z < 0 || z >= WFO_MAX_ZONES) continue;
      double s[3] = {0, 0, 0};
      for (int i = 0; i < d->n; i++)
        if (id[i] == z)
          for (int k = 0; k < 3; k++) s[k] += m[i] * x_a[3 * i + k];
      for (int k = 0; k < 3; k++) d->react[z][k] = s[k];
    }
    d->react_pending = 0;
  }
  for (int i = 0; i < d->n; i++) {
    for (int b = 0; b < d->nbcs; b++) {
      const bc_t *bc = &d->bcs[b];
      if (id[i] != bc->zone) continue;
      if (bc->type == 0) {
        if (bc->value_type == 0) {
          for (int k = 0; k < 3; k++) { x_a[3 * i + k] = 0.0; v[3 * i + k] = bc->value[k]; }
        } else if (bc->value_type == 1) {
          for (int j = 0; j < d->namps; j++) {
            if (d->amps[j].id != bc->amp_id) continue;
            double val = bc->amp_factor * amp_value(&d->amps[j], d->Time);
            for (int k = 0; k < 3; k++) { x_a[3 * i + k] = 0.0; v[3 * i + k] = val * bc->value[k]; }
          }
        }
      } else if (bc->type == 2) {
        T[i] = bc->T;
      }
    }
  }
  /* rigid meshes follow the velocity BC of their zone (Source/WeldForm.cpp:94-119) */
  for (int b = 0; b < d->nbcs; b++) {
    const bc_t *bc = &d->bcs[b];
    if (bc->type != 0) continue;
    for (int m = 0; m < d->nmesh; m++) {
      if (d->mesh[m].id != bc->zone) continue;
      if (bc->value_type == 0) {
        for (int k = 0; k < 3; k++) {
          d->mesh[m].mv[k] = bc->value[k];
          d->mesh[m].mw[k] = bc->value_ang[k];   /* SetRotAxisVel, constant BCs only */
        }
      } else if (bc->value_type == 1) {
        for (int j = 0; j < d->namps; j++) {
          if (d->amps[j].id != bc->amp_id) continue;
          double val = bc->amp_factor * amp_value(&d->amps[j], d->Time);
          for (int k = 0; k < 3; k++) d->mesh[m].mv[k] = val * bc->value[k];
        }
      }
    }
  }
  return 0;
}

int wfo_reaction(const wfo_domain *d, int zone, double out[3]) {
  if (zone < 0 || zone >= WFO_MAX_ZONES) return fail("zone out of range");
  for (int k = 0; k < 3; k++) out[k] = d->react[zone][k];
  return 0;
}

/* ------------------------------------------------------------------ */
/* InitialChecks (Source/Domain.cpp:2097-2113): EOS pressure (PresEq 0) */
static void initial_checks(wfo_domain *d) {
  for (int i = 0; i < d->n; i++)
    d->f[WFO_P][i] = d->f[WFO_P0][i] + (d->f[WFO_CS][i] * d->f[WFO_CS][i]) * (d->f[WFO_RHO][i] - d->f[WFO_REFRHO][i]);
}

/* CellInitiate, non-periodic (Source/Neighbour.cpp:10-139); in 2D one
 * layer of cells along z (CellNo[2] = 1, CellSize[2] = 0, :56-75) */
static int cell_initiate(wfo_domain *d) {
  const double *x = d->f[WFO_X];
  if (!(norm3(d->TRPR) > 0.0) && !(norm3(d->BLPF) > 0.0)) {
    for (int k = 0; k < 3; k++) d->BLPF[k] = d->TRPR[k] = x[k];
    d->hmax = d->f[WFO_H][0];
    for (int i = 0; i < d->n; i++) {
      for (int k = 0; k < 3; k++) {
        if (x[3 * i + k] > d->TRPR[k]) d->TRPR[k] = x[3 * i + k];
        if (x[3 * i + k] < d->BLPF[k]) d->BLPF[k] = x[3 * i + k];
      }
      if (d->f[WFO_H][i] > d->hmax) d->hmax = d->f[WFO_H][i];
    }
  }
  for (int k = 0; k < 3; k++) {
    if (d->DomMax[k] > d->TRPR[k]) d->TRPR[k] = d->DomMax[k];
    if (d->DomMin[k] < d->BLPF[k]) d->BLPF[k] = d->DomMin[k];
  }
  for (int k = 0; k < 3; k++) { d->TRPR[k] += d->hmax / 2; d->BLPF[k] -= d->hmax / 2; }
  for (int k = 0; k < d->dim; k++) {
    double e = (d->TRPR[k] - d->BLPF[k]) / (d->Cellfac * d->hmax);
    if ((double)(ceil(e) - e) < (d->hmax / 10.0)) d->CellNo[k] = (int)ceil(e);
    else d->CellNo[k] = (int)floor(e);
    if (d->CellNo[k] <= 0) return fail("CellInitiate: zero cells along an axis");
  }
  if (d->dim == 2) d->CellNo[2] = 1;
  for (int k = 0; k < 3; k++) d->CellSize[k] = (d->TRPR[k] - d->BLPF[k]) / d->CellNo[k];
  if (d->dim == 2) d->CellSize[2] = 0.0;
  int64_t nc = (int64_t)d->CellNo[0] * d->CellNo[1] * d->CellNo[2];
  if (nc > (int64_t)1 << 30) return fail("CellInitiate: too many cells");
  free(d->HOC);
  d->HOC = (int *)xcalloc((size_t)nc, sizeof(int));
  for (int64_t c = 0; c < nc; c++) d->HOC[c] = -1;
  return 0;
}

#define HOC_AT(d, i, j, k) (d)->HOC[((int64_t)(i) * (d)->CellNo[1] + (j)) * (d)->CellNo[2] + (k)]

/* CellReset + ListGenerate (Source/Neighbour.cpp:141-284); in 2D every
 * particle is in the one z layer.  The cell pairs below are the 3D
 * stencil, which over one layer visits the same pairs of cells. */
static void list_generate(wfo_domain *d) {
  int64_t nc = (int64_t)d->CellNo[0] * d->CellNo[1] * d->CellNo[2];
  for (int64_t c = 0; c < nc; c++) d->HOC[c] = -1;
  const double *x = d->f[WFO_X];
  for (int a = 0; a < d->n; a++) {
    int c[3];
    for (int k = 0; k < 3; k++) {
      if (k == 2 && d->dim == 2) {
        c[k] = 0;
        continue;
      }
      c[k] = (int)(floor((x[3 * a + k] - d->BLPF[k]) / d->CellSize[k]));
      if (c[k] < 0) c[k] = 0;
      if (c[k] >= d->CellNo[k]) c[k] = d->CellNo[k] - 1;
    }
    int temp = HOC_AT(d, c[0], c[1], c[2]);
    HOC_AT(d, c[0], c[1], c[2]) = a;
    d->LL[a] = temp;
  }
}

/* CheckRadius (Source/Neighbour.cpp:361-372) */
static inline int check_radius(const wfo_domain *d, int p1, int p2) {
  const double *x = d->f[WFO_X];
  double h = (d->f[WFO_H][p1] + d->f[WFO_H][p2]) / 2;
  double xij[3] = {x[3 * p1] - x[3 * p2], x[3 * p1 + 1] - x[3 * p2 + 1], x[3 * p1 + 2] - x[3 * p2 + 2]};
  double rij = norm3(xij);
  return (rij / h) <= d->Cellfac;
}

typedef struct { int *p; int64_t n, cap; int64_t other; int *r; int64_t nr, capr; } pairbuf_t;
static void push2(int **p, int64_t *n, int64_t *cap, int a, int c) {
  if (*n == *cap) {
    *cap = *cap ? 2 * *cap : 4096;
    *p = (int *)realloc(*p, (size_t)*cap * 2 * sizeof(int));
    if (!*p) { fprintf(stderr, "wf_oracle: out of memory\n"); abort(); }
  }
  (*p)[2 * *n] = a; (*p)[2 * *n + 1] = c; (*n)++;
}
static void pb_push(pairbuf_t *b, int a, int c) { push2(&b->p, &b->n, &b->cap, a, c); }

static inline int is_mesh_id(const wfo_domain *d, int id) {
  for (int m = 0; m < d->nmesh; m++)
    if (d->mesh[m].id == id) return 1;
  return 0;
}

/* AllocateNbPair without contact (Source/Neighbour.cpp:374-417): both free
 * and same Material -> SMPairs.  Free/fixed and cross-material pairs feed
 * paths this oracle does not model; they are only counted. */
static inline void allocate_pair(const wfo_domain *d, int t1, int t2, pairbuf_t *b) {
  if (d->contact) {
    /* contact branch (Source/Neighbour.cpp:375-396): a pair with a
     * rigid-mesh particle goes to RIGPairs, without a radius test, when the
     * other particle carries the free-surface ID; it never reaches SMPairs */
    const int *id = d->fi[WFO_ID];
    if (is_mesh_id(d, id[t1]) || is_mesh_id(d, id[t2])) {
      if (id[t1] == d->id_free_surf || id[t2] == d->id_free_surf) push2(&b->r, &b->nr, &b->capr, t1, t2);
      return;
    }
  }
  if (!check_radius(d, t1, t2)) return;
  const int *fr = d->fi[WFO_FREE], *mt = d->fi[WFO_MATERIAL];
  if (fr[t1] || fr[t2]) {
    if (mt[t1] == mt[t2] && fr[t1] && fr[t2]) pb_push(b, t1, t2);
    else b->other++;
  }
}

/* YZPlaneCellsNeighbourSearch (Source/Neighbour.cpp:418-477) */
static void yz_plane_search(const wfo_domain *d, int q1, pairbuf_t *b) {
  const int *CN = d->CellNo;
  const int *LL = d->LL;
  for (int q3 = 0; q3 < CN[2]; q3++)
    for (int q2 = 0; q2 < CN[1]; q2++) {
      int temp1 = HOC_AT(d, q1, q2, q3);
      if (temp1 == -1) continue;
      while (temp1 != -1) {
        int temp2 = LL[temp1];
        while (temp2 != -1) { allocate_pair(d, temp1, temp2, b); temp2 = LL[temp2]; }
        if (q1 + 1 < CN[0]) {
          temp2 = HOC_AT(d, q1 + 1, q2, q3);
          while (temp2 != -1) { allocate_pair(d, temp1, temp2, b); temp2 = LL[temp2]; }
        }
        if (q2 + 1 < CN[1]) {
          for (int i = q1 - 1; i <= q1 + 1; i++)
            if (i < CN[0] && i >= 0) {
              temp2 = HOC_AT(d, i, q2 + 1, q3);
              while (temp2 != -1) { allocate_pair(d, temp1, temp2, b); temp2 = LL[temp2]; }
            }
        }
        if (q3 + 1 < CN[2]) {
          for (int j = q2 - 1; j <= q2 + 1; j++)
            for (int i = q1 - 1; i <= q1 + 1; i++)
              if (i < CN[0] && i >= 0 && j < CN[1] && j >= 0) {
                temp2 = HOC_AT(d, i, j, q3 + 1);
                while (temp2 != -1) { allocate_pair(d, temp1, temp2, b); temp2 = LL[temp2]; }
              }
        }
        temp1 = LL[temp1];
      }
    }
}

/* MainNeighbourSearch (Source/Neighbour.cpp:346-359).  Planes are searched
 * in parallel into per-plane buffers that are concatenated in q1 order,
 * which is exactly the Nproc=1 SMPairs order. */
static void main_neighbour_search(wfo_domain *d) {
  int nx = d->CellNo[0];
  pairbuf_t *bufs = (pairbuf_t *)xcalloc((size_t)nx, sizeof(pairbuf_t));
#pragma omp parallel for schedule(dynamic) num_threads(g_threads)
  for (int q1 = 0; q1 < nx; q1++) yz_plane_search(d, q1, &bufs[q1]);
  int64_t tot = 0, other = 0;
  for (int q1 = 0; q1 < nx; q1++) { tot += bufs[q1].n; other += bufs[q1].other; }
  if (tot > d->cap_pairs) {
    free(d->pairs);
    d->cap_pairs = tot;
    d->pairs = (int *)xcalloc((size_t)tot * 2, sizeof(int));
  }
  int64_t off = 0;
  for (int q1 = 0; q1 < nx; q1++) {
    if (bufs[q1].n) memcpy(d->pairs + 2 * off, bufs[q1].p, (size_t)bufs[q1].n * 2 * sizeof(int));
    off += bufs[q1].n;
    free(bufs[q1].p);
    for (int64_t k = 0; k < bufs[q1].nr; k++) push2(&d->rig, &d->nrig, &d->caprig, bufs[q1].r[2 * k], bufs[q1].r[2 * k + 1]);
    free(bufs[q1].r);
  }
  free(bufs);
  d->npairs = tot;
  d->n_other_pairs = other;
  d->nb_cleared = 0;
  d->nb_searches++;
}

/* CalcPairPosList (Source/Neighbour.cpp:617-694) with an overflow check
 * the reference lacks; SaveNeighbourData (:499-532). */
static int calc_pair_pos_list(wfo_domain *d) {
  int64_t P = d->npairs;
  free(d->pair_force); free(d->pair_densinc); free(d->pair_sr); free(d->pair_rr);
  d->pair_force = (double *)xcalloc((size_t)P * 3, sizeof(double));
  d->pair_densinc = (double *)xcalloc((size_t)P, sizeof(double));
  d->pair_sr = (double *)xcalloc((size_t)P * 9, sizeof(double));
  d->pair_rr = (double *)xcalloc((size_t)P * 9, sizeof(double));
  memset(d->ipair, 0, sizeof(int) * d->n);
  memset(d->jpair, 0, sizeof(int) * d->n);
  memset(d->fi[WFO_NB], 0, sizeof(int) * d->n);
  for (int64_t pp = 0; pp < P; pp++) {
    int a = d->pairs[2 * pp], b = d->pairs[2 * pp + 1];
    int i = a < b ? a : b, j = a < b ? b : a;
    if (d->ipair[i] + d->jpair[i] >= WFO_MAX_NB || d->ipair[j] + d->jpair[j] >= WFO_MAX_NB)
      return fail("MAX_NB_PER_PART exceeded");
    d->Anei[(int64_t)i * WFO_MAX_NB + d->ipair[i]] = j;
    d->Anei[(int64_t)j * WFO_MAX_NB + WFO_MAX_NB - 1 - d->jpair[j]] = i;
    d->Aref[(int64_t)i * WFO_MAX_NB + d->ipair[i]] = (int)pp;
    d->Aref[(int64_t)j * WFO_MAX_NB + WFO_MAX_NB - 1 - d->jpair[j]] = (int)pp;
    d->ipair[i]++;
    d->jpair[j]++;
    d->fi[WFO_NB][a]++;
    d->fi[WFO_NB][b]++;
  }
  return 0;
}

/* ClearNbData (Source/Neighbour.cpp:479-497) */
static void clear_nb_data(wfo_domain *d) {
  d->npairs = 0;
  d->nrig = 0;
  d->ncont = 0;
  list_generate(d);
  d->nb_cleared = 1;
}

int wfo_neighbour_search(wfo_domain *d) {
  if (!d->HOC && cell_initiate(d)) return -1;
  clear_nb_data(d);
  main_neighbour_search(d);
  if (calc_pair_pos_list(d)) return -1;
  return d->contact ? contact_nb_update(d) : 0;
}
int64_t wfo_npairs(const wfo_domain *d) { return d->npairs; }
const int *wfo_pairs(const wfo_domain *d) { return d->pairs; }

/* ------------------------------------------------------------------ */
/* Symmetric 3x3 eigen-decomposition by cyclic Jacobi rotations (Golub & Van
 * Loan, Matrix Computations, 4th ed., Alg. 8.5.2-8.5.3): A = V diag(w) V^T
 * with the eigenvectors in the columns of V.  It stands in for GSL's
 * gsl_eigen_symmv, which the reference reaches through Rotation / Eig
 * (Source/Functions.cpp:381-395, External/matvec.h:776); GSL is absent, and
 * its use here -- the tensile-instability stress, a spectral function of
 * sigma -- does not depend on the eigenvectors' signs or order. */
void wfo_sym3_eig(const double *Ain, double *w, double *V) {
  double A[3][3];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      A[a][b] = 0.5 * (Ain[3 * a + b] + Ain[3 * b + a]);
      V[3 * a + b] = a == b ? 1.0 : 0.0;
    }
  for (int sweep = 0; sweep < 64; sweep++) {
    double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
    double dia = A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2];
    if (off == 0.0 || off <= 1e-40 * dia) break;
    for (int p = 0; p < 2; p++)
      for (int q = p + 1; q < 3; q++) {
        if (A[p][q] == 0.0) continue;
        double th = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
        double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 3; k++) { /* A <- A J */
          double akp = A[k][p], akq = A[k][q];
          A[k][p] = c * akp - s * akq;
          A[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; k++) { /* A <- J^T A */
          double apk = A[p][k], aqk = A[q][k];
          A[p][k] = c * apk - s * aqk;
          A[q][k] = s * apk + c * aqk;
        }
        A[p][q] = A[q][p] = 0.0;
        for (int k = 0; k < 3; k++) { /* V <- V J */
          double vkp = V[3 * k + p], vkq = V[3 * k + q];
          V[3 * k + p] = c * vkp - s * vkq;
          V[3 * k + q] = s * vkp + c * vkq;
        }
      }
  }
  for (int k = 0; k < 3; k++) w[k] = A[k][k];
}

/* The tensile-instability stress of a particle, 3D branch of
 * StartAcceleration (Source/Domain.cpp:1545-1563): principal stresses
 * s_k > 0 become -TI s_k / rho^2, the others 0, rotated back with the
 * eigenvectors (Vec Val, then (Vec Val) Vec^T). */
static void tensile_instability(wfo_domain *d, int i) {
  const double *sig = d->f[WFO_SIGMA] + 9 * (int64_t)i;
  double rho2 = d->f[WFO_RHO][i] * d->f[WFO_RHO][i];
  double pc = d->f[WFO_TI][i] / (rho2);
  double w[3], V[9], temp[9];
  wfo_sym3_eig(sig, w, V);
  for (int k = 0; k < 3; k++) w[k] = w[k] > 0 ? -pc * w[k] : 0.0;
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) temp[3 * a + b] = V[3 * a + b] * w[b];
  double *tir = d->f[WFO_TIR] + 9 * (int64_t)i;
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      double s = 0.0;
      for (int k = 0; k < 3; k++) s += temp[3 * a + k] * V[3 * b + k];
      tir[3 * a + b] = s;
    }
}

/* StartAcceleration with a = 0 (Source/Domain.cpp:1519-1597); in 2D the
 * velocity's z part is reset too (:1592; the 2D tensile-instability branch
 * is not restated: TI is refused in 2D) */
int wfo_start_acceleration(wfo_domain *d) {
  for (int i = 0; i < d->n; i++) {
    if (d->dim == 2) d->f[WFO_V][3 * i + 2] = 0.0;
    if (d->fi[WFO_FREE][i] && d->f[WFO_TI][i] > 0.0) tensile_instability(d, i);
    if (!d->fi[WFO_FREE][i]) {
      d->f[WFO_P][i] = 0.0;
      memset(d->f[WFO_SIGMA] + 9 * (int64_t)i, 0, 9 * sizeof(double));
      memset(d->f[WFO_SHEAR] + 9 * (int64_t)i, 0, 9 * sizeof(double));
    }
    for (int k = 0; k < 3; k++) d->f[WFO_A][3 * i + k] = 0.0;
    d->f[WFO_DRHO][i] = 0.0;
    memset(d->f[WFO_SR] + 9 * (int64_t)i, 0, 9 * sizeof(double));
    memset(d->f[WFO_RR] + 9 * (int64_t)i, 0, 9 * sizeof(double));
  }
  return 0;
}

/* pair geometry shared by the three passes: P1 = min index, P2 = max */
typedef struct { int p1, p2; double h, xij[3], vij[3], rij, GK; } pairgeo_t;
static inline void pair_geo(const wfo_domain *d, int64_t p, pairgeo_t *g) {
  int a = d->pairs[2 * p], b = d->pairs[2 * p + 1];
  g->p1 = a < b ? a : b;
  g->p2 = a < b ? b : a;
  const double *x = d->f[WFO_X], *v = d->f[WFO_V];
  g->h = (d->f[WFO_H][g->p1] + d->f[WFO_H][g->p2]) / 2;
  for (int k = 0; k < 3; k++) {
    g->xij[k] = x[3 * g->p1 + k] - x[3 * g->p2 + k];
    g->vij[k] = v[3 * g->p1 + k] - v[3 * g->p2 + k];
  }
  g->rij = norm3(g->xij);
  g->GK = wfo_grad_kernel_dim(g->rij / g->h, g->h, d->dim);
}

/* CalcAccel, Nishimura branch, GradientType 0
 * (Source/InteractionAlt.cpp:12-235); the tensile-instability term
 * TIij = (W(r/h) / W(d0/h))^((n_1+n_2)/2) (TIR_1 + TIR_2), d0 the mean
 * TIInitDist, when either particle has TI > 0 (:104-110) */
static void calc_accel_pairs(wfo_domain *d) {
  const double *rho = d->f[WFO_RHO], *sig = d->f[WFO_SIGMA];
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int64_t p = 0; p < d->npairs; p++) {
    pairgeo_t g;
    pair_geo(d, p, &g);
    int p1 = g.p1, p2 = g.p2;
    double Alpha = (d->f[WFO_ALPHA][p1] + d->f[WFO_ALPHA][p2]) / 2.0;
    double Beta = (d->f[WFO_BETA][p1] + d->f[WFO_BETA][p2]) / 2.0;
    double di = rho[p1], dj = rho[p2];
    double PI_s = 0.0; /* PIij = PI_s * I */
    if (Alpha != 0.0 || Beta != 0.0) {
      double MUij = g.h * dot3(g.vij, g.xij) / (g.rij * g.rij + 0.01 * g.h * g.h);
      double Ci = d->f[WFO_CS][p1], Cj = d->f[WFO_CS][p2];
      double Cij = 0.5 * (Ci + Cj);
      if (dot3(g.vij, g.xij) < 0) PI_s = (Alpha * Cij * MUij + Beta * MUij * MUij) / (0.5 * (di + dj));
    }
    double si = 1.0 / (di * di), sj = 1.0 / (dj * dj);
    double tis = 0.0;
    const double *ti = d->f[WFO_TI], *tir = d->f[WFO_TIR];
    int ti_on = ti[p1] > 0.0 || ti[p2] > 0.0;
    if (ti_on) {
      double K = wfo_kernel_dim(g.rij / g.h, g.h, d->dim);
      double K0 = wfo_kernel_dim((d->f[WFO_TIDIST][p1] + d->f[WFO_TIDIST][p2]) / (2.0 * g.h), g.h, d->dim);
      tis = pow((K / K0), (d->f[WFO_TIN][p1] + d->f[WFO_TIN][p2]) / 2.0);
    }
    double M[9];
    for (int a = 0; a < 3; a++)
      for (int c = 0; c < 3; c++) {
        double piac = (a == c) ? 1.0 * PI_s : 0.0 * PI_s;
        double tiac = ti_on ? tis * (tir[9 * (int64_t)p1 + 3 * a + c] + tir[9 * (int64_t)p2 + 3 * a + c]) : 0.0;
        M[3 * a + c] = ((sig[9 * (int64_t)p1 + 3 * a + c] * si + sig[9 * (int64_t)p2 + 3 * a + c] * sj) + piac) + tiac;
      }
    double gv[3] = {g.GK * g.xij[0], g.GK * g.xij[1], g.GK * g.xij[2]};
    for (int c = 0; c < 3; c++)
      d->pair_force[3 * p + c] = gv[0] * M[c] + gv[1] * M[3 + c] + gv[2] * M[6 + c];
    /* plane strain: the pair force has no z part (:184) */
    if (d->dim == 2) d->pair_force[3 * p + 2] = 0.0;
  }
}

/* AccelReduction (Source/InteractionAlt.cpp:237-260) */
static void accel_reduction(wfo_domain *d) {
  const double *m = d->f[WFO_MASS];
  double *acc = d->f[WFO_A];
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int i = 0; i < d->solid_count; i++) {
    double a[3] = {0., 0., 0.};
    const int *an = d->Anei + (int64_t)i * WFO_MAX_NB, *ar = d->Aref + (int64_t)i * WFO_MAX_NB;
    for (int n = 0; n < d->ipair[i]; n++)
      for (int k = 0; k < 3; k++) a[k] = a[k] + m[an[n]] * d->pair_force[3 * (int64_t)ar[n] + k];
    for (int n = 0; n < d->jpair[i]; n++)
      for (int k = 0; k < 3; k++)
        a[k] = a[k] - m[an[WFO_MAX_NB - 1 - n]] * d->pair_force[3 * (int64_t)ar[WFO_MAX_NB - 1 - n] + k];
    for (int k = 0; k < 3; k++) acc[3 * i + k] = a[k];
  }
}

/* g = (GK L) x_ij, Mult(GKc, xij, vc) of the corrected passes
 * (Source/InteractionAlt.cpp:118-133): L the particle's gradCorrM */
static inline void corr_grad(const wfo_domain *d, int i, double GK, const double *xij, double *g) {
  const double *L = d->f[WFO_GCM] + 9 * (int64_t)i;
  for (int a = 0; a < 3; a++) g[a] = (GK * L[3 * a]) * xij[0] + (GK * L[3 * a + 1]) * xij[1] + (GK * L[3 * a + 2]) * xij[2];
}

/* pair geometry of the Locking passes: P1 = first, P2 = second as stored */
static inline void pair_geo_locking(const wfo_domain *d, int64_t p, pairgeo_t *g) {
  g->p1 = d->pairs[2 * p];
  g->p2 = d->pairs[2 * p + 1];
  const double *x = d->f[WFO_X], *v = d->f[WFO_V];
  g->h = (d->f[WFO_H][g->p1] + d->f[WFO_H][g->p2]) / 2;
  for (int k = 0; k < 3; k++) {
    g->xij[k] = x[3 * g->p1 + k] - x[3 * g->p2 + k];
    g->vij[k] = v[3 * g->p1 + k] - v[3 * g->p2 + k];
  }
  g->rij = norm3(g->xij);
  g->GK = wfo_grad_kernel_dim(g->rij / g->h, g->h, d->dim);
}

/* The damage weight of a pair in the Locking passes
 * (Source/InteractionAlt.cpp:54-60,558-563): 1 - D_pair when D_pair > 0,
 * clamped to 1, the pair skipped when it is not positive.  With the
 * Johnson-Cook criterion Domain::CalcDamage sets D_pair = (D_1 + D_2) / 2
 * after every step's particle update while it is below 1 (and keeps 1
 * after, which the mean of two capped, non-decreasing values reaches only
 * when both are 1), so it is formed here from the particles' dam_D
 * (Source/Damage.cpp:141-150); no pair is added after the first step's
 * search when damage is on. */
static inline double pair_damage_factor(const wfo_domain *d, int p1, int p2) {
  if (!d->model_damage) return 1.0;
  const double Dp = (d->f[WFO_DAMD][p1] + d->f[WFO_DAMD][p2]) / 2.0;
  double dam_f = 1.0;
  if (Dp > 0.0) dam_f = 1.0 - Dp;
  if (dam_f > 1.0) dam_f = 1.0;
  return dam_f;
}

/* CalcAccel, Locking branch (nonlock_sum false), GradientType 0
 * (Source/InteractionAlt.cpp:12-235): each pair in pair order adds
 * mj (g1 . M) to particle 1 and subtracts mi (g2 . M) from particle 2, with
 * g = GK x_ij, or g_k = (GK L_k) x_ij under the kernel-gradient
 * correction; a starts from StartAcceleration's 0 */
static void calc_accel_locking(wfo_domain *d) {
  const double *rho = d->f[WFO_RHO], *sig = d->f[WFO_SIGMA], *m = d->f[WFO_MASS];
  double *acc = d->f[WFO_A];
  for (int64_t p = 0; p < d->npairs; p++) {
    pairgeo_t g;
    pair_geo_locking(d, p, &g);
    int p1 = g.p1, p2 = g.p2;
    double Alpha = (d->f[WFO_ALPHA][p1] + d->f[WFO_ALPHA][p2]) / 2.0;
    double Beta = (d->f[WFO_BETA][p1] + d->f[WFO_BETA][p2]) / 2.0;
    double di = rho[p1], dj = rho[p2];
    double PI_s = 0.0;
    if (Alpha != 0.0 || Beta != 0.0) {
      double MUij = g.h * dot3(g.vij, g.xij) / (g.rij * g.rij + 0.01 * g.h * g.h);
      double Cij = 0.5 * (d->f[WFO_CS][p1] + d->f[WFO_CS][p2]);
      if (dot3(g.vij, g.xij) < 0) PI_s = (Alpha * Cij * MUij + Beta * MUij * MUij) / (0.5 * (di + dj));
    }
    double si = 1.0 / (di * di), sj = 1.0 / (dj * dj);
    double tis = 0.0;
    const double *ti = d->f[WFO_TI], *tir = d->f[WFO_TIR];
    int ti_on = ti[p1] > 0.0 || ti[p2] > 0.0;
    if (ti_on) {
      double K = wfo_kernel_dim(g.rij / g.h, g.h, d->dim);
      double K0 = wfo_kernel_dim((d->f[WFO_TIDIST][p1] + d->f[WFO_TIDIST][p2]) / (2.0 * g.h), g.h, d->dim);
      tis = pow((K / K0), (d->f[WFO_TIN][p1] + d->f[WFO_TIN][p2]) / 2.0);
    }
    double M[9];
    for (int a = 0; a < 3; a++)
      for (int c = 0; c < 3; c++) {
        double piac = (a == c) ? 1.0 * PI_s : 0.0 * PI_s;
        double tiac = ti_on ? tis * (tir[9 * (int64_t)p1 + 3 * a + c] + tir[9 * (int64_t)p2 + 3 * a + c]) : 0.0;
        M[3 * a + c] = ((sig[9 * (int64_t)p1 + 3 * a + c] * si + sig[9 * (int64_t)p2 + 3 * a + c] * sj) + piac) + tiac;
      }
    double g1[3], g2[3];
    if (d->grad_corr) {
      corr_grad(d, p1, g.GK, g.xij, g1);
      corr_grad(d, p2, g.GK, g.xij, g2);
    } else {
      for (int k = 0; k < 3; k++) g1[k] = g2[k] = g.GK * g.xij[k];
    }
    const double dam_f = pair_damage_factor(d, p1, p2);
    if (!(dam_f > 0.0)) continue;
    for (int c = 0; c < 3; c++) {
      /* plane strain: temp(2) = 0 (Source/InteractionAlt.cpp:184), which
       * the corrected-gradient terms do not see */
      if (c == 2 && d->dim == 2 && !d->grad_corr) continue;
      double t1 = g1[0] * M[c] + g1[1] * M[3 + c] + g1[2] * M[6 + c];
      double t2 = g2[0] * M[c] + g2[1] * M[3 + c] + g2[2] * M[6 + c];
      acc[3 * p1 + c] += dam_f * m[p2] * t1;
      acc[3 * p2 + c] -= dam_f * m[p1] * t2;
    }
  }
}

int wfo_calc_accel(wfo_domain *d) {
  if (!d->nonlock_sum) {
    for (int i = 0; i < 3 * d->n; i++) d->f[WFO_A][i] = 0.0;
    calc_accel_locking(d);
    return 0;
  }
  calc_accel_pairs(d);
  accel_reduction(d);
  return 0;
}

/* Domain::LastComputeAcceleration's pair loop (Source/Domain.cpp:1700-1709):
 * Domain::CalcForce2233 (Source/Interaction.cpp:25-333) for every pair in
 * pair order, P1 = first, P2 = second, free-free pairs, GradientType 0.
 * a, dDensity, StrainRate and RotationRate are the Locking passes' sums
 * (the same terms, added pair by pair in the same order: m_j g1.M into a,
 * m_j (d_i/d_j) v_ij.g1 into dDensity, (float)(m_j/d_j) times the rate
 * tensors), so they are formed by those functions; the fused loop's own
 * terms follow: ZWab += (float)(m_j/d_j) K, VXSPH += XSPH m_j /
 * ((d_i + d_j)/2) K (-v_ij) (v_ij for particle 2), and SumDen += m_j K when
 * the Shepard sum is due.  Everything starts from StartAcceleration's
 * zeros. */
static void calc_dens_locking(wfo_domain *d);
static void calc_rates_locking(wfo_domain *d);
int wfo_calc_force2233(wfo_domain *d, double xsph, int shepard) {
  for (int i = 0; i < 3 * d->n; i++) d->f[WFO_A][i] = 0.0;
  calc_accel_locking(d);
  calc_dens_locking(d);
  calc_rates_locking(d);
  const double *m = d->f[WFO_MASS], *rho = d->f[WFO_RHO];
  double *zw = d->f[WFO_ZWAB], *vx = d->f[WFO_VXSPH], *sd = d->f[WFO_SUMDEN];
  for (int i = 0; i < d->n; i++) {
    zw[i] = sd[i] = 0.0;
    for (int k = 0; k < 3; k++) vx[3 * i + k] = 0.0;
  }
  for (int64_t p = 0; p < d->npairs; p++) {
    pairgeo_t g;
    pair_geo_locking(d, p, &g);
    const int p1 = g.p1, p2 = g.p2;
    const double K = wfo_kernel_dim(g.rij / g.h, g.h, d->dim);
    const double di = rho[p1], dj = rho[p2], mi = m[p1], mj = m[p2];
    const float mj_dj = (float)(mj / dj), mi_di = (float)(mi / di);
    zw[p1] += mj_dj * K;
    zw[p2] += mi_di * K;
    if (xsph != 0.0)
      for (int k = 0; k < 3; k++) {
        vx[3 * p1 + k] += xsph * mj / (0.5 * (di + dj)) * K * (-g.vij[k]);
        vx[3 * p2 + k] += xsph * mi / (0.5 * (di + dj)) * K * g.vij[k];
      }
    if (shepard) {
      sd[p1] += mj * K;
      sd[p2] += mi * K;
    }
  }
  return 0;
}

/* Domain::CalcGradCorrMatrix (Source/Domain.cpp:1832-1940): for every
 * pair in pair order, m = (GK x_ij) (x) x_ij; particle 1 subtracts
 * (mj/dj) m, particle 2 (mi/di) m; in 2D the sum's zz entry is set to 1
 * (:1911-1915); then each free particle's sum is inverted (InvLog,
 * External/matvec.h:673-693: |det| < 1e-10 fails and the identity is
 * taken); fixed particles get the identity */
int wfo_calc_grad_corr(wfo_domain *d) {
  double *T = (double *)xcalloc((size_t)d->n * 9, sizeof(double));
  const double *rho = d->f[WFO_RHO], *m = d->f[WFO_MASS];
  for (int64_t p = 0; p < d->npairs; p++) {
    pairgeo_t g;
    pair_geo_locking(d, p, &g);
    double gx[3] = {g.GK * g.xij[0], g.GK * g.xij[1], g.GK * g.xij[2]};
    double w1 = m[g.p2] / rho[g.p2], w2 = m[g.p1] / rho[g.p1];
    for (int a = 0; a < 3; a++)
      for (int c = 0; c < 3; c++) {
        double mm = gx[a] * g.xij[c];
        T[9 * (int64_t)g.p1 + 3 * a + c] = T[9 * (int64_t)g.p1 + 3 * a + c] - w1 * mm;
        T[9 * (int64_t)g.p2 + 3 * a + c] = T[9 * (int64_t)g.p2 + 3 * a + c] - w2 * mm;
      }
  }
  double *L = d->f[WFO_GCM];
  for (int i = 0; i < d->solid_count; i++) {
    double *Li = L + 9 * (int64_t)i;
    if (d->dim == 2) T[9 * (int64_t)i + 8] = 1.0;
    const double *M = T + 9 * (int64_t)i;
    int ok = d->fi[WFO_FREE][i];
    if (ok) {
      double det = M[0] * (M[4] * M[8] - M[5] * M[7]) - M[1] * (M[3] * M[8] - M[5] * M[6]) +
                   M[2] * (M[3] * M[7] - M[4] * M[6]);
      if (fabs(det) < 1.0e-10) {
        ok = 0;
      } else {
        Li[0] = (M[4] * M[8] - M[5] * M[7]) / det;
        Li[1] = (M[2] * M[7] - M[1] * M[8]) / det;
        Li[2] = (M[1] * M[5] - M[2] * M[4]) / det;
        Li[3] = (M[5] * M[6] - M[3] * M[8]) / det;
        Li[4] = (M[0] * M[8] - M[2] * M[6]) / det;
        Li[5] = (M[2] * M[3] - M[0] * M[5]) / det;
        Li[6] = (M[3] * M[7] - M[4] * M[6]) / det;
        Li[7] = (M[1] * M[6] - M[0] * M[7]) / det;
        Li[8] = (M[0] * M[4] - M[1] * M[3]) / det;
      }
    }
    if (!ok)
      for (int k = 0; k < 9; k++) Li[k] = (k % 4 == 0) ? 1.0 : 0.0;
  }
  free(T);
  return 0;
}

/* CalcDensInc, Locking branch (Source/InteractionAlt.cpp:483-600): pair
 * order, particle 1 adds mj (di/dj) (v_ij . g1), particle 2 mi (dj/di)
 * (v_ij . g2) */
static void calc_dens_locking(wfo_domain *d) {
  const double *m = d->f[WFO_MASS], *rho = d->f[WFO_RHO];
  double *drho = d->f[WFO_DRHO];
  for (int i = 0; i < d->n; i++) drho[i] = 0.0;
  for (int64_t p = 0; p < d->npairs; p++) {
    pairgeo_t g;
    pair_geo_locking(d, p, &g);
    double g1[3], g2[3];
    if (d->grad_corr) {
      corr_grad(d, g.p1, g.GK, g.xij, g1);
      corr_grad(d, g.p2, g.GK, g.xij, g2);
    } else {
      for (int k = 0; k < 3; k++) g1[k] = g2[k] = g.GK * g.xij[k];
    }
    double t1 = dot3(g.vij, g1), t2 = dot3(g.vij, g2);
    double di = rho[g.p1], dj = rho[g.p2];
    const double dam_f = pair_damage_factor(d, g.p1, g.p2);
    if (!(dam_f > 0.0)) continue;
    drho[g.p1] += dam_f * m[g.p2] * (di / dj) * t1;
    drho[g.p2] += dam_f * m[g.p1] * (dj / di) * t2;
  }
}

/* CalcDensInc + DensReduction (Source/InteractionAlt.cpp:483-640) */
int wfo_calc_dens(wfo_domain *d) {
  if (!d->nonlock_sum) {
    calc_dens_locking(d);
    return 0;
  }
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int64_t p = 0; p < d->npairs; p++) {
    pairgeo_t g;
    pair_geo(d, p, &g);
    double gv[3] = {g.GK * g.xij[0], g.GK * g.xij[1], g.GK * g.xij[2]};
    d->pair_densinc[p] = dot3(g.vij, gv);
  }
  const double *m = d->f[WFO_MASS], *rho = d->f[WFO_RHO];
  double *drho = d->f[WFO_DRHO];
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int i = 0; i < d->solid_count; i++) {
    const int *an = d->Anei + (int64_t)i * WFO_MAX_NB, *ar = d->Aref + (int64_t)i * WFO_MAX_NB;
    double s = 0.;
    for (int n = 0; n < d->ipair[i]; n++) s += m[an[n]] / rho[an[n]] * d->pair_densinc[ar[n]];
    for (int n = 0; n < d->jpair[i]; n++) {
      int j = an[WFO_MAX_NB - 1 - n];
      s += m[j] / rho[j] * d->pair_densinc[ar[WFO_MAX_NB - 1 - n]];
    }
    s *= rho[i];
    drho[i] = s;
  }
  return 0;
}

/* CalcRateTensors, Locking branch (Source/InteractionAlt.cpp:263-440):
 * pair order; the weight is mj/dj rounded to float (`float mj_dj`,
 * :412, :425); without the correction the pair's StrainRate / RotationRate
 * (-0.5 GK times the v (x) x forms), with it -0.5 (G +- G^T) for
 * G = v_ij (x) g_k, g_k = (GK L_k) x_ij */
static void calc_rates_locking(wfo_domain *d) {
  const double *m = d->f[WFO_MASS], *rho = d->f[WFO_RHO];
  for (int i = 0; i < d->n; i++) {
    memset(d->f[WFO_SR] + 9 * (int64_t)i, 0, 9 * sizeof(double));
    memset(d->f[WFO_RR] + 9 * (int64_t)i, 0, 9 * sizeof(double));
  }
  for (int64_t p = 0; p < d->npairs; p++) {
    pairgeo_t g;
    pair_geo_locking(d, p, &g);
    const double *va = g.vij, *xi = g.xij;
    double SR[2][9], RR[2][9];
    if (!d->grad_corr) {
      double S[9], R[9];
      S[0] = 2.0 * va[0] * xi[0];
      S[1] = va[0] * xi[1] + va[1] * xi[0];
      S[2] = va[0] * xi[2] + va[2] * xi[0];
      S[3] = S[1];
      S[4] = 2.0 * va[1] * xi[1];
      S[5] = va[1] * xi[2] + va[2] * xi[1];
      S[6] = S[2];
      S[7] = S[5];
      S[8] = 2.0 * va[2] * xi[2];
      R[0] = R[4] = R[8] = 0.0;
      R[1] = va[0] * xi[1] - va[1] * xi[0];
      R[2] = va[0] * xi[2] - va[2] * xi[0];
      R[5] = va[1] * xi[2] - va[2] * xi[1];
      R[3] = -R[1];
      R[6] = -R[2];
      R[7] = -R[5];
      double s = -0.5 * g.GK;
      for (int k = 0; k < 9; k++) SR[0][k] = SR[1][k] = s * S[k], RR[0][k] = RR[1][k] = s * R[k];
    } else {
      for (int q = 0; q < 2; q++) {
        double gk[3], G[9];
        corr_grad(d, q ? g.p2 : g.p1, g.GK, xi, gk);
        for (int a = 0; a < 3; a++)
          for (int c = 0; c < 3; c++) G[3 * a + c] = va[a] * gk[c];
        for (int a = 0; a < 3; a++)
          for (int c = 0; c < 3; c++) {
            SR[q][3 * a + c] = -0.5 * (G[3 * a + c] + G[3 * c + a]);
            RR[q][3 * a + c] = -0.5 * (G[3 * a + c] - G[3 * c + a]);
          }
      }
    }
    float w1 = (float)(m[g.p2] / rho[g.p2]), w2 = (float)(m[g.p1] / rho[g.p1]);
    double *S1 = d->f[WFO_SR] + 9 * (int64_t)g.p1, *R1 = d->f[WFO_RR] + 9 * (int64_t)g.p1;
    double *S2 = d->f[WFO_SR] + 9 * (int64_t)g.p2, *R2 = d->f[WFO_RR] + 9 * (int64_t)g.p2;
    for (int k = 0; k < 9; k++) {
      S1[k] = S1[k] + w1 * SR[0][k];
      R1[k] = R1[k] + w1 * RR[0][k];
      S2[k] = S2[k] + w2 * SR[1][k];
      R2[k] = R2[k] + w2 * RR[1][k];
    }
  }
}

/* CalcRateTensors + RateTensorsReduction (Source/InteractionAlt.cpp:263-479) */
int wfo_calc_rates(wfo_domain *d) {
  if (!d->nonlock_sum) {
    calc_rates_locking(d);
    return 0;
  }
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int64_t p = 0; p < d->npairs; p++) {
    pairgeo_t g;
    pair_geo(d, p, &g);
    const double *va = g.vij, *xi = g.xij;
    double SR[9], RR[9];
    SR[0] = 2.0 * va[0] * xi[0];
    SR[1] = va[0] * xi[1] + va[1] * xi[0];
    SR[2] = va[0] * xi[2] + va[2] * xi[0];
    SR[3] = SR[1];
    SR[4] = 2.0 * va[1] * xi[1];
    SR[5] = va[1] * xi[2] + va[2] * xi[1];
    SR[6] = SR[2];
    SR[7] = SR[5];
    SR[8] = 2.0 * va[2] * xi[2];
    RR[0] = RR[4] = RR[8] = 0.0;
    RR[1] = va[0] * xi[1] - va[1] * xi[0];
    RR[2] = va[0] * xi[2] - va[2] * xi[0];
    RR[5] = va[1] * xi[2] - va[2] * xi[1];
    RR[3] = -RR[1];
    RR[6] = -RR[2];
    RR[7] = -RR[5];
    double s = -0.5 * g.GK;
    for (int k = 0; k < 9; k++) {
      d->pair_sr[9 * p + k] = SR[k] * s;
      d->pair_rr[9 * p + k] = RR[k] * s;
    }
  }
  const double *m = d->f[WFO_MASS], *rho = d->f[WFO_RHO];
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int i = 0; i < d->solid_count; i++) {
    const int *an = d->Anei + (int64_t)i * WFO_MAX_NB, *ar = d->Aref + (int64_t)i * WFO_MAX_NB;
    double *SRi = d->f[WFO_SR] + 9 * (int64_t)i, *RRi = d->f[WFO_RR] + 9 * (int64_t)i;
    for (int n = 0; n < d->ipair[i]; n++) {
      double mjdj = m[an[n]] / rho[an[n]];
      for (int k = 0; k < 9; k++) {
        SRi[k] = SRi[k] + d->pair_sr[9 * (int64_t)ar[n] + k] * mjdj;
        RRi[k] = RRi[k] + d->pair_rr[9 * (int64_t)ar[n] + k] * mjdj;
      }
    }
    for (int n = 0; n < d->jpair[i]; n++) {
      int j = an[WFO_MAX_NB - 1 - n];
      int64_t pr = ar[WFO_MAX_NB - 1 - n];
      double mjdj = m[j] / rho[j];
      for (int k = 0; k < 9; k++) {
        SRi[k] = SRi[k] + d->pair_sr[9 * pr + k] * mjdj;
        RRi[k] = RRi[k] + d->pair_rr[9 * pr + k] * mjdj;
      }
    }
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* Materials (Source/Material.cpp:6-72)                                 */
static double jc_yield(const wfo_material *M, double strain, double sr, double T) {
  double T_h = (T - M->Tt) / (M->Tm - M->Tt);
  double f = 1.0;
  if (sr > M->eps_0) f = (1.0 + M->C * log(sr / M->eps_0));
  return (M->A + M->B * pow(strain, M->n)) * f * (1.0 - pow(T_h, M->jm));
}
static double jc_tangent(const wfo_material *M, double pl, double sr, double T) {
  double T_h = (T - M->Tt) / (M->Tm - M->Tt);
  double f = 1.0;
  if (sr > M->eps_0) f = (1.0 + M->C * log(sr / M->eps_0));
  if (pl > 0.) return M->n * M->B * pow(pl, M->n - 1.) * f * (1.0 - pow(T_h, M->jm));
  return M->E;
}
static double hol_yield(const wfo_material *M, double strain) {
  double eps0 = M->sy0 / M->E, eps1 = pow(M->sy0 / M->K, 1. / M->m);
  if (strain + eps0 > eps1) return M->K * pow(strain + eps0, M->m);
  return M->sy0;
}
static double hol_tangent(const wfo_material *M, double strain) {
  double eps0 = M->sy0 / M->E, eps1 = pow(M->sy0 / M->K, 1. / M->m);
  if (strain + eps0 > eps1) return M->K * M->m * pow(strain + eps0, (M->m - 1.0));
  return 0.;
}

/* Mult(Mat,Mat) term order of External/matvec.h:845-850 */
static inline void mat_mult(const double *A, const double *B, double *M) {
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++)
      M[3 * a + b] = A[3 * a + 2] * B[6 + b] + A[3 * a + 1] * B[3 + b] + A[3 * a] * B[b];
}

/* Particle::CalcStressStrain with Fail = 1 (Source/Particle.cpp:960-1067) */
static void stress_strain(wfo_domain *d, int i, double dt) {
  double *S = d->f[WFO_SHEAR] + 9 * (int64_t)i, *SR = d->f[WFO_SR] + 9 * (int64_t)i;
  double *RR = d->f[WFO_RR] + 9 * (int64_t)i, *Sig = d->f[WFO_SIGMA] + 9 * (int64_t)i;
  const wfo_material *M = &d->mats[d->fi[WFO_MATIDX][i]];
  int model = d->fi[WFO_MODEL][i];
  double rho = d->f[WFO_RHO][i];
  double Cs = d->f[WFO_CS][i];
  double P = d->f[WFO_P0][i] + (Cs * Cs) * (rho - d->f[WFO_REFRHO][i]);
  d->f[WFO_P][i] = P;
  double RT[9], SRT[9], RS[9];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) RT[3 * a + b] = RR[3 * b + a];
  mat_mult(S, RT, SRT);
  mat_mult(RR, S, RS);
  double dep = 0., sig_trial = 0.;
  double G = d->f[WFO_G][i];
  double c3 = 1.0 / 3.0 * (SR[0] + SR[4] + SR[8]);
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      int k = 3 * a + b;
      double Iab = (a == b) ? 1.0 : 0.0;
      S[k] = dt * ((2.0 * G) * (SR[k] - Iab * c3) + SRT[k] + RS[k]) + S[k];
    }
  d->f[WFO_EFFSR][i] = sqrt(0.5 * ((SR[0] - SR[4]) * (SR[0] - SR[4]) + (SR[4] - SR[8]) * (SR[4] - SR[8]) +
                                   (SR[8] - SR[0]) * (SR[8] - SR[0])) +
                            3.0 * (SR[1] * SR[1] + SR[5] * SR[5] + SR[6] * SR[6]));
  double *Sy = &d->f[WFO_SIGMAY][i], *pl = &d->f[WFO_PLSTRAIN][i], *Ep = &d->f[WFO_EP][i];
  double J2 = 0.5 * (S[0] * S[0] + 2.0 * S[1] * S[3] + 2.0 * S[2] * S[6] + S[4] * S[4] + 2.0 * S[5] * S[7] + S[8] * S[8]);
  sig_trial = sqrt(3.0 * J2);
  double ratio = *Sy / sig_trial;
  double fac = (1.0 < ratio) ? 1.0 : ratio; /* std::min(ratio, 1.0) */
  for (int k = 0; k < 9; k++) S[k] = S[k] * fac;
  if (model == WFO_HOLLOMON) *Sy = hol_yield(M, *pl);
  else if (model == WFO_JOHNSON_COOK) *Sy = jc_yield(M, *pl, d->f[WFO_EFFSR][i], d->f[WFO_T][i]);
  if (sig_trial > *Sy) {
    if (model == WFO_HOLLOMON) d->f[WFO_ET][i] = hol_tangent(M, *pl);
    else if (model == WFO_JOHNSON_COOK) d->f[WFO_ET][i] = jc_tangent(M, *pl, d->f[WFO_EFFSR][i], d->f[WFO_T][i]);
    if (*Ep < 0) *Ep = 1. * M->E;
    dep = (sig_trial - *Sy) / (3. * G + *Ep);
    *pl += dep;
    d->f[WFO_DPLSTRAIN][i] = dep;
    if (model == WFO_BILINEAR) *Sy += dep * *Ep;
  }
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      double Iab = (a == b) ? 1.0 : 0.0;
      Sig[3 * a + b] = Iab * (-P) + S[3 * a + b];
    }
  if (dep > 0.0) {
    double f = dep / *Sy;
    double *inc = d->f[WFO_EPLINC] + 9 * (int64_t)i, *epl = d->f[WFO_EPL] + 9 * (int64_t)i;
    inc[0] = f * (Sig[0] - 0.5 * (Sig[4] + Sig[8]));
    inc[4] = f * (Sig[4] - 0.5 * (Sig[0] + Sig[8]));
    inc[8] = f * (Sig[8] - 0.5 * (Sig[0] + Sig[4]));
    inc[1] = inc[3] = 1.5 * f * (Sig[1]);
    inc[2] = inc[6] = 1.5 * f * (Sig[2]);
    inc[5] = inc[7] = 1.5 * f * (Sig[5]);
    for (int k = 0; k < 9; k++) epl[k] = epl[k] + inc[k];
  }
  double *E = d->f[WFO_STRAIN] + 9 * (int64_t)i;
  for (int k = 0; k < 9; k++) E[k] = dt * SR[k] + E[k];
}

int wfo_stress_all(wfo_domain *d, double dt) {
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int i = 0; i < d->solid_count; i++) stress_strain(d, i, dt);
  return 0;
}

/* JohnsonCookDamage::CalcFractureStrain (Source/Material.cpp:76-92): the
 * strain at fracture at plastic strain rate `sr`, triaxiality `sig_as`
 * and temperature T, with the material's T_t, T_m and eps_0 */
static double jc_fracture_strain(const wfo_material *M, const double *D, double sr, double sig_as, double T) {
  double T_h = (T - M->Tt) / (M->Tm - M->Tt);
  double f_sr = 1.0;
  double f_triax = D[0];
  if (sr > M->eps_0) f_sr = 1.0 + log(sr / M->eps_0) * D[3];
  if (sig_as < 1.5) f_triax = D[0] + D[1] * exp(D[2] * sig_as);
  double ef = f_triax * f_sr * (1.0 + D[4] * T_h);
  if (ef > 10.0) ef = 10.0;
  return ef;
}

/* Domain::CalcDamage, Johnson-Cook criterion (Source/Damage.cpp:9-46):
 * after the stress update, a solid particle whose (never reset)
 * delta_pl_strain is positive and whose von Mises stress
 * (Particle::CalculateEquivalentStress, Source/Particle.cpp:623-633) is
 * above 1e-3 adds delta_pl_strain / eps_f(delta_pl_strain / dt,
 * -P / Sigma_eq, T) to dam_D, capped at 1.  T is the particle's with the
 * thermal solver; the reference's T is otherwise an uninitialised private
 * copy of a local set to 0, taken as 0 here.  The pair values follow from
 * the particles' (pair_damage_factor). */
static void calc_damage(wfo_domain *d) {
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int i = 0; i < d->solid_count; i++) {
    const int mi = d->fi[WFO_MATIDX][i];
    if (d->dam_model[mi] != 1) continue;
    const double dpl = d->f[WFO_DPLSTRAIN][i];
    if (!(dpl > 0.0)) continue;
    const double *S = d->f[WFO_SHEAR] + 9 * (int64_t)i;
    const double J2 = 0.5 * (S[0] * S[0] + 2.0 * S[1] * S[3] + 2.0 * S[2] * S[6] + S[4] * S[4] + 2.0 * S[5] * S[7] +
                             S[8] * S[8]);
    const double seq = sqrt(3.0 * J2);
    if (!(seq > 1.0e-3)) continue;
    const double T = d->thermal_solver ? d->f[WFO_T][i] : 0.0;
    const double ef = jc_fracture_strain(&d->mats[mi], d->dam_par[mi], dpl / d->deltat, -d->f[WFO_P][i] / seq, T);
    double D = d->f[WFO_DAMD][i] + dpl / ef;
    if (D > 1.0) D = 1.0;
    d->f[WFO_DAMD][i] = D;
  }
}

/* Domain::CalcPlasticWorkHeat (Source/ThermalStuct.cpp:14-39) with
 * Particle::CalcPlasticWorkHeat (Source/Particle.cpp:922-939) */
static void plastic_work_heat(wfo_domain *d, double dt) {
  if (!d->pl_heating) return;
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int p = 0; p < d->solid_count; p++) {
    if (d->f[WFO_DPLSTRAIN][p] > 0.0) {
      const double *Sig = d->f[WFO_SIGMA] + 9 * (int64_t)p, *inc = d->f[WFO_EPLINC] + 9 * (int64_t)p;
      double dp[9];
      for (int k = 0; k < 9; k++) dp[k] = inc[k] * (1. / dt);
      d->f[WFO_QPL][p] = Sig[0] * dp[0] + 2.0 * Sig[1] * dp[3] + 2.0 * Sig[2] * dp[6] + Sig[4] * dp[4] +
                         2.0 * Sig[5] * dp[7] + Sig[8] * dp[8];
    }
    d->f[WFO_PSENERGY][p] += d->f[WFO_QPL][p] * dt;
  }
  double pl_sum = 0.;
  for (int p = 0; p < d->solid_count; p++) pl_sum += d->f[WFO_MASS][p] / d->f[WFO_RHO][p] * d->f[WFO_QPL][p];
  d->plastic_work += pl_sum * dt;
}

/* per-particle merge of the forward and backward Nishimura lists back into
 * global pair order p, i.e. the order an Nproc=1 pair loop visits them */
#define MERGE_BEGIN(d, i)                                                     \
  const int *ar_ = (d)->Aref + (int64_t)(i) * WFO_MAX_NB;                     \
  int nf_ = (d)->ipair[i], nb_ = (d)->jpair[i], f_ = 0, b_ = 0;              \
  while (f_ < nf_ || b_ < nb_) {                                             \
    int64_t p_;                                                               \
    if (b_ >= nb_ || (f_ < nf_ && ar_[f_] < ar_[WFO_MAX_NB - 1 - b_])) {      \
      p_ = ar_[f_++];                                                         \
    } else {                                                                  \
      p_ = ar_[WFO_MAX_NB - 1 - b_++];                                        \
    }
#define MERGE_END }

/* CalcConvHeat (Source/Thermal.cpp:339-389): dS2 = (m/rho)^(2/3) in 3D,
 * (m/rho)^(1/3) in plane strain (:354-366) */
static void conv_heat(wfo_domain *d) {
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int i = 0; i < d->n; i++) {
    double dens = d->f[WFO_RHO][i];
    double dS2 = pow(d->f[WFO_MASS][i] / dens, d->dim == 2 ? 1.0 / 3.0 : 2.0 / 3.0);
    if (d->fi[WFO_THBC][i] == 1)
      d->f[WFO_QCONV][i] = dens * d->f[WFO_HCONV][i] * dS2 * (d->f[WFO_TINF][i] - d->f[WFO_T][i]) / d->f[WFO_MASS][i];
  }
}

/* CalcTempInc without gradient correction (Source/Thermal.cpp:53-219):
 * temp[first] += m, temp[second] -= m with m using the second particle's
 * m/rho; accumulated per particle in pair order. */
static void temp_inc(wfo_domain *d) {
  const double *x = d->f[WFO_X], *T = d->f[WFO_T], *kT = d->f[WFO_KT];
  const double *ms = d->f[WFO_MASS], *rho = d->f[WFO_RHO], *hh = d->f[WFO_H];
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int i = 0; i < d->n; i++) {
    double acc = 0.0;
    MERGE_BEGIN(d, i)
      int p1 = d->pairs[2 * p_], p2 = d->pairs[2 * p_ + 1];
      double xij[3] = {x[3 * p1] - x[3 * p2], x[3 * p1 + 1] - x[3 * p2 + 1], x[3 * p1 + 2] - x[3 * p2 + 2]};
      double h = (hh[p1] + hh[p2]) / 2.0;
      double GK = wfo_grad_kernel_dim(norm3(xij) / h, h, d->dim);
      double mj = ms[p2], dj = rho[p2];
      double gv[3] = {GK * xij[0], GK * xij[1], GK * xij[2]};
      double m = mj / dj * 4. * (kT[p1] * kT[p2]) / (kT[p1] + kT[p2]) * (T[p1] - T[p2]) * dot3(xij, gv) /
                 (norm3(xij) * norm3(xij));
      if (p1 == i) acc += m; else acc -= m;
    MERGE_END
    d->temp_work[i] = acc;
  }
  for (int i = 0; i < d->solid_count; i++) {
    double dd = d->f[WFO_RHO][i];
    double f = 1. / (dd * d->f[WFO_CPT][i]);
    d->f[WFO_DTDT][i] = f * (d->temp_work[i] + d->f[WFO_QCONV][i] + d->f[WFO_QSRC][i] +
                             d->f[WFO_QPL][i] * d->pl_work_heat_frac + d->f[WFO_QCONT][i]);
    /* Source/Thermal.cpp:175-189 */
    if (d->contact && d->cont_heat_fric) d->f[WFO_DTDT][i] += f * d->f[WFO_QFRIC][i];
  }
  /* the contact heat totals (Source/Thermal.cpp:194-210) */
  double fr_temp = 0., cond_temp = 0.;
  if (d->contact)
    for (int i = 0; i < d->solid_count; i++) {
      const double dd = d->f[WFO_RHO][i];
      fr_temp += d->f[WFO_QFRIC][i] * d->f[WFO_MASS][i] / dd;
      if (d->cont_heat_cond) cond_temp += d->f[WFO_QCONT][i] * d->f[WFO_MASS][i] / dd;
    }
  for (int i = d->solid_count; i < d->n; i++) d->f[WFO_DTDT][i] = 0.;
  d->contact_friction_work += fr_temp * d->deltat;
  d->accum_cont_heat_cond += cond_temp * d->deltat;
}

/* ThermalCalcs (Source/Thermal.cpp:721-748) with CalcThermalExpStrainRate
 * (Source/ThermalStuct.cpp:7-12, Source/Particle.cpp:953-957) */
int wfo_thermal_calcs(wfo_domain *d, double dt) {
  if (!d->thermal_solver) return 0;
  d->m_maxT = 0.;
  d->m_minT = 1000.;
  for (int i = 0; i < d->solid_count; i++) {
    d->f[WFO_T][i] += dt * d->f[WFO_DTDT][i];
    if (d->f[WFO_T][i] > d->m_maxT) d->m_maxT = d->f[WFO_T][i];
    if (d->f[WFO_T][i] < d->m_minT) d->m_minT = d->f[WFO_T][i];
  }
  conv_heat(d);
  temp_inc(d);
  for (int p = 0; p < d->n; p++) {
    double s = d->f[WFO_THEXP][p] * d->f[WFO_DTDT][p];
    double *SR = d->f[WFO_SR] + 9 * (int64_t)p;
    for (int a = 0; a < 3; a++) SR[4 * a] = SR[4 * a] + 1.0 * s;
  }
  return 0;
}

/* CalcKinEnergyEqn (Source/Domain.cpp:1752-1808): the pair terms are
 * antisymmetric, so the sum is zero up to rounding; kept for fidelity. */
static void kin_energy(wfo_domain *d) {
  const double *x = d->f[WFO_X], *v = d->f[WFO_V], *P = d->f[WFO_P];
  const double *ms = d->f[WFO_MASS], *rho = d->f[WFO_RHO], *hh = d->f[WFO_H];
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int i = 0; i < d->n; i++) {
    double acc = 0.0;
    MERGE_BEGIN(d, i)
      int p1 = d->pairs[2 * p_], p2 = d->pairs[2 * p_ + 1];
      double xij[3], vij[3];
      for (int k = 0; k < 3; k++) { xij[k] = x[3 * p1 + k] - x[3 * p2 + k]; vij[k] = v[3 * p1 + k] - v[3 * p2 + k]; }
      double h = (hh[p1] + hh[p2]) / 2.0;
      double GK = wfo_grad_kernel_dim(norm3(xij) / h, h, d->dim);
      double di = rho[p1], dj = rho[p2], mj = ms[p2];
      double gv[3] = {GK * xij[0], GK * xij[1], GK * xij[2]};
      double temp = 0.5 * mj * (P[p1] / (di * di) + P[p2] / (dj * dj)) * dot3(vij, gv);
      if (p1 == i) acc += temp; else acc -= temp;
    MERGE_END
    d->dkin_work[i] = acc;
  }
  double inc = 0.;
  for (int i = 0; i < d->n; i++) inc += d->dkin_work[i];
  d->kin_energy_sum += inc * d->deltat;
}

/* CalcIntEnergyEqn (Source/Domain.cpp:1810-1829, Source/Particle.cpp:941-950) */
static void int_energy(wfo_domain *d) {
  double inc = 0.;
  const int max = d->contact ? d->mesh[0].first : d->n;
  for (int i = 0; i < max; i++) {
    const double *S = d->f[WFO_SHEAR] + 9 * (int64_t)i, *SR = d->f[WFO_SR] + 9 * (int64_t)i;
    double de = d->f[WFO_MASS][i] / d->f[WFO_RHO][i] *
                (S[0] * SR[0] + 2.0 * S[1] * SR[3] + 2.0 * S[2] * SR[6] + S[4] * SR[4] + 2.0 * S[5] * SR[7] + S[8] * SR[8]);
    inc += de;
  }
  d->int_energy_sum += inc * d->deltat;
}

int wfo_plastic_work_heat(wfo_domain *d, double dt) {
  plastic_work_heat(d, dt);
  return 0;
}

int wfo_energies(wfo_domain *d) {
  kin_energy(d);
  int_energy(d);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Rigid meshes and contact (Source/Mesh.cpp:70-300, Source/Contact.cpp:10-132,
 * Source/Contact_Wang.cpp:15-348, Source/Domain.cpp:1415-1491)        */

static void cross3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

/* TriMesh::CalcNormals, 3D (Source/Mesh.cpp:229-243) */
static void mesh_normals(mesh_t *M) {
  for (int e = 0; e < M->nelem; e++) {
    const double *n0 = M->node + 3 * M->el[3 * e], *n1 = M->node + 3 * M->el[3 * e + 1],
                 *n2 = M->node + 3 * M->el[3 * e + 2];
    double u[3], v[3], w[3];
    for (int k = 0; k < 3; k++) { u[k] = n1[k] - n0[k]; v[k] = n2[k] - n0[k]; }
    cross3(u, v, w);
    double nw = norm3(w);
    for (int k = 0; k < 3; k++) M->normal[3 * e + k] = w[k] / nw;
  }
}

/* TriMesh::UpdatePlaneCoeff (Source/Mesh.cpp:222-227) */
static void mesh_plane_coeff(mesh_t *M) {
  for (int e = 0; e < M->nelem; e++)
    M->pplane[e] = dot3(M->node + 3 * M->el[3 * e + M->nfar[e]], M->normal + 3 * e);
}

static int mesh_finish(wfo_domain *d, mesh_t *M, int first, int id, double hfac);

static mesh_t *mesh_alloc(wfo_domain *d, int id, int first, int nnodes, int nelem) {
  mesh_t *M = &d->mesh[d->nmesh];
  memset(M, 0, sizeof *M);
  M->id = id;
  M->first = first;
  M->nnodes = nnodes;
  M->nelem = nelem;
  M->node = (double *)xcalloc((size_t)3 * nnodes, sizeof(double));
  M->node_v = (double *)xcalloc((size_t)3 * nnodes, sizeof(double));
  M->el = (int *)xcalloc((size_t)3 * nelem, sizeof(int));
  M->centroid = (double *)xcalloc((size_t)3 * nelem, sizeof(double));
  M->normal = (double *)xcalloc((size_t)3 * nelem, sizeof(double));
  M->pplane = (double *)xcalloc((size_t)nelem, sizeof(double));
  M->radius = (double *)xcalloc((size_t)nelem, sizeof(double));
  M->nfar = (int *)xcalloc((size_t)nelem, sizeof(int));
  return M;
}

/* TriMesh(NastranReader &, flipnormals) (Source/Mesh.cpp:17-64), 3D, from
 * the nodes and triangles after the driver's flip / scale / translation:
 * centroid (n0 + n1 + n2) / 3, normal = (n1 - n0) x (n2 - n0) normalised
 * unless given; then CalcSpheres and AddTrimeshParticles */
int wfo_add_mesh(wfo_domain *d, int first, int id, int nnodes, const double *nodes, int nelem, const int *el,
                 const double *normals, double hfac) {
  if (d->nmesh >= WFO_MAX_MESH) return fail("too many meshes");
  if (nnodes < 3 || nelem < 1) return fail("empty mesh");
  if (first < 0 || first + nelem > d->n) return fail("rigid particles exceed the domain size");
  for (int k = 0; k < 3 * nelem; k++)
    if (el[k] < 0 || el[k] >= nnodes) return fail("element node out of range");
  mesh_t *M = mesh_alloc(d, id, first, nnodes, nelem);
  memcpy(M->node, nodes, sizeof(double) * 3 * (size_t)nnodes);
  memcpy(M->el, el, sizeof(int) * 3 * (size_t)nelem);
  for (int e = 0; e < nelem; e++) {
    const double *a = M->node + 3 * el[3 * e], *b = M->node + 3 * el[3 * e + 1], *c = M->node + 3 * el[3 * e + 2];
    for (int k = 0; k < 3; k++) M->centroid[3 * e + k] = (a[k] + b[k] + c[k]) / 3.;
  }
  /* the initial normals: given (the reference's file meshes start from the
   * unflipped winding) or from the winding */
  if (normals)
    memcpy(M->normal, normals, sizeof(double) * 3 * (size_t)nelem);
  else
    mesh_normals(M);
  return mesh_finish(d, M, first, id, hfac);
}

/* TriMesh::AxisPlaneMesh, 3D (Source/Mesh.cpp:82-202) followed by
 * CalcSpheres (:206-220, whose loop leaves nfar = the last node), then
 * Domain::AddTrimeshParticles (Source/Contact.cpp:10-37) with h =
 * hfac * radius, hfac passed as float. */
int wfo_add_plane_mesh(wfo_domain *d, int first, int id, int axis, int flip, const double p1[3], const double p2[3],
                       int dens, double hfac) {
  if (d->nmesh >= WFO_MAX_MESH) return fail("too many meshes");
  if (dens < 1) return fail("mesh density must be >= 1");
  const int nelem = 2 * dens * dens, nnodes = (dens + 1) * (dens + 1);
  if (first < 0 || first + nelem > d->n) return fail("rigid particles exceed the domain size");
  mesh_t *M = mesh_alloc(d, id, first, nnodes, nelem);
  double p[3];
  for (int k = 0; k < 3; k++) p[k] = p2[k] - p1[k];
  int dir[3];
  if (axis == 0) { dir[0] = 1; dir[1] = 2; }
  else if (axis == 1) { dir[0] = 0; dir[1] = 2; }
  else { dir[0] = 0; dir[1] = 1; }
  dir[2] = axis;
  double x3 = p1[dir[2]], x2 = p1[dir[1]];
  double dl = p[dir[0]] / dens;
  int nn = 0;
  for (int j = 0; j < dens + 1; j++) {
    double x1 = p1[dir[0]];
    for (int i = 0; i < dens + 1; i++) {
      double v[3];
      v[dir[0]] = x1; v[dir[1]] = x2; v[dir[2]] = x3;
      for (int k = 0; k < 3; k++) M->node[3 * nn + k] = v[k];
      nn++;
      x1 += dl;
    }
    x2 += dl;
  }
  int el = 0;
  for (int j = 0; j < dens; j++)
    for (int i = 0; i < dens; i++) {
      int n[4];
      n[0] = (dens + 1) * j + i; n[1] = n[0] + 1;
      n[2] = (dens + 1) * (j + 1) + i; n[3] = n[2] + 1;
      int con[2][3];
      if (flip) {
        con[0][0] = n[0]; con[0][1] = n[1]; con[0][2] = n[2];
        con[1][0] = n[1]; con[1][1] = n[3]; con[1][2] = n[2];
      } else {
        con[0][0] = n[0]; con[0][1] = n[2]; con[0][2] = n[1];
        con[1][0] = n[1]; con[1][1] = n[2]; con[1][2] = n[3];
      }
      for (int e = 0; e < 2; e++) {
        for (int k = 0; k < 3; k++) M->el[3 * el + k] = con[e][k];
        const double *a = M->node + 3 * con[e][0], *b = M->node + 3 * con[e][1], *c = M->node + 3 * con[e][2];
        for (int k = 0; k < 3; k++) M->centroid[3 * el + k] = (a[k] + b[k] + c[k]) / 3.;
        el++;
      }
    }
  mesh_normals(M);
  if (flip)
    for (int k = 0; k < 3 * nelem; k++) M->normal[k] = -M->normal[k];
  return mesh_finish(d, M, first, id, hfac);
}

/* CalcSpheres (Source/Mesh.cpp:206-220: nfar = the last node) and
 * Domain::AddTrimeshParticles (Source/Contact.cpp:10-37) for a mesh whose
 * nodes, elements, centroids and normals are set */
static int mesh_finish(wfo_domain *d, mesh_t *M, int first, int id, double hfac) {
  const int nelem = M->nelem;
  for (int e = 0; e < nelem; e++) {
    double mx = 0.;
    for (int q = 0; q < 3; q++) {
      double rv[3];
      for (int k = 0; k < 3; k++) rv[k] = M->node[3 * M->el[3 * e + q] + k] - M->centroid[3 * e + k];
      if (norm3(rv) > mx) mx = norm3(rv);
      M->nfar[e] = q;
    }
    M->radius[e] = mx;
  }
  mesh_plane_coeff(M);
  const double hf = (double)(float)hfac;
  for (int e = 0; e < nelem; e++) {
    int r = first + e;
    for (int k = 0; k < 3; k++) {
      d->f[WFO_X][3 * r + k] = M->centroid[3 * e + k];
      d->f[WFO_V][3 * r + k] = 0.0;
      d->f[WFO_NORMAL][3 * r + k] = M->normal[3 * e + k];
    }
    d->f[WFO_H][r] = hf * M->radius[e];
    d->f[WFO_MASS][r] = 0.0;
    d->f[WFO_RHO][r] = d->f[WFO_REFRHO][r] = 0.0;
    d->fi[WFO_ID][r] = id;
    d->fi[WFO_FREE][r] = 1;
    d->fi[WFO_ELEMENT][r] = e;
    d->fi[WFO_MESHIDX][r] = d->nmesh;
  }
  d->nmesh++;
  d->contact = 1;
  return nelem;
}

/* Domain::CalculateSurface (Source/Domain.cpp:1415-1491): mass-weighted
 * sum of x_ij over SMPairs, scaled by the mean particle mass over ALL
 * particles (rigid ones count with mass 0); |n| >= h/4 and Nb <= 46 (12
 * in 2D, :1472-1474) marks a free-surface particle, which takes ID 1. */
static int calculate_surface(wfo_domain *d, int id) {
  d->id_free_surf = id;
  double totmass = 0.;
  for (int i = 0; i < d->n; i++) totmass += d->f[WFO_MASS][i];
  totmass /= d->n;
  const int maxid = d->contact ? d->mesh[0].first : d->n;
  double *nrm = d->f[WFO_NORMAL];
  const double *x = d->f[WFO_X], *m = d->f[WFO_MASS];
  for (int i = 0; i < maxid; i++) {
    for (int k = 0; k < 3; k++) nrm[3 * i + k] = 0.;
    d->fi[WFO_ID][i] = d->fi[WFO_IDORIG][i];
  }
  for (int64_t p = 0; p < d->npairs; p++) {
    int a = d->pairs[2 * p], b = d->pairs[2 * p + 1];
    for (int k = 0; k < 3; k++) {
      double xij = x[3 * a + k] - x[3 * b + k];
      nrm[3 * a + k] += m[b] * xij;
      nrm[3 * b + k] -= m[a] * xij;
    }
  }
  int surf = 0;
  for (int i = 0; i < maxid; i++) {
    for (int k = 0; k < 3; k++) nrm[3 * i + k] *= 1. / totmass;
    if (norm3(nrm + 3 * i) >= 0.25 * d->f[WFO_H][i] && d->fi[WFO_NB][i] <= (d->dim == 2 ? 12 : 46)) {
      d->fi[WFO_ID][i] = id;
      surf++;
    }
  }
  if (surf == 0) return fail("ERROR: No external particles found. Please check particle masses");
  return 0;
}

/* Domain::ContactNbSearch (Source/Contact.cpp:57-132): RIGPairs whose
 * particles are one rigid and one free-surface particle closer than
 * h1 + h2 become contact pairs.  They are then ordered by (solid particle,
 * rigid particle): the reference visits them in RIGPairs order, which
 * differs only where one particle has two contact points inside
 * triangles (a point on a shared edge). */
static int cmp_pair(const void *a, const void *b) {
  const int *x = (const int *)a, *y = (const int *)b;
  if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
  if (x[1] != y[1]) return x[1] < y[1] ? -1 : 1;
  return 0;
}
static void contact_nb_search(wfo_domain *d) {
  const int *id = d->fi[WFO_ID];
  const double *x = d->f[WFO_X], *h = d->f[WFO_H];
  for (int64_t a = 0; a < d->nrig; a++) {
    int p1 = d->rig[2 * a], p2 = d->rig[2 * a + 1];
    if (!(is_mesh_id(d, id[p1]) || is_mesh_id(d, id[p2]))) continue;
    if (!(id[p1] == d->id_free_surf || id[p2] == d->id_free_surf)) continue;
    double xij[3] = {x[3 * p1] - x[3 * p2], x[3 * p1 + 1] - x[3 * p2 + 1], x[3 * p1 + 2] - x[3 * p2 + 2]};
    if (norm3(xij) < h[p1] + h[p2]) {
      /* store as (SPH particle, rigid particle) */
      int s = is_mesh_id(d, id[p1]) ? p2 : p1, r = is_mesh_id(d, id[p1]) ? p1 : p2;
      push2(&d->cont, &d->ncont, &d->capcont, s, r);
    }
  }
  d->nrig = 0;
  qsort(d->cont, (size_t)d->ncont, 2 * sizeof(int), cmp_pair);
}

/* ContactNbUpdate (Source/Domain.cpp:2481-2487) */
static int contact_nb_update(wfo_domain *d) {
  if (calculate_surface(d, 1)) return -1;
  contact_nb_search(d);
  return 0;
}

/* The Coulomb coefficients of the particle's pairs: the domain's
 * friction_sta / friction_dyn, or its material's; with friction_function
 * Linear both are friction_m T + friction_b (Source/Contact_Wang.cpp:241-244,
 * Source/Contact_LS.cpp:217-220) */
static void pair_friction(const wfo_domain *d, int P1, double *fs, double *fd) {
  const double *mfr = d->mat_fr[d->fi[WFO_MATIDX][P1]];
  *fs = mfr[0] >= 0. ? mfr[0] : d->friction_sta;
  *fd = mfr[0] >= 0. ? mfr[1] : d->friction_dyn;
  if (d->friction_linear) *fs = *fd = d->friction_m * d->f[WFO_T][P1] + d->friction_b;
}

/* the inside test of the contact point qj against triangle e of mesh M
 * (Fraser 3-147: every edge's cross product with qj - node_i along the
 * normal is >= 0) */
static int inside_tri(const mesh_t *M, int e, const double *qj, const double *n2) {
  for (int ii = 0; ii < 3; ii++) {
    int jj = ii + 1;
    if (jj > 2) jj = 0;
    const double *ni = M->node + 3 * M->el[3 * e + ii], *nj = M->node + 3 * M->el[3 * e + jj];
    double ea[3], eb[3], c[3];
    for (int k = 0; k < 3; k++) { ea[k] = nj[k] - ni[k]; eb[k] = qj[k] - ni[k]; }
    cross3(ea, eb, c);
    if (dot3(c, n2) < 0.0) return 0;
  }
  return 1;
}

/* Domain::CalcContactForces, Fraser's law (Source/Contact.cpp:223-554),
 * 3D meshes.  Only pairs approaching (delta_ = -n.v_r > 0) whose time to
 * contact deltat_cont = (h + pplane - n.x) / (-delta_) is below the step
 * and whose contact point x + v deltat_cont - h n lies inside the triangle
 * count; the penetration is (deltat - deltat_cont) delta_ and the normal
 * force (PFAC cont_stiff delta - psi delta_) n is ASSIGNED to contforce.
 * Under Leapfrog the force is not added to a (the solver's loop that would
 * add it is commented out, Source/SolverLeapfrog.cpp:244-247): contforce
 * and contact_force_sum are outputs.  Friction: with fric_type Fr_Bound the
 * tangential part of a is removed; else a sliding pair (|v_t| > 1e-3)
 * ASSIGNS contforce = -friction_dyn |contforce| t / m.  The static branch
 * reads a tgforce that the reference never sets before it (an
 * uninitialised variable carried across pairs); it is taken as zero here,
 * which makes that branch add nothing.  Domain::min_force_ts takes the
 * running minimum of dt_fext whenever it drops to a positive value. */
static void contact_forces_fraser(wfo_domain *d) {
  const double dt = d->deltat;
  const double *x = d->f[WFO_X], *v = d->f[WFO_V], *h = d->f[WFO_H], *ms = d->f[WFO_MASS], *nrmv = d->f[WFO_NORMAL];
  double *a = d->f[WFO_A], *cf = d->f[WFO_CONTFORCE];
  for (int i = 0; i < 3 * d->n; i++) cf[i] = 0.;
  double min_force_ts_ = 1000.;
  d->contact_force_sum = 0.;
  for (int64_t q = 0; q < d->ncont; q++) {
    const int P1 = d->cont[2 * q], P2 = d->cont[2 * q + 1];
    const double *n2 = nrmv + 3 * P2;
    double vr[3];
    for (int k = 0; k < 3; k++) vr[k] = v[3 * P1 + k] - v[3 * P2 + k];
    const double delta_ = -dot3(n2, vr);
    if (!(delta_ > 0)) continue;
    const mesh_t *M = &d->mesh[d->fi[WFO_MESHIDX][P2]];
    const int e = d->fi[WFO_ELEMENT][P2];
    const double deltat_cont = (h[P1] + M->pplane[e] - dot3(n2, x + 3 * P1)) / (-delta_);
    if (!(deltat_cont < dt)) continue;
    double qj[3];
    for (int k = 0; k < 3; k++) qj[k] = x[3 * P1 + k] + (v[3 * P1 + k] * deltat_cont) - (h[P1] * n2[k]);
    if (!inside_tri(M, e, qj, n2)) continue;
    const double m1 = ms[P1];
    const double delta = (dt - deltat_cont) * delta_;
    const double kij = d->PFAC * d->cont_stiff[P1];
    const double omega = sqrt(kij / m1);
    const double psi_cont = 2. * m1 * omega * d->DFAC;
    const double normal_cf = (kij * delta - psi_cont * delta_);
    for (int k = 0; k < 3; k++) cf[3 * P1 + k] = normal_cf * n2[k];
    d->contact_force_sum += norm3(cf + 3 * P1);
    double tgvr[3], atg[3];
    for (int k = 0; k < 3; k++) tgvr[k] = vr[k] + delta_ * n2[k];
    const double norm_tgvr = norm3(tgvr);
    const double an = dot3(a + 3 * P1, n2);
    for (int k = 0; k < 3; k++) atg[k] = a[3 * P1 + k] - an * n2[k];
    const double dt_fext = 1. * (m1 * 2. * norm3(v + 3 * P1) / norm3(cf + 3 * P1));
    if (dt_fext < min_force_ts_) {
      min_force_ts_ = dt_fext;
      if (dt_fext > 0) d->min_force_ts = min_force_ts_;
    }
    if (d->fric_bound) {
      for (int k = 0; k < 3; k++) a[3 * P1 + k] -= atg[k];
    } else if (d->friction_dyn > 0. && norm_tgvr > 1.e-3) {
      const double ncf = norm3(cf + 3 * P1);
      for (int k = 0; k < 3; k++) {
        const double tgforce = d->friction_dyn * ncf * (tgvr[k] / norm_tgvr);
        cf[3 * P1 + k] = -tgforce / m1;
      }
    }
  }
}

/* Domain::CalcContactForces2, Seo's law (Source/Contact.cpp:562-770):
 * penetration of the current position, dist = n.x - pplane < h, contact
 * point x - dist n inside the triangle, normal force
 * (PFAC cont_stiff (h - dist) - psi delta_) n ASSIGNED to contforce (not
 * added to a).  Static friction (ref_accel): if |m a_t| < friction_sta
 * |contforce|, contforce -= PFAC cont_stiff (v_r dt)_t / m.  No dynamic
 * branch.  min_force_ts as in Fraser's law. */
static void contact_forces_seo(wfo_domain *d) {
  const double dt = d->deltat;
  const double *x = d->f[WFO_X], *v = d->f[WFO_V], *h = d->f[WFO_H], *ms = d->f[WFO_MASS], *nrmv = d->f[WFO_NORMAL];
  const double *a = d->f[WFO_A];
  double *cf = d->f[WFO_CONTFORCE];
  for (int i = 0; i < 3 * d->n; i++) cf[i] = 0.;
  double min_force_ts_ = 1000.;
  d->contact_force_sum = 0.;
  for (int64_t q = 0; q < d->ncont; q++) {
    const int P1 = d->cont[2 * q], P2 = d->cont[2 * q + 1];
    const double *n2 = nrmv + 3 * P2;
    double vr[3];
    for (int k = 0; k < 3; k++) vr[k] = v[3 * P1 + k] - v[3 * P2 + k];
    const double delta_ = -dot3(n2, vr);
    const mesh_t *M = &d->mesh[d->fi[WFO_MESHIDX][P2]];
    const int e = d->fi[WFO_ELEMENT][P2];
    const double dist = dot3(n2, x + 3 * P1) - M->pplane[e];
    if (!(dist < h[P1])) continue;
    double qj[3];
    for (int k = 0; k < 3; k++) qj[k] = x[3 * P1 + k] - dist * n2[k];
    if (!inside_tri(M, e, qj, n2)) continue;
    const double m1 = ms[P1];
    const double delta = h[P1] - dist;
    const double kij = d->PFAC * d->cont_stiff[P1];
    const double omega = sqrt(kij / m1);
    const double psi_cont = 2. * m1 * omega * d->DFAC;
    for (int k = 0; k < 3; k++) cf[3 * P1 + k] = (kij * delta - psi_cont * delta_) * n2[k];
    d->contact_force_sum += norm3(cf + 3 * P1);
    double atg[3];
    const double an = dot3(a + 3 * P1, n2);
    for (int k = 0; k < 3; k++) atg[k] = a[3 * P1 + k] - an * n2[k];
    const double dt_fext = 1. * (m1 * 2. * norm3(v + 3 * P1) / norm3(cf + 3 * P1));
    if (dt_fext < min_force_ts_) {
      min_force_ts_ = dt_fext;
      if (dt_fext > 0) d->min_force_ts = min_force_ts_;
    }
    double du[3], dtg[3], tgforce[3], ref_tg[3];
    for (int k = 0; k < 3; k++) du[k] = vr[k] * dt;
    const double dn = dot3(du, n2);
    for (int k = 0; k < 3; k++) dtg[k] = du[k] - dn * n2[k];
    for (int k = 0; k < 3; k++) tgforce[k] = kij * dtg[k];
    for (int k = 0; k < 3; k++) ref_tg[k] = atg[k] * m1;
    if (d->friction_sta > 0.)
      if (norm3(ref_tg) < d->friction_sta * norm3(cf + 3 * P1))
        for (int k = 0; k < 3; k++) cf[3 * P1 + k] -= tgforce[k] / m1;
  }
}

/* Domain::CalcContactForcesLS (Source/Contact_LS.cpp:15-292), 3D meshes:
 * the Wang geometry (predicted position, contact point from the current
 * one), penalty 2 m / dt^2 PFAC, normal force ASSIGNED to contforce (not
 * added to a); the contact heat fields are reset.  Friction: the trial
 * force fstar = tgforce_old - PFAC m v_r / dt, returned to the cone
 * friction_sta |contforce| and added; Particle::tgforce, the old force,
 * is never assigned by the reference (read uninitialised), taken as zero
 * here.  contact_reaction_sum and the external work as in Wang's law;
 * Domain::min_force_ts follows the running minimum of dt_fext. */
static void contact_forces_ls(wfo_domain *d) {
  const double dt = d->deltat;
  const double *x = d->f[WFO_X], *v = d->f[WFO_V], *h = d->f[WFO_H], *ms = d->f[WFO_MASS], *nrmv = d->f[WFO_NORMAL];
  const double *a = d->f[WFO_A];
  double *cf = d->f[WFO_CONTFORCE];
  double *qfric = d->f[WFO_QFRIC], *cshear = d->f[WFO_CSHEAR], *fhfl = d->f[WFO_FHFL], *qcont = d->f[WFO_QCONT];
  for (int i = 0; i < d->n; i++) {
    for (int k = 0; k < 3; k++) cf[3 * i + k] = 0.;
    qfric[i] = cshear[i] = fhfl[i] = qcont[i] = 0.;
  }
  double min_force_ts_ = 1000.;
  d->contact_force_sum = 0.;
  d->contact_reaction_sum = 0.;
  double ext_step = 0.;
  for (int64_t q = 0; q < d->ncont; q++) {
    const int P1 = d->cont[2 * q], P2 = d->cont[2 * q + 1];
    const double *n2 = nrmv + 3 * P2;
    double vr[3];
    for (int k = 0; k < 3; k++) vr[k] = v[3 * P1 + k] - v[3 * P2 + k];
    const double delta_ = -dot3(n2, vr);
    const mesh_t *M = &d->mesh[d->fi[WFO_MESHIDX][P2]];
    const int e = d->fi[WFO_ELEMENT][P2];
    double x_pred[3];
    for (int k = 0; k < 3; k++) x_pred[k] = x[3 * P1 + k] + v[3 * P1 + k] * dt + a[3 * P1 + k] * dt * dt / 2.0;
    const double dist = dot3(n2, x_pred) - M->pplane[e];
    if (!(dist < h[P1])) continue;
    double qj[3];
    for (int k = 0; k < 3; k++) qj[k] = x[3 * P1 + k] - dist * n2[k];
    if (!inside_tri(M, e, qj, n2)) continue;
    const double m1 = ms[P1];
    const double delta = h[P1] - dist;
    const double kij = 2.0 * m1 / (dt * dt) * d->PFAC;
    const double omega = sqrt(kij / m1);
    const double psi_cont = 2. * m1 * omega * d->DFAC;
    for (int k = 0; k < 3; k++) cf[3 * P1 + k] = (kij * delta - psi_cont * delta_) * n2[k];
    const double dt_fext = 1. * (m1 * 2. * norm3(v + 3 * P1) / norm3(cf + 3 * P1));
    if (dt_fext < min_force_ts_) {
      min_force_ts_ = dt_fext;
      d->min_force_ts = min_force_ts_;
    }
    double fr_sta, fr_dyn;
    pair_friction(d, P1, &fr_sta, &fr_dyn);
    if (fr_sta > 0.) {
      double kdeltae[3], fstar[3], fT[3];
      for (int k = 0; k < 3; k++) kdeltae[k] = d->PFAC * m1 * vr[k] / dt;
      const double fy = fr_sta * norm3(cf + 3 * P1);
      for (int k = 0; k < 3; k++) fstar[k] = 0.0 - kdeltae[k];
      const double nfs = norm3(fstar);
      if (nfs > fy)
        for (int k = 0; k < 3; k++) fT[k] = fy * fstar[k] / nfs;
      else
        for (int k = 0; k < 3; k++) fT[k] = fstar[k];
      for (int k = 0; k < 3; k++) cf[3 * P1 + k] += fT[k];
    }
    d->contact_force_sum += norm3(cf + 3 * P1);
    d->contact_reaction_sum += dot3(a + 3 * P1, n2) * m1;
    ext_step += dot3(cf + 3 * P1, v + 3 * P2);
  }
  d->ext_forces_work += ext_step * dt;
}

/* the law Domain::contact_alg selects (Source/SolverLeapfrog.cpp:234-239;
 * SolveDiffUpdateKickDrift always calls Wang's, Source/SolverKickDrift.cpp:211) */
static void contact_forces(wfo_domain *d);
static void contact_law(wfo_domain *d, int leapfrog) {
  if (!leapfrog || d->contact_alg == 1) contact_forces(d);
  else if (d->contact_alg == 0) contact_forces_fraser(d);
  else if (d->contact_alg == 2) contact_forces_seo(d);
  else if (d->contact_alg == 4) contact_forces_ls(d);
}

/* Domain::CalcContactForcesWang (Source/Contact_Wang.cpp:15-348), 3D,
 * no contact heat: penalty kij = 2 m / dt^2 on the penetration of the
 * predicted position into the triangle's plane, damping DFAC, Coulomb
 * friction on the predicted tangential slip.  contforce is assigned per
 * pair while a accumulates; the static branch subtracts tgforce / m from
 * contforce (both as in the reference). */
static void contact_forces(wfo_domain *d) {
  const double dt = d->deltat;
  double *x = d->f[WFO_X], *v = d->f[WFO_V], *a = d->f[WFO_A], *cf = d->f[WFO_CONTFORCE];
  const double *h = d->f[WFO_H], *ms = d->f[WFO_MASS], *nrmv = d->f[WFO_NORMAL];
  double *qfric = d->f[WFO_QFRIC], *cshear = d->f[WFO_CSHEAR], *fhfl = d->f[WFO_FHFL], *qcont = d->f[WFO_QCONT];
  const double *T = d->f[WFO_T], *rho = d->f[WFO_RHO];
  /* the reset of every particle (Source/Contact_Wang.cpp:29-41) */
  for (int i = 0; i < d->n; i++) {
    for (int k = 0; k < 3; k++) cf[3 * i + k] = 0.;
    qfric[i] = cshear[i] = fhfl[i] = qcont[i] = 0.;
  }
  /* min_force_ts_ (Source/Contact_Wang.cpp:17): the running minimum of
   * the pairs' dt_fext, which Domain::min_force_ts follows whenever it
   * drops */
  double min_force_ts_ = 1000.;
  memset(d->m_contact_force, 0, sizeof d->m_contact_force);
  memset(d->g_contact_force, 0, sizeof d->g_contact_force);
  d->contact_force_sum = 0.;
  d->contact_reaction_sum = 0.;
  double ext_step = 0.;
  for (int64_t q = 0; q < d->ncont; q++) {
    const int P1 = d->cont[2 * q], P2 = d->cont[2 * q + 1];
    const double *n2 = nrmv + 3 * P2;
    double vr[3];
    for (int k = 0; k < 3; k++) vr[k] = v[3 * P1 + k] - v[3 * P2 + k];
    const double delta_ = -dot3(n2, vr);
    const int m = d->fi[WFO_MESHIDX][P2];
    const mesh_t *M = &d->mesh[m];
    const int e = d->fi[WFO_ELEMENT][P2];
    double x_pred[3];
    for (int k = 0; k < 3; k++) x_pred[k] = x[3 * P1 + k] + v[3 * P1 + k] * dt + a[3 * P1 + k] * dt * dt / 2.0;
    const double dist = dot3(n2, x_pred) - M->pplane[e];
    if (!(dist < h[P1])) continue;
    double qj[3];
    for (int k = 0; k < 3; k++) qj[k] = x[3 * P1 + k] - dist * n2[k];
    int inside = 1;
    for (int ii = 0; ii < 3 && inside; ii++) {
      int jj = ii + 1;
      if (jj > 2) jj = 0;
      const double *ni = M->node + 3 * M->el[3 * e + ii], *nj = M->node + 3 * M->el[3 * e + jj];
      double ea[3], eb[3], c[3];
      for (int k = 0; k < 3; k++) { ea[k] = nj[k] - ni[k]; eb[k] = qj[k] - ni[k]; }
      cross3(ea, eb, c);
      if (dot3(c, n2) < 0.0) inside = 0;
    }
    if (!inside) continue;
    const double m1 = ms[P1];
    const double delta = h[P1] - dist;
    const double kij = 2.0 * m1 / (dt * dt);
    const double omega = sqrt(kij / m1);
    const double psi_cont = 2. * m1 * omega * d->DFAC;
    for (int k = 0; k < 3; k++) cf[3 * P1 + k] = (kij * delta - psi_cont * delta_) * n2[k];
    /* dt_fext with contact_force_factor 1 (Source/Contact_Wang.cpp:226-233),
     * on the normal force before the friction */
    const double dt_fext = 1. * (m1 * 2. * norm3(v + 3 * P1) / norm3(cf + 3 * P1));
    if (dt_fext < min_force_ts_) {
      min_force_ts_ = dt_fext;
      d->min_force_ts = min_force_ts_;
    }
    for (int k = 0; k < 3; k++) a[3 * P1 + k] += cf[3 * P1 + k] / m1;
    /* Fraser 3-119 (Source/Contact_Wang.cpp:245-247): the contact area */
    const double dens = rho[P1];
    const double dS2 = pow(m1 / dens, 2.0 / 3.0);
    double fr_sta, fr_dyn;
    pair_friction(d, P1, &fr_sta, &fr_dyn);
    if (fr_sta > 0.) {
      double du[3], dtg[3], tgf[3];
      for (int k = 0; k < 3; k++) du[k] = x_pred[k] - x[3 * P1 + k] - v[3 * P2 + k] * dt;
      const double dn = dot3(du, n2);
      for (int k = 0; k < 3; k++) dtg[k] = du[k] - dn * n2[k];
      for (int k = 0; k < 3; k++) tgf[k] = kij * dtg[k];
      if (norm3(tgf) < fr_sta * norm3(cf + 3 * P1)) {
        for (int k = 0; k < 3; k++) {
          cf[3 * P1 + k] -= tgf[k] / m1;
          a[3 * P1 + k] -= tgf[k] / m1;
        }
      } else {
        const double ncf = norm3(cf + 3 * P1), ntg = norm3(tgf);
        double tgd[3];
        for (int k = 0; k < 3; k++) tgd[k] = fr_dyn * ncf * tgf[k] / ntg;
        for (int k = 0; k < 3; k++) {
          cf[3 * P1 + k] -= tgd[k];
          a[3 * P1 + k] -= tgd[k] / m1;
        }
        /* friction work as heat in the sliding branch
         * (Source/Contact_Wang.cpp:291-300) */
        if (d->cont_heat_fric) {
          const double abs_fv = fabs(dot3(tgd, vr));
          qfric[P1] = abs_fv * dens / m1;
          cshear[P1] = norm3(tgd) / dS2;
          fhfl[P1] = abs_fv / dS2;
        }
      }
    }
    d->contact_force_sum += norm3(cf + 3 * P1);
    for (int k = 0; k < 3; k++) d->m_contact_force[m][k] += cf[3 * P1 + k];
    for (int k = 0; k < 3; k++) d->g_contact_force[d->fi[WFO_MATIDX][P1]][m][k] += cf[3 * P1 + k];
    d->contact_reaction_sum += dot3(a + 3 * P1, n2) * m1;
    ext_step += dot3(cf + 3 * P1, v + 3 * P2);
    /* contact conductance, Fraser 3.121 (Source/Contact_Wang.cpp:320-324) */
    if (d->cont_heat_cond) qcont[P1] = dens * d->contact_hc * dS2 * (T[P2] - T[P1]) / m1;
  }
  d->ext_forces_work += ext_step * dt;
}

/* TriMesh::Update with m_w = 0 (Source/Mesh.cpp:273-296) and
 * Domain::UpdateContactParticles (Source/Contact.cpp:41-55) */
static void mesh_update(wfo_domain *d) {
  const double dt = d->deltat;
  for (int m = 0; m < d->nmesh; m++) {
    mesh_t *M = &d->mesh[m];
    /* TriMesh::Update (Source/Mesh.cpp:273-296): node_v = m_v + m_w x node,
     * taken at the node's position before it moves */
    for (int nd = 0; nd < M->nnodes; nd++) {
      double vr[3];
      cross3(M->mw, M->node + 3 * nd, vr);
      for (int k = 0; k < 3; k++) {
        M->node_v[3 * nd + k] = M->mv[k] + vr[k];
        M->node[3 * nd + k] += M->node_v[3 * nd + k] * dt;
      }
    }
    for (int e = 0; e < M->nelem; e++)
      for (int k = 0; k < 3; k++) {
        double c = 0.;
        for (int q = 0; q < 3; q++) c += M->node[3 * M->el[3 * e + q] + k];
        M->centroid[3 * e + k] = c / 3;
      }
    mesh_normals(M);
    mesh_plane_coeff(M);
    for (int e = 0; e < M->nelem; e++) {
      int r = M->first + e;
      for (int k = 0; k < 3; k++) {
        double s = 0.;
        for (int q = 0; q < 3; q++) s += M->node_v[3 * M->el[3 * e + q] + k];
        d->f[WFO_V][3 * r + k] = s / 3;
        d->f[WFO_A][3 * r + k] = 0.;
        d->f[WFO_NORMAL][3 * r + k] = M->normal[3 * e + k];
      }
    }
  }
}

int wfo_contact_force(const wfo_domain *d, int mesh, double out[3]) {
  if (mesh < 0 || mesh >= d->nmesh) return fail("mesh index out of range");
  for (int k = 0; k < 3; k++) out[k] = d->m_contact_force[mesh][k];
  return 0;
}
int wfo_group_contact_force(const wfo_domain *d, int group, int mesh, double out[3]) {
  if (mesh < 0 || mesh >= d->nmesh) return fail("mesh index out of range");
  if (group < 0 || group >= 256) return fail("group index out of range");
  for (int k = 0; k < 3; k++) out[k] = d->g_contact_force[group][mesh][k];
  return 0;
}
int64_t wfo_ncontact_pairs(const wfo_domain *d) { return d->ncont; }
const int *wfo_contact_pairs(const wfo_domain *d) { return d->cont; }

/* ------------------------------------------------------------------ */
/* Solver set-up (Source/SolverLeapfrog.cpp:24-103)                     */
/* CheckMinTSVel (Source/Domain.cpp:265-299): smallest neighbour distance of
 * each particle over its Nishimura lists, dt >= CFL d_min / (Cs + |v|) */
static void check_min_ts_vel(wfo_domain *d) {
  const double *x = d->f[WFO_X], *v = d->f[WFO_V], *cs = d->f[WFO_CS];
  d->deltatmin = d->deltatint;
  for (int i = 0; i < d->n; i++) {
    double mn = 1000.;
    for (int k = 0; k < d->ipair[i] + d->jpair[i]; k++) {
      int j = k < d->ipair[i] ? d->Anei[(int64_t)i * WFO_MAX_NB + k]
                              : d->Anei[(int64_t)i * WFO_MAX_NB + WFO_MAX_NB - 1 - (k - d->ipair[i])];
      double xij[3] = {x[3 * j] - x[3 * i], x[3 * j + 1] - x[3 * i + 1], x[3 * j + 2] - x[3 * i + 2]};
      double dd = norm3(xij);
      if (dd < mn) mn = dd;
    }
    double test = d->CFL * mn / (cs[i] + norm3(v + 3 * i));
    if (d->deltatmin > test) d->deltatmin = test;
  }
}

/* CheckMinTSAccel (Source/Domain.cpp:304-319): dt >= sqrt_h_a sqrt(h / |a|) */
static void check_min_ts_accel(wfo_domain *d) {
  d->deltatmin = d->deltatint;
  for (int i = 0; i < d->n; i++) {
    double test = d->sqrt_h_a * sqrt(d->f[WFO_H][i] / norm3(d->f[WFO_A] + 3 * i));
    if (d->deltatmin > test) d->deltatmin = test;
  }
}

/* AdaptiveTimeStep (Source/Domain.cpp:206-239); the contact criterion
 * (auto_ts bit 2) takes min_force_ts when it is the smaller */
static void adaptive_time_step(wfo_domain *d) {
  if (d->deltatint > d->deltatmin) {
    if (d->deltat < d->deltatmin) d->deltat = 2.0 * d->deltat * d->deltatmin / (d->deltat + d->deltatmin);
    else d->deltat = d->deltatmin;
  } else {
    if (d->deltatint != d->deltat) d->deltat = 2.0 * d->deltat * d->deltatint / (d->deltat + d->deltatint);
    else d->deltat = d->deltatint;
  }
  if (d->contact && (d->auto_ts & 4) && d->min_force_ts < d->deltat) d->deltat = d->min_force_ts;
  if (d->deltat < d->manual_min_ts) d->deltat = d->manual_min_ts;
}

/* the time-step update of both solvers: prev_deltat = deltat, then the
 * velocity and/or acceleration check (the later one wins) */
static double update_time_step(wfo_domain *d) {
  double prev = d->deltat;
  if (d->auto_ts & 1) check_min_ts_vel(d);
  if (d->auto_ts & 2) check_min_ts_accel(d);
  if (d->auto_ts & 3) adaptive_time_step(d);
  return prev;
}

int wfo_solve_init(wfo_domain *d, double dt) {
  if (d->n <= 0) return fail("Particle Count is Null");
  if (d->dim == 2) {
    /* plane strain: the 2D tensile-instability branch and the 2D contact
     * (whose dS2 reads another particle, Source/Contact_Wang.cpp:259) are
     * not restated */
    if (d->contact) return fail("2D: rigid contact is not restated");
    for (int i = 0; i < d->n; i++)
      if (d->f[WFO_TI][i] > 0.0) return fail("2D: the tensile-instability term is not restated");
  }
  d->deltat = dt;
  d->deltatint = d->deltatmin = dt;
  initial_checks(d);
  memset(d->TRPR, 0, sizeof d->TRPR);
  memset(d->BLPF, 0, sizeof d->BLPF);
  if (cell_initiate(d)) return -1;
  list_generate(d);
  if (d->contact) {
    /* Source/SolverLeapfrog.cpp:80-100: ID_orig, one search, Nb counts
     * and the free surface before the loop */
    for (int i = 0; i < d->n; i++) d->fi[WFO_IDORIG][i] = d->fi[WFO_ID][i];
    d->min_force_ts = dt;   /* Source/SolverLeapfrog.cpp:95 */
    if (d->contact_alg == 0 || d->contact_alg == 2) {
      /* Particle::cont_stiff: SolveDiffUpdateLeapFrog never sets it; the
       * formula SolveDiffUpdateKickDrift applies before its loop
       * (Source/SolverKickDrift.cpp:79-84, Fraser 3-153) */
      free(d->cont_stiff);
      d->cont_stiff = (double *)xcalloc((size_t)d->n, sizeof(double));
      for (int i = 0; i < d->n; i++) {
        const double rho = d->f[WFO_RHO][i], Cs = d->f[WFO_CS][i], G = d->f[WFO_G][i];
        const double bulk = Cs * Cs * rho;
        const double dS = pow(d->f[WFO_MASS][i] / rho, 0.33333);
        d->cont_stiff[i] = 9. * bulk * G / (3. * bulk + G) * dS;
      }
    }
    clear_nb_data(d);
    main_neighbour_search(d);
    if (calc_pair_pos_list(d)) return -1;
    if (calculate_surface(d, 1)) return -1;
  }
  /* CalculateSurface resets ID to ID_orig, which SolveDiffUpdateLeapFrog
   * sets only with contact; with damage (and no contact) it is taken from
   * ID here as well */
  if (d->model_damage && !d->contact)
    for (int i = 0; i < d->n; i++) d->fi[WFO_IDORIG][i] = d->fi[WFO_ID][i];
  clear_nb_data(d);
  d->isfirst = 1; d->isyielding = 0; d->ts_i = 0;
  d->initialized = 1;
  return 0;
}

static double max_pl_strain(const wfo_domain *d) {
  double mx = 0;
  for (int i = 0; i < d->n; i++)
    if (d->f[WFO_PLSTRAIN][i] > mx) mx = d->f[WFO_PLSTRAIN][i];
  return mx;
}

/* neighbour schedule at the top of a step (SolverLeapfrog.cpp:160-201) */
static int schedule_search(wfo_domain *d, double mx) {
  /* with damage: no search on first yielding, and ts_i = 1 after the first
   * step, so the first step's search is the last (SolverLeapfrog.cpp:160-171) */
  if (d->model_damage && !d->isfirst) d->ts_i = 1;
  if (!d->model_damage && mx > MIN_PS_FOR_NBSEARCH && !d->isyielding) {
    clear_nb_data(d);
    main_neighbour_search(d);
    if (calc_pair_pos_list(d)) return -1;
    if (d->contact && contact_nb_update(d)) return -1;
    d->isyielding = 1;
  }
  if (mx > MIN_PS_FOR_NBSEARCH || d->isfirst) {
    if (d->ts_i == 0 && d->nb_cleared) {
      main_neighbour_search(d);
      if (calc_pair_pos_list(d)) return -1;
      if (d->grad_corr) wfo_calc_grad_corr(d);
      if (d->contact && contact_nb_update(d)) return -1;
    }
  }
  /* with damage the free surface is refreshed every step, after the
   * search (SolverLeapfrog.cpp:204-206) */
  if (d->model_damage && calculate_surface(d, 1)) return -1;
  /* SolverLeapfrog.cpp:207-212: the correction again on the first step */
  if (d->isfirst && d->grad_corr) wfo_calc_grad_corr(d);
  return 0;
}

/* neighbour schedule at the end of a step (SolverLeapfrog.cpp:347-356) */
static void schedule_end(wfo_domain *d, double mx) {
  if (mx > MIN_PS_FOR_NBSEARCH) {
    if (d->ts_i == (d->ts_nb_inc - 1)) clear_nb_data(d);
    d->ts_i++;
    if (d->ts_i > (d->ts_nb_inc - 1)) d->ts_i = 0;
  }
}

static void density_update(wfo_domain *d) {
  for (int i = 0; i < d->n; i++) d->f[WFO_RHO][i] += d->deltat * d->f[WFO_DRHO][i];
}

static void drift(wfo_domain *d) {
  double *x = d->f[WFO_X], *v = d->f[WFO_V], *xp = d->f[WFO_XPREV], *u = d->f[WFO_DISP];
  for (int i = 0; i < d->n; i++)
    for (int k = 0; k < 3; k++) {
      xp[3 * i + k] = x[3 * i + k];
      double du = (v[3 * i + k] + 0.0) * d->deltat;
      u[3 * i + k] += du;
      x[3 * i + k] += du;
    }
}

static void step_tail(wfo_domain *d, double mx) {
  plastic_work_heat(d, d->deltat);
  wfo_thermal_calcs(d, d->deltat);
  kin_energy(d);
  int_energy(d);
  d->steps++;
  d->Time += d->deltat;
  d->isfirst = 0;
  if (d->contact) mesh_update(d);
  schedule_end(d, mx);
}

/* One iteration of SolveDiffUpdateLeapFrog's loop
 * (Source/SolverLeapfrog.cpp:124-362), fixed time step, no contact. */
int wfo_step_leapfrog(wfo_domain *d) {
  if (!d->initialized) return fail("wfo_solve_init not called");
  wfo_start_acceleration(d);
  double mx = max_pl_strain(d);
  d->max_pl_strain = mx;
  if (schedule_search(d, mx)) return -1;
  wfo_calc_accel(d);
  d->react_pending = 1;
  wfo_general_after(d);
  if (d->contact) contact_law(d, 1);
  double *v = d->f[WFO_V], *a = d->f[WFO_A];
  if (d->isfirst)
    for (int i = 0; i < 3 * d->n; i++) v[i] -= a[i] * 0.5 * d->deltat;
  wfo_general_after(d);
  double prev = update_time_step(d);
  for (int i = 0; i < 3 * d->n; i++) v[i] += a[i] * 0.5 * (prev + d->deltat);
  wfo_general_after(d);
  wfo_calc_dens(d);
  density_update(d);
  drift(d);
  wfo_calc_rates(d);
  wfo_stress_all(d, d->deltat);
  /* SolverLeapfrog.cpp:315 */
  if (d->model_damage) calc_damage(d);
  step_tail(d, mx);
  return 0;
}

/* One iteration of SolveDiffUpdateKickDrift's loop
 * (Source/SolverKickDrift.cpp:117-330), fixed time step, no contact. */
int wfo_step_kickdrift(wfo_domain *d) {
  if (!d->initialized) return fail("wfo_solve_init not called");
  if (d->model_damage) return fail("damage is modelled by SolveDiffUpdateLeapFrog only");
  wfo_start_acceleration(d);
  double mx = max_pl_strain(d);
  d->max_pl_strain = mx;
  if (schedule_search(d, mx)) return -1;
  wfo_calc_accel(d);
  d->react_pending = 1;
  wfo_general_after(d);
  if (d->contact) contact_forces(d);
  double *v = d->f[WFO_V], *a = d->f[WFO_A];
  double prev = update_time_step(d);
  for (int i = 0; i < 3 * d->n; i++) v[i] += a[i] * 0.5 * prev * 1.;
  wfo_general_after(d);
  wfo_calc_dens(d);
  for (int i = 0; i < d->n; i++) d->f[WFO_RHO][i] += d->deltat * d->f[WFO_DRHO][i] * 1.;
  drift(d);
  for (int i = 0; i < 3 * d->n; i++) v[i] += a[i] * d->deltat / 2. * 1.;
  wfo_general_after(d);
  wfo_calc_rates(d);
  wfo_stress_all(d, d->deltat);
  step_tail(d, mx);
  return 0;
}
