/* Runs the oracle's solver paths on small cases, for a build with the
 * address and undefined-behaviour sanitizers (tests/test_sanitizers.py):
 * Leapfrog with Johnson-Cook damage, the Locking sums, heat conduction and a
 * rigid plane under each contact law; KickDrift with the Nishimura sums and
 * the kernel-gradient correction.  Exit status 0 when every call succeeded. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "wf_oracle.h"

static int fails = 0;
#define CHECK(x)                                                      \
  do {                                                                \
    if ((x) != 0) {                                                   \
      fprintf(stderr, "%s:%d %s: %s\n", __FILE__, __LINE__, #x, wfo_last_error()); \
      fails++;                                                        \
    }                                                                 \
  } while (0)

/* a JC steel billet of R 3 mm, L 8 mm, spacing 1 mm, plus a rigid plane of
 * 2 x 6^2 triangles one h above it when `contact` */
static wfo_domain *billet(int contact, int law, int damage, int thermal, int locking, int grad_corr) {
  const double V[3] = {0, 0, 0}, r = 0.5e-3, h = 1.2 * 2 * r, rho = 7850.0, E = 200e9, nu = 0.29;
  const int64_t ns = wfo_gen_cylinder(V, 3e-3, 8e-3, r, NULL, 0);
  const int dens = 6, nr = contact ? 2 * dens * dens : 0;
  wfo_domain *d = wfo_create((int)ns + nr);
  double *x = wfo_f64(d, WFO_X);
  wfo_gen_cylinder(V, 3e-3, 8e-3, r, x, ns);
  const double K = E / (3 * (1 - 2 * nu)), G = E / (2 * (1 + nu)), Cs = sqrt(K / rho);
  double zmax = -1e9;
  for (int64_t i = 0; i < ns; i++) {
    wfo_f64(d, WFO_H)[i] = h;
    wfo_f64(d, WFO_MASS)[i] = M_PI * 3e-3 * 3e-3 * 8e-3 * rho / (double)ns;
    wfo_f64(d, WFO_RHO)[i] = wfo_f64(d, WFO_REFRHO)[i] = rho;
    wfo_f64(d, WFO_CS)[i] = Cs;
    wfo_f64(d, WFO_G)[i] = G;
    wfo_f64(d, WFO_ALPHA)[i] = 1.0;
    wfo_f64(d, WFO_SIGMAY)[i] = 792e6;
    wfo_f64(d, WFO_T)[i] = 300.0;
    wfo_f64(d, WFO_KT)[i] = 50.0;
    wfo_f64(d, WFO_CPT)[i] = 460.0;
    wfo_i32(d, WFO_MODEL)[i] = 2;
    if (x[3 * i + 2] > zmax) zmax = x[3 * i + 2];
  }
  for (int i = 0; i < (int)ns + nr; i++) wfo_i32(d, WFO_FREE)[i] = 1;
  wfo_material m = {0};
  m.model = 2;
  m.E = E, m.nu = nu, m.A = 792e6, m.B = 510e6, m.n = 0.26, m.C = 0.014, m.eps_0 = 1.0, m.jm = 1.03;
  m.Tm = 1793.0, m.Tt = 293.0;
  CHECK(wfo_set_material(d, 0, &m));
  const double lo[3] = {-1, -1, -1}, mid[3] = {1, 1, 2.6e-3}, hi0[3] = {-1, -1, 5.4e-3}, hi1[3] = {1, 1, 1};
  if (wfo_assign_zone(d, lo, mid, 2) < 0) fails++;
  if (wfo_assign_zone(d, hi0, hi1, 3) < 0) fails++;
  const double zero[3] = {0, 0, 0}, vdown[3] = {0, 0, -20.0};
  CHECK(wfo_add_velocity_bc(d, 2, 0, zero, 0, 1.0));
  CHECK(wfo_add_velocity_bc(d, 3, 0, vdown, 0, 1.0));
  CHECK(wfo_set_param(d, "solid_count", (double)ns));
  CHECK(wfo_set_param(d, "nonlock_sum", locking ? 0 : 1));
  CHECK(wfo_set_param(d, "grad_corr", grad_corr));
  if (thermal) {
    CHECK(wfo_set_param(d, "thermal_solver", 1));
    CHECK(wfo_set_param(d, "pl_heating", 1));
  }
  if (damage) {
    const double D[5] = {0.2, 0.3, -2.0, 0.01, 0.4};
    CHECK(wfo_set_material_damage(d, 0, 1, D));
    CHECK(wfo_set_param(d, "model_damage", 1));
  }
  if (contact) {
    const double p1[3] = {-3.75e-3, -3.75e-3, zmax + 1.02 * h}, p2[3] = {3.75e-3, 3.75e-3, zmax + 1.02 * h};
    if (wfo_add_plane_mesh(d, (int)ns, 10, 2, 0, p1, p2, dens, 1.1) < 0) fails++;
    CHECK(wfo_add_velocity_bc(d, 10, 0, vdown, 0, 1.0));
    CHECK(wfo_set_param(d, "contact", 1));
    CHECK(wfo_set_param(d, "contact_alg", law));
    CHECK(wfo_set_param(d, "friction_sta", 0.2));
    CHECK(wfo_set_param(d, "friction_dyn", 0.2));
  }
  return d;
}

int main(void) {
  wfo_set_threads(2);
  const double dt = 0.4 * 1.2e-3 / 5000.0;
  const int laws[4] = {1, 0, 2, 4};  /* the fifth run: Wang without damage */
  for (int k = 0; k < 5; k++) {
    wfo_domain *d = billet(1, laws[k % 4], k == 0, k == 0, 1, 0);
    CHECK(wfo_solve_init(d, dt));
    for (int s = 0; s < 40; s++) CHECK(wfo_step_leapfrog(d));
    double f[3];
    CHECK(wfo_contact_force(d, 0, f));
    double dmax = 0;
    for (int i = 0; i < wfo_n(d); i++) dmax = fmax(dmax, wfo_f64(d, WFO_DAMD)[i]);
    printf("leapfrog law %d: contact pairs %lld, contact_force_sum %g, plane force %g, max dam_D %g\n", laws[k % 4],
           (long long)wfo_ncontact_pairs(d), wfo_get_param(d, "contact_force_sum"), f[2], dmax);
    wfo_destroy(d);
  }
  wfo_domain *d = billet(0, 1, 0, 0, 1, 1);
  CHECK(wfo_solve_init(d, dt));
  for (int s = 0; s < 20; s++) CHECK(wfo_step_kickdrift(d));
  wfo_destroy(d);
  d = billet(0, 1, 0, 0, 0, 0);
  CHECK(wfo_solve_init(d, dt));
  for (int s = 0; s < 20; s++) CHECK(wfo_step_leapfrog(d));
  wfo_destroy(d);
  printf("done, %d failures\n", fails);
  return fails != 0;
}
