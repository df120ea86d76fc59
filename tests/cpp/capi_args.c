/* The C-ABI's host-side argument checking and generators, for a build with
 * the address and undefined-behaviour sanitizers (tests/test_sanitizers.py):
 * every call with a null or invalid argument must return an error code
 * (never crash), the error text must be readable, and the cylinder
 * generator must fill exactly the count it reports.  Needs no device: with
 * none present wf_create returns WF_EHIP after validating its arguments. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "weldform_hip.h"

static int fails = 0;
#define TRUE(cond)                                                          \
  do {                                                                      \
    if (!(cond)) {                                                          \
      fprintf(stderr, "%s:%d %s does not hold\n", __FILE__, __LINE__, #cond); \
      fails++;                                                              \
    }                                                                       \
  } while (0)
#define EXPECT(call, code)                                                               \
  do {                                                                                   \
    int rc_ = (call);                                                                    \
    if (rc_ != (code)) {                                                                 \
      fprintf(stderr, "%s:%d %s returned %d, expected %d\n", __FILE__, __LINE__, #call, rc_, code); \
      fails++;                                                                           \
    }                                                                                    \
  } while (0)

int main(void) {
  TRUE(wf_abi_version() > 0);
  int32_t ndev = -1;
  if (wf_device_count(&ndev) != WF_OK) ndev = 0;  /* no HIP runtime device: an error code */
  EXPECT(wf_device_count(NULL), WF_EINVAL);
  wf_config cfg;
  memset(&cfg, 0, sizeof cfg);
  cfg.precision = 7;
  cfg.max_nb = 64;
  cfg.cellfac = 2.0;
  cfg.ts_nb_inc = 5;
  void *ctx = NULL;
  EXPECT(wf_create(&cfg, 10, &ctx), WF_EINVAL);   /* unknown precision */
  cfg.precision = WF_FP64;
  cfg.max_nb = 500;
  EXPECT(wf_create(&cfg, 10, &ctx), WF_EINVAL);   /* neighbour capacity */
  cfg.max_nb = 64;
  EXPECT(wf_create(NULL, 10, &ctx), WF_EINVAL);
  EXPECT(wf_create(&cfg, 10, NULL), WF_EINVAL);
  EXPECT(wf_create(&cfg, 0, &ctx), WF_EINVAL);
  cfg.ts_nb_inc = 0;
  EXPECT(wf_create(&cfg, 10, &ctx), WF_EINVAL);
  cfg.ts_nb_inc = 5;
  if (ndev == 0) EXPECT(wf_create(&cfg, 10, &ctx), WF_EHIP);
  const char *e = wf_last_error(NULL);
  TRUE(e && strlen(e) > 0);
  /* every entry point refuses a null context */
  double buf[64] = {0}, v3[3] = {0, 0, 0};
  int32_t ib[64] = {0}, g3[3];
  int64_t cnt = 0;
  float fb[64];
  EXPECT(wf_destroy(NULL), WF_OK);  /* like free(NULL) */
  EXPECT(wf_set_field(NULL, WF_F_X, buf), WF_EINVAL);
  EXPECT(wf_get_field(NULL, WF_F_X, buf), WF_EINVAL);
  EXPECT(wf_set_int_field(NULL, WF_I_ID, ib), WF_EINVAL);
  EXPECT(wf_get_int_field(NULL, WF_I_ID, ib), WF_EINVAL);
  EXPECT(wf_add_amplitude(NULL, 1, 2, buf, buf), WF_EINVAL);
  EXPECT(wf_add_velocity_bc(NULL, 3, 0, v3, 0, 1.0), WF_EINVAL);
  EXPECT(wf_add_temperature_bc(NULL, 3, 300.0), WF_EINVAL);
  EXPECT(wf_set_bc_angular_velocity(NULL, 3, v3), WF_EINVAL);
  EXPECT(wf_add_rigid_plane(NULL, 10, 2, 0, v3, v3, 4, 1.1), WF_EINVAL);
  EXPECT(wf_set_contact(NULL, 0.2, 0.2, 0.0), WF_EINVAL);
  EXPECT(wf_set_material_friction(NULL, 0, 0.2, 0.2), WF_EINVAL);
  EXPECT(wf_set_contact_law(NULL, 2, 0.5, 0), WF_EINVAL);
  EXPECT(wf_set_friction_function(NULL, 1, 0.0, 0.1), WF_EINVAL);
  EXPECT(wf_set_material_damage(NULL, 0, 1, buf), WF_EINVAL);
  EXPECT(wf_set_contact_heat(NULL, 1, 1, 1e4), WF_EINVAL);
  EXPECT(wf_set_rigid_temperature(NULL, 1, buf), WF_EINVAL);
  EXPECT(wf_init(NULL, 1e-8), WF_EINVAL);
  EXPECT(wf_step(NULL, WF_LEAPFROG, 1), WF_EINVAL);
  EXPECT(wf_step_part(NULL, WF_LEAPFROG, 0), WF_EINVAL);
  EXPECT(wf_prepare(NULL, WF_LEAPFROG), WF_EINVAL);
  EXPECT(wf_sync(NULL), WF_EINVAL);
  EXPECT(wf_get_reaction_history(NULL, 3, buf, 1, &cnt), WF_EINVAL);
  EXPECT(wf_get_contact_force_history(NULL, 0, buf, 1, &cnt), WF_EINVAL);
  EXPECT(wf_set_group_sums(NULL, -1, 16), WF_EINVAL);
  EXPECT(wf_get_group_history(NULL, 1, 0, 0, buf, 1, &cnt), WF_EINVAL);
  EXPECT(wf_rebuild_neighbours(NULL), WF_EINVAL);
  EXPECT(wf_get_neighbour_pairs(NULL, ib, 1, &cnt), WF_EINVAL);
  EXPECT(wf_eval(NULL, 0, 1e-8), WF_EINVAL);
  EXPECT(wf_calc_force2233(NULL, 0.5, 1), WF_EINVAL);
  EXPECT(wf_set_timing(NULL, 1), WF_EINVAL);
  EXPECT(wf_get_cell_grids(NULL, g3, g3), WF_EINVAL);
  EXPECT(wf_pack_xdmf(NULL, fb, ib), WF_EINVAL);
  EXPECT(wf_pack_xdmf_noalloc(NULL, fb, ib), WF_EINVAL);
  EXPECT(wf_pack_csv(NULL, buf), WF_EINVAL);
  EXPECT(wf_adj_backward(NULL, buf), WF_EINVAL);
  EXPECT(wf_adj_restore(NULL), WF_EINVAL);
  TRUE(wf_pack_xdmf_noalloc_max() > 0);
  /* the generator: count, then exactly that many positions */
  const double V[3] = {0, 0, 0};
  int64_t n = wf_gen_cylinder(V, 7.5e-3, 30e-3, 0.5e-3, NULL, 0);
  TRUE(n == 4524);
  double *x = (double *)malloc((size_t)n * 3 * sizeof(double));
  TRUE(wf_gen_cylinder(V, 7.5e-3, 30e-3, 0.5e-3, x, n) == n);
  free(x);
  /* a smaller buffer: the first cap positions, the full count returned */
  x = (double *)malloc((size_t)(n - 100) * 3 * sizeof(double));
  TRUE(wf_gen_cylinder(V, 7.5e-3, 30e-3, 0.5e-3, x, n - 100) == n);
  TRUE(wf_gen_cylinder(V, -1.0, 30e-3, 0.5e-3, NULL, 0) < 0);
  free(x);
  printf("capi argument checks: %d failures\n", fails);
  return fails != 0;
}
