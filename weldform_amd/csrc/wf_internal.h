// wf_internal.h -- device data layout and launcher declarations shared by
// the HIP kernels (wf_kernels.hip) and the C-ABI layer (wf_capi.hip).
//
// HBM layout (internal particle order = Morton order of the initial cell
// grid, fixed for the lifetime of a context):
//   xh[2]   double4  (x, y, z, h)          positions, double-buffered
//   vm      double4  (vx, vy, vz, mass)
//   acc     double4  (ax, ay, az, 0)
//   srec    StressRec (64 B) = (sigma [6], rho, 1/rho^2): the stress snapshot
//           the acceleration gather reads for every neighbour
//   rho[2], vol[2]   double (vol = mass/rho), double-buffered
//   per-particle own state as component-major arrays (coalesced, never
//   gathered): S[6], strain[6], epl[6], eplinc[6], sr[6], rr[3], scalars
//   nbr     int32 neighbour list in groups of four: entry k of particle i
//           sits at ((k/4) * npad + i) * 4 + k%4, so lane i of a wavefront
//           reads entries 4g..4g+3 as one int4 and the wave's load is one
//           coalesced 1 KiB run
//   stage   int32 [ntiles][kStageMax], nbrl uint16 [max_nb][npad]: the LDS
//           staging lists of the mixed-precision pair passes
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/weldform_hip.h"

namespace wf {

constexpr int kBlock = 128;
constexpr int kScanChunk = 1024;  // cells per block of the counting sort's cell-start scan
// LDS staging of the mixed-precision pair passes: a tile is one block of
// kBlock consecutive particles; its staged set (the tile's particles and all
// their neighbours, sorted by index) holds at most kStageMax particles.
// 896 keeps 99.7 % of the tiles of a cubic lattice with 2h = 2.4 dx staged
// (sets of 622-932, median 733) while three acceleration blocks, four rate
// blocks and five density blocks fit a CU's 160 KB of LDS.
constexpr int kStageMax = 896;
constexpr int kStagePow2 = 1024;  // the power of two >= kStageMax the set sort pads to

// neighbour rows are read in batches of kNbBatch entries, the next batch
// prefetched while the current one is processed
constexpr int kNbBatch = 8;
constexpr int kMaxLanes = 4;     // threads per particle in a pair pass
constexpr int kLanesMixed = 2;   // defaults (tuned on MI355X)
constexpr int kLanesFP64 = 2;
__host__ __device__ inline long long nb_slot(long long npad, int k, long long i) {
  return ((long long)(k >> 2) * npad + i) * 4 + (k & 3);
}
// spare entries per particle past a full row (max_nb entries) that the
// pair passes' row prefetch may touch: a batch of kNbBatch per lane, or
// three groups of four per lane
constexpr int kRowSpare = 16 * kMaxLanes;
static_assert(kRowSpare >= kNbBatch * kMaxLanes, "row spare covers the batch prefetch");
// entries allocated per particle
__host__ __device__ inline int nb_rows(int max_nb) {
  return (max_nb + kNbBatch - 1) / kNbBatch * kNbBatch + kRowSpare;
}
constexpr int kMaxBc = 8;        // velocity / temperature BC slots per context
constexpr int kMaxZones = 8;     // velocity-BC zones with a reaction history
constexpr int kMaxMesh = 4;      // rigid contact planes
constexpr int kMaxContact = 32;  // contact partners per particle
constexpr int kMaxMat = 256;
constexpr int kMaxNb = 128;

// symmetric 3x3 as (xx, yy, zz, xy, xz, yz)
enum { XX = 0, YY = 1, ZZ = 2, XY = 3, XZ = 4, YZ = 5 };

struct __align__(64) StressRec {
  double s[6];   // Cauchy stress Sigma (symmetric)
  double rho;    // density the record was formed with
  double irho2;  // 1 / (rho * rho)
};

// FP32 record gathered by the mixed-precision acceleration pass: the
// stress already divided by rho^2 (the factor the pair term applies), formed
// in FP64 and rounded once
struct __align__(32) StressRecF {
  float t[6];    // sigma / rho^2
  float rho;
  float pad;
};

struct DevMat {
  int model;
  int dam_model;               // damage: 0 none, 1 Johnson-Cook (dam: D1..D5)
  double E, G, Cs, ref_density, P0, alpha, beta, Ep;
  double hol_K, hol_m, hol_sy0, hol_eps0, hol_eps1;
  double jc_A, jc_B, jc_n, jc_C, jc_eps0, jc_m, jc_Tm, jc_Tt;
  double k_T, cp_T, th_exp;
  double ti, ti_n, ti_dist;    // tensile instability: TI, TIn, TIInitDist
  // Coulomb friction of the material's particles against the rigid
  // surfaces (wf_set_material_friction); negative: the context's
  double fr_sta, fr_dyn;
  double dam[5];
  // origin of the material's FP32 position shadows: the centre of its
  // particles' initial bounding box (pairs form within one material, so the
  // pair differences see one origin; a batch of separated specimens keeps
  // each specimen's shadows as fine as a lone specimen's)
  double xc[3];
};

// device-side solver state (one per context)
struct DevFlags {
  int do_search;     // neighbour search in this step
  int yielded;       // sticky: some particle has pl_strain > 1e-6
  int isyielding;    // reference's isyielding
  int ts_i;          // reference's ts_i
  int cleared;       // reference's m_isNbDataCleared
  int step_yield;    // yielded value sampled at the top of the step
  int overflow;      // a neighbour row was truncated
  int nan;           // non-finite acceleration seen
  int max_count;     // largest neighbour count of the last build
  int gc_due;        // the kernel-gradient correction is formed in this step
  int ts_nb_inc;     // reference's ts_nb_inc (search period after yield)
  long long nb_searches;
  long long directed_pairs;
  long long step;    // completed steps since wf_init (reaction history slot)
  double time;       // Domain::Time at the start of the next step
  double dt;         // Domain::deltat (changes only with the adaptive step)
  double dt_prev;    // prev_deltat of the step
  double dtint;      // Domain::deltatint (the initial dt)
  double min_force_ts;  // Domain::min_force_ts (the contact step criterion)
  unsigned hmax_bits;   // largest FP32 h of the search in progress (bits of a positive float)
  unsigned fin_ticket;  // blocks of the end-of-step reduction done (0 between steps)
  unsigned scan_ticket; // blocks of the cell-start scan done (0 between searches)
};

// partial-sum slots per block
enum {
  P_INT = 0,         // sum vol * S:SR       (internal energy rate)
  P_KIN,             // sum of kinetic energy pair terms
  P_PLW,             // sum vol * q_plheat   (plastic work rate)
  P_MAXT, P_MINT,    // max / min temperature
  P_REACT,           // + 3*zone + comp
  P_CFORCE = P_REACT + 3 * kMaxZones,   // + 3*mesh + comp (contact force per plane)
  P_CFSUM = P_CFORCE + 3 * kMaxMesh,    // sum |contforce|
  P_CREACT,                             // sum m a.n of contact particles
  P_EXTW,                               // sum contforce . v_rigid
  P_DTMIN,                              // min time-step criterion
  P_FTS,                                // min dt_fext of the contact pairs
  P_FRW,                                // sum q_fric_work m / rho
  P_CHC,                                // sum q_cont_conv m / rho
  P_COUNT
};

enum { S_KIN = 0, S_INT, S_PLW, S_MAXT, S_MINT, S_MAXPL, S_CFSUM, S_CREACT, S_EXTW, S_FRW, S_CHC, S_COUNT };

constexpr int kMaxAmp = 8;
constexpr int kMaxAmpPts = 16;

// amplitude table (Source/Domain.h:69-89)
struct DevAmp {
  int id, n;
  double t[kMaxAmpPts], v[kMaxAmpPts];
};

// boundary conditions, constant for a context; amplitude-driven velocities
// are evaluated on the device at the step's time, so a step has no
// host-varying argument and can be captured in a graph
struct BcTable {
  int n;
  int zone[kMaxBc];
  int type[kMaxBc];            // 0 velocity, 2 temperature
  int value_type[kMaxBc];      // 0 constant, 1 amplitude
  int amp[kMaxBc];             // index into the amplitude table, -1 if none
  double amp_factor[kMaxBc];
  double v[kMaxBc][3];         // constant value / amplitude direction
  double w[kMaxBc][3];         // angular velocity a constant BC gives its rigid mesh
  double T[kMaxBc];
  int nz;                      // reaction zones
  int rzone[kMaxZones];
};

struct Dev {
  int n;
  int dim;  // Domain::Dimension: 3, or 2 (plane strain, one z-plane)
  long long npad;
  int max_nb;
  int ncells;
  int cellno[3];
  double blpf[3];
  double cellsize[3];
  double cellfac;
  double pl_frac;
  int thermal;
  int pl_heating;
  int mixed;
  int nblocks;
  // state
  double4 *xh[2];
  double4 *vm;
  double4 *acc;
  double4 *disp;
  StressRec *srec;
  double *rho[2];
  double *vol[2];
  double *S, *strain, *epl, *eplinc, *sr, *rr;   // [6|3][npad]
  double *pl, *dpl, *sigy, *effsr, *pres, *drho;
  double *T, *dTdt, *qpl, *psen, *qconv, *hconv, *tinf, *qsrc;
  int *id, *mat, *group, *thbc, *perm;
  // mixed-precision shadows of the gathered inputs (written by the FP64
  // integrators); positions are stored relative to the material's origin
  // DevMat::xc to keep their rounding at the specimen's scale
  float4 *xhf[2];              // (x - xc, y - yc, z - zc, mass/rho)
  float4 *vmf;                 // (vx, vy, vz, mass)
  StressRecF *srecf;
  float *prf;                  // P / rho^2 (EOS of the current density)
  // neighbours.  cell: each particle's cell in the reference grid (blpf,
  // cellsize, cellno: CellInitiate's), which the rigid binning, the contact
  // partners and the thermal pair orientation use; scell: its cell in the
  // search grid (same box, scellsize >= Cellfac x the largest SPH h), which
  // the SPH neighbour build uses and the counting sort below is over
  int *cell, *cell_count, *cell_start, *sorted;
  // counting sort: each particle's rank within its cell (atomic order), the
  // cells' in-chunk offsets, the chunks' totals and offsets (chunks of cells
  // scanned by one block each) and the blocks' shares of the largest h
  int *srank, *cell_loc, *ctot, *coff;
  unsigned *hred;
  int *scell;
  int scellno[3];
  double scellsize[3];
  int sncells;
  // cell-ordered candidate records of the last counting sort (position p of
  // `sorted`): offsets of the particle from its cell's lower corner and h in
  // FP32, and (particle index, group)
  float4 *sxyz;
  int2 *sinfo;
  // per search cell: where its z-slices 1, 2, 3 start in the cell order (a
  // cell's records are ordered by the quarter of the cell height they lie
  // in, then by index), and the cell's end
  int4 *soff;
  unsigned char *sslice;  // each particle's z-slice in its search cell
  unsigned *hblk;         // per block of particles: the largest FP32 h (bits)
  // rigid (contact) particles binned in the same cells: counts, starts,
  // cursors [ncells + 1], cell of each rigid particle and the cell order
  int *rcell, *rcount, *rstart, *rsorted;
  int *rrank, *rloc, *rtot, *roff;  // the rigid grid's counting sort, as above
  int *nnb, *nbr;
  // LDS staging (mixed precision): per tile the sorted staged particle list
  // stage[t][0..stage_cnt[t]) (stage_cnt = -1: tile gathers from HBM), each
  // neighbour entry as a position in that list (nbrl, same layout as nbr)
  // and each particle's own position (selfl)
  int stage_max, ntiles;
  int lanes_mx, lanes_f64;     // threads per particle in the pair passes
  int auto_ts;                 // bit 0 velocity, bit 1 acceleration criterion
  int ti_on;                   // some material has TI > 0
  int beta_on;                 // some material has a quadratic viscosity term (beta != 0)
  int gc_on;                   // kernel-gradient correction (Domain::gradKernelCorr)
  int locking;                 // Locking sums: float m/rho weights in the rate gather
  double *gcm;                 // gradCorrM, symmetric [6][npad] (FP64)
  double *tir;                 // tensile-instability stress [6][npad] (FP64)
  float *tirf;                 // its FP32 copy (mixed precision)
  double cfl, sqrt_h_a, min_ts;
  int *stage, *stage_cnt;
  uint16_t *nbrl, *selfl;
  // rigid contact planes (all meshes concatenated): nodes, elements
  // (n0, n1, n2, nfar as global node indices), element normal + pplane,
  // one rigid particle per element (x, h) and its velocity; per plane the
  // node / element ranges and its velocity m_v
  int contact, nmesh, nrig, nnode;
  int mesh_zone[kMaxMesh], mesh_node0[kMaxMesh], mesh_nnode[kMaxMesh], mesh_elem0[kMaxMesh], mesh_nelem[kMaxMesh];
  double4 *node, *enorm, *rx, *rv, *mv;
  double4 *mw, *nodev;         // mesh angular velocity [kMaxMesh]; node velocities of the last update
  int4 *elem;
  int *rmesh;
  double totmass;              // mean particle mass over solid and rigid particles
  double fr_sta, fr_dyn, dfac;
  // the contact law (Domain::contact_alg, Source/Domain.h:64: 0 Fraser,
  // 1 Wang, 2 Seo, 4 LSDyna), PFAC, fric_type == Fr_Bound, friction
  // function Linear (fr = fr_m T + fr_b) and, for Fraser's and Seo's laws,
  // Particle::cont_stiff per particle (Fraser 3-153, formed at set-up)
  int contact_law, fric_bound, fric_linear;
  // Domain::model_damage and Particle::dam_D (Johnson-Cook damage)
  int damage;
  double *damD;
  // CalcForce2233's own outputs (allocated by the first wf_calc_force2233):
  // Particle::ZWab, VXSPH, SumDen
  double *zwab, *sumden;
  double4 *vxsph;
  double pfac, fr_m, fr_b;
  double *cstiff;
  // contact heat (Source/Contact_Wang.cpp:245-324): conduction with
  // conductance hc from the rigid particles' temperatures rT, friction work
  // as heat; per particle q_fric_work, friction_hfl, cshearabs and
  // q_cont_conv (allocated when either is on)
  int heat_cond, heat_fric;
  double hc;
  double *rT;
  double *qfric, *fhfl, *cshear, *qcont;
  int *id_orig, *ccnt, *cidx;  // contact partners [kMaxContact][npad]
  double *pmass;               // particle mass (vm.w), for the free-surface sums
  double4 *cforce;
  double *cf_hist;             // [hist_cap][kMaxMesh][3]
  // sums per particle group (material index, wf_set_group_sums): each
  // group's particles in internal order (goff, gidx), per step the reaction
  // of velocity-BC zone gzone (sum m a at the observer) and the contact force
  // of each rigid mesh on the group, kept for the last ghist_cap steps
  int ngroups, gzone;
  long long ghist_cap;
  int *goff, *gidx;
  double4 *gcf;                // [kMaxMesh][npad] a particle's contact force per mesh (where ccnt > 0)
  double *gr_hist;             // [ghist_cap][ngroups][3]
  double *gc_hist;             // [ghist_cap][ngroups][kMaxMesh][3]
  // materials, flags, reductions
  const DevMat *mats;
  const DevAmp *amps;
  DevFlags *flags;
  double *partials;            // [P_COUNT][nblocks]
  double *fin;                 // [P_COUNT] the reduced slots of the step
  double *scalars;             // [S_COUNT]
  double *react_hist;          // [hist_cap][kMaxZones][3]
  long long hist_cap;
};

// ---------------------------------------------------------------------------
// adjoint of the FP64 Leapfrog step (wf_adjoint.hip).  A segment of steps is
// replayed from a checkpoint with a tape: slot s holds the state at the
// start of step s of the segment and that step's intermediates; the list
// pointers name the neighbour / contact lists in effect for the step.
constexpr int kAdjParams = 7;  // A, B, n, C, m (Johnson-Cook), friction_sta, friction_dyn
struct AdjSlot {
  double4 *xh, *vm;            // (x, h), (v, m) at the step's start
  double *rho, *S, *sigy, *pl; // rho, S [6][npad], Sy, pl at the step's start
  double4 *apre;               // acceleration after the pair pass (before contact)
  double *drho, *sr, *rr;      // density rate, strain rate [6], rotation rate [3]
  int *nbr, *nnb;              // neighbour list in effect
  int *cidx, *ccnt, *id;       // contact partners and particle ids in effect
  double4 *node, *enorm, *rv;  // rigid planes at the step's contact evaluation
};
struct AdjBufs {
  double4 *xb, *vb;            // adjoint of the state x, v
  double4 *ab, *vpb;           // adjoint of the acceleration and of the kicked velocity
  double *rhob, *Sb, *syb, *plb;
  double4 *rsb;                // adjoints of the strain rate and rotation rate, 3 records per
                               // particle: (xx, yy, zz, xy), (xz, yz, w01, w02), (w12, 0, 0, 0)
  double *db;                  // adjoint of the density-rate sum
  double4 *tsig;               // the step's sigma / rho^2, 2 records per particle:
                               // (xx, yy, zz, xy), (xz, yz, rho, 1 / rho^2)
  double *thb;                 // per-particle parameter adjoints [kAdjParams][npad]
  double *part;                // [kAdjParams][nblocks] block partials
};
struct AdjSeeds {
  double cf[kMaxMesh][3];      // d loss / d plane contact force of the step
  double react[kMaxZones][3];  // d loss / d zone reaction of the step
};
void launch_adj_step(const Dev &d, hipStream_t st, const AdjSlot &s0, const AdjSlot &s1, const AdjBufs &b,
                     const BcTable &bc, const AdjSeeds &seeds, double dt, double kick, int adj_mat);
void launch_adj_reduce(const Dev &d, hipStream_t st, const AdjBufs &b, double *out);
void launch_tape_lists(const Dev &d, hipStream_t st, const AdjSlot &s, int force, int *ver, int slot);

// launchers (wf_kernels.hip); all enqueue on `st`
void launch_schedule(const Dev &d, hipStream_t st, int force);
void launch_plane_strain(const Dev &d, hipStream_t st);
void launch_neighbours(const Dev &d, hipStream_t st, int cur);
void launch_gradcorr(const Dev &d, hipStream_t st, int cur);
void launch_accel(const Dev &d, hipStream_t st, int cur);
void launch_tir(const Dev &d, hipStream_t st, int cur);
// dt < 0 in a launcher: the kernel reads the step from the device flags
// (adaptive time step).  phase 0: the whole kick; 1: BCs and the first-step
// half kick; 2: the kick with (dt_prev, dt) and the BCs.
// kick phases: K_WHOLE the whole kick; K_FIRST BCs and the first-step half
// kick; K_SECOND the kick with (dt_prev, dt) and the BCs; K_OBSERVE only the
// reaction observer.  K_WHOLE and K_FIRST observe the reaction too (before
// the BCs zero a); the _NOOBS forms do not (it was observed before a host
// hook ran).
enum { K_WHOLE = 0, K_FIRST, K_SECOND, K_OBSERVE, K_WHOLE_NOOBS, K_FIRST_NOOBS };
void launch_kick(const Dev &d, hipStream_t st, const BcTable &bc, int scheme, double dt_prev, double dt, int phase);
void launch_time_step(const Dev &d, hipStream_t st, int cur);
void launch_kick2(const Dev &d, hipStream_t st, double dt);
void launch_dens_drift(const Dev &d, hipStream_t st, int cur, double dt, int do_update);
void launch_rates(const Dev &d, hipStream_t st, int buf);
void launch_stress(const Dev &d, hipStream_t st, int buf, double dt);
void launch_thermal(const Dev &d, hipStream_t st, int buf, double dt);
void launch_finalize(const Dev &d, hipStream_t st, double dt, int nz);
void launch_max_pl(const Dev &d, hipStream_t st);
void launch_shadows(const Dev &d, hipStream_t st, int cur);
void launch_surface(const Dev &d, hipStream_t st, int cur, int setup);
// Domain::CalcForce2233 over every particle's row (FP64)
void launch_force2233(const Dev &d, hipStream_t st, int cur, double xsph, int shepard);
// law: Domain::contact_alg (0 Fraser, 1 Wang, 2 Seo, 4 LSDyna)
void launch_contact(const Dev &d, hipStream_t st, const BcTable &bc, int cur, double dt, int law);
// per-group sums of the step: what 0 the zone reaction (after the
// acceleration pass), 1 the contact forces per mesh (after contact)
void launch_group_sums(const Dev &d, hipStream_t st, int what);
void launch_rigid_drift(const Dev &d, hipStream_t st, double dt);
void launch_mesh_update(const Dev &d, hipStream_t st, double dt);
// output tables (wf_output.hip)
void launch_pack_xdmf(const Dev &d, hipStream_t st, int cur, float *f, int *iv);
void launch_pack_csv(const Dev &d, hipStream_t st, int cur, double *rows);
// the static device scratch of wf_pack_xdmf_noalloc: its device address on
// `device` (looked up on the first call per device)
constexpr size_t kOutScratchBytes = (size_t)64 << 20;
constexpr int kMaxDevices = 64;
hipError_t out_scratch(int device, void **ptr);

}  // namespace wf
