// wf_capi.hip -- C-ABI of libweldform_hip.so (declared in include/weldform_hip.h).
//
// Host side of the device path: staging of particle fields in the caller's
// order, the reference's cell grid (CellInitiate), the Morton permutation,
// device allocation, and the per-step launch sequence of the Leapfrog and
// KickDrift loops.  No host<->device synchronisation happens inside
// wf_step: the neighbour-search schedule lives on the device.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "wf_internal.h"

namespace {

using namespace wf;

thread_local std::string g_err;

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(call)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      throw Error(WF_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));               \
  } while (0)

const int kFieldDim[WF_NFIELDS] = {3, 3, 3, 3, 1, 1, 1, 9, 9, 9, 9, 9, 9, 1, 1,
                                   1, 1, 1, 1, 1, 1, 1, 9, 1, 1, 1, 1, 1, 3, 9, 1,
                                   1, 3, 1};

struct BcHost {
  int zone, type, value_type, amp_id;
  double value[3], amp_factor, T, w[3];
};
struct AmpHost {
  int id;
  std::vector<double> t, v;
  // amplitude::getValAtTime (Source/Domain.h:75-87)
  double at(double time) const {
    int i = (int)t.size() - 1;
    bool end = false;
    while (!end) {
      i--;
      if (i == 0 || time > t[i]) end = true;
    }
    return v[i] + (v[i + 1] - v[i]) / (t[i + 1] - t[i]) * (time - t[i]);
  }
};

enum { T_ACCEL = 0, T_DENS, T_RATES, T_NBR, T_STEP, T_SURF, T_COUNT };

// rigid plane mesh (TriMesh::AxisPlaneMesh + CalcSpheres, Source/Mesh.cpp:82-220)
struct MeshHost {
  int zone = 0;
  std::vector<double> node;        // 3 per node
  std::vector<int> el;             // 3 per element
  std::vector<int> nfar;
  std::vector<double> centroid, normal, pplane, radius, h;
};

struct Ctx {
  wf_config cfg{};
  std::vector<MeshHost> meshes;
  double fr_sta = 0.0, fr_dyn = 0.0, dfac = 0.0;
  // wf_set_contact_law / wf_set_friction_function (PFAC defaults to the JSON
  // driver's penaltyFactor default, Source/WeldForm.cpp:685)
  int contact_law = 1, fric_bound = 0, fric_linear = 0;
  int dim = 3;  // Domain::Dimension (wf_set_dimension)
  double pfac = 0.5, fr_m = 0.0, fr_b = 0.0;
  int heat_cond = 0, heat_fric = 0;
  double heat_hc = 0.0;
  std::vector<double> rigid_T;     // one per rigid particle (mesh order), 0 unless set
  int64_t n = 0;
  long long npad = 0;
  hipStream_t st = nullptr;
  std::vector<double> hf[WF_NFIELDS];
  std::vector<int32_t> hi[WF_NIFIELDS];
  std::vector<wf_material> mats;
  std::vector<char> mat_set;
  std::vector<BcHost> bcs;
  std::vector<AmpHost> amps;
  bool inited = false;
  bool built = false;
  std::vector<int32_t> perm, inv;  // perm[internal] = original
  std::vector<double> mat_xc;      // per-material FP32 shadow origin [kMaxMat][3]
  std::vector<double> mat_fr;      // per-material friction (static, dynamic) [kMaxMat][2]; < 0: unset
  std::vector<int> mat_dam;        // per-material damage model (0 none, 1 Johnson-Cook) [kMaxMat]
  std::vector<double> mat_dpar;    // its D1..D5 [kMaxMat][5]
  int gsum_zone = -2;              // wf_set_group_sums: -2 off, -1 contact forces only, else the zone
  long long gsum_cap = 0;
  Dev d{};
  std::vector<std::pair<void *, size_t>> allocs;  // (pointer, bytes) of the solver state
  DevMat *dmats = nullptr;
  int cur = 0;
  double dt = 0.0;  // the initial step (Domain::deltatint); the device flags hold the current one
  long long steps = 0;
  int next_part = 0;  // the part wf_step_part expects next (0: a step boundary)
  std::vector<int> rzones;
  BcTable bc{};
  bool bc_dirty = true;
  DevAmp *damps = nullptr;
  hipGraphExec_t gexec[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // [scheme][cur]
  std::string err;
  // timing
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;
  struct Rec {
    int k;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;
  double tsum[T_COUNT] = {0};
  long long tcnt[T_COUNT] = {0};
  size_t ev_next = 0;

  template <class T>
  T *alloc(size_t count) {
    void *p = nullptr;
    if (count == 0) count = 1;
    HIPCHK(hipMalloc(&p, count * sizeof(T)));
    HIPCHK(hipMemsetAsync(p, 0, count * sizeof(T), st));
    allocs.push_back({p, count * sizeof(T)});
    return (T *)p;
  }
  // adjoint storage: checkpoints, tape, adjoint buffers (not part of the
  // snapshots)
  struct Adj {
    wf_adj_config cfg{};
    bool ready = false;       // a window's forward has run
    bool window = false;      // snapshot 0 holds a window's start
    long long step0 = 0;      // steps completed before the window
    size_t snap_bytes = 0;
    std::vector<char *> ckpt; // device snapshots of every state allocation
    std::vector<std::pair<int, long long>> ckh;  // (cur, steps) at each snapshot
    std::vector<double> hist, target;
    double loss = 0.0;
    int seg = 0;              // tape slots (checkpoint interval)
    std::vector<AdjSlot> slot;  // seg + 1 state slots; intermediates and lists in the first seg
    int *ver = nullptr;
    AdjBufs buf{};
    double *grad = nullptr;
    bool have_grad = false;
    std::vector<void *> mem;
  } adj;
  // output staging (wf_pack_*): allocated on first use, not snapshotted
  void *outbuf = nullptr;
  size_t out_cap = 0;
  template <class T>
  T *adj_alloc(size_t count) {
    void *p = nullptr;
    if (count == 0) count = 1;
    HIPCHK(hipMalloc(&p, count * sizeof(T)));
    HIPCHK(hipMemsetAsync(p, 0, count * sizeof(T), st));
    adj.mem.push_back(p);
    return (T *)p;
  }
  void drop_graphs() {
    for (auto &g : gexec)
      for (auto &e : g)
        if (e) {
          (void)hipGraphExecDestroy(e);
          e = nullptr;
        }
  }
  ~Ctx() {
    drop_graphs();
    for (auto &e : ev_pool) (void)hipEventDestroy(e);
    for (auto &p : allocs) (void)hipFree(p.first);
    for (void *p : adj.mem) (void)hipFree(p);
    if (outbuf) (void)hipFree(outbuf);
    if (st) (void)hipStreamDestroy(st);
  }
};

int fail(Ctx *c, int code, const std::string &m) {
  g_err = m;
  if (c) c->err = m;
  return code;
}

template <class F>
int guard(Ctx *c, F &&f) {
  try {
    return f();
  } catch (const Error &e) {
    return fail(c, e.code, e.what());
  } catch (const std::bad_alloc &) {
    return fail(c, WF_ENOMEM, "host out of memory");
  } catch (const std::exception &e) {
    return fail(c, WF_EINVAL, e.what());
  }
}

Ctx *as_ctx(void *p) {
  if (!p) throw Error(WF_EINVAL, "null context");
  return (Ctx *)p;
}

// ---------------------------------------------------------------------------
// device <-> host helpers (internal order <-> caller's order)

template <class T>
std::vector<T> dget(Ctx *c, const T *p, size_t count) {
  std::vector<T> h(count);
  HIPCHK(hipMemcpyAsync(h.data(), p, count * sizeof(T), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return h;
}
template <class T>
void dput(Ctx *c, T *p, const std::vector<T> &h) {
  HIPCHK(hipMemcpyAsync(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
}

// symmetric storage index of (a,b)
inline int sym_index(int a, int b) {
  if (a == b) return a;
  if (a > b) std::swap(a, b);
  return a == 0 ? (b == 1 ? XY : XZ) : YZ;
}

// component-major [6][npad] from 9-per-particle caller data
void put_sym(Ctx *c, double *dev, const double *h9) {
  std::vector<double> t((size_t)6 * c->npad, 0.0);
  for (int64_t k = 0; k < c->n; ++k) {
    const double *m = h9 + 9 * (size_t)c->perm[k];
    const int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
    for (int q = 0; q < 6; ++q) t[(size_t)q * c->npad + k] = m[3 * ia[q] + ib[q]];
  }
  dput(c, dev, t);
}
void get_sym(Ctx *c, const double *dev, double *h9) {
  auto t = dget(c, dev, (size_t)6 * c->npad);
  for (int64_t k = 0; k < c->n; ++k) {
    double *m = h9 + 9 * (size_t)c->perm[k];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) m[3 * a + b] = t[(size_t)sym_index(a, b) * c->npad + k];
  }
}
void put_scalar(Ctx *c, double *dev, const double *h) {
  std::vector<double> t((size_t)c->npad, 0.0);
  for (int64_t k = 0; k < c->n; ++k) t[k] = h[c->perm[k]];
  dput(c, dev, t);
}
void get_scalar(Ctx *c, const double *dev, double *h) {
  auto t = dget(c, dev, (size_t)c->npad);
  for (int64_t k = 0; k < c->n; ++k) h[c->perm[k]] = t[k];
}
// double4 arrays: replace components [c0, c0+nc) from caller data of width w
void put_d4(Ctx *c, double4 *dev, const double *h, int c0, int nc, int w) {
  auto t = dget(c, dev, (size_t)c->npad);
  for (int64_t k = 0; k < c->n; ++k) {
    double *e = &t[k].x;
    for (int q = 0; q < nc; ++q) e[c0 + q] = h[(size_t)w * c->perm[k] + q];
  }
  dput(c, dev, t);
}
void get_d4(Ctx *c, const double4 *dev, double *h, int c0, int nc, int w) {
  auto t = dget(c, dev, (size_t)c->npad);
  for (int64_t k = 0; k < c->n; ++k) {
    const double *e = &t[k].x;
    for (int q = 0; q < nc; ++q) h[(size_t)w * c->perm[k] + q] = e[c0 + q];
  }
}

// ---------------------------------------------------------------------------
// set-up: CellInitiate (Source/Neighbour.cpp:10-139), 3D, non-periodic

void cell_initiate(Ctx *c) {
  const std::vector<double> &x = c->hf[WF_F_X], &h = c->hf[WF_F_H];
  double TRPR[3], BLPF[3], hmax = h[0];
  for (int k = 0; k < 3; ++k) TRPR[k] = BLPF[k] = x[k];
  for (int64_t i = 0; i < c->n; ++i) {
    for (int k = 0; k < 3; ++k) {
      if (x[3 * i + k] > TRPR[k]) TRPR[k] = x[3 * i + k];
      if (x[3 * i + k] < BLPF[k]) BLPF[k] = x[3 * i + k];
    }
    if (h[i] > hmax) hmax = h[i];
  }
  // mixed-precision shadows store positions relative to the body's centre
  // rigid-plane particles follow the solid ones in the reference's particle
  // list and count for the bounding box and h_max
  for (const MeshHost &M : c->meshes)
    for (size_t e = 0; e < M.h.size(); ++e) {
      for (int k = 0; k < 3; ++k) {
        if (M.centroid[3 * e + k] > TRPR[k]) TRPR[k] = M.centroid[3 * e + k];
        if (M.centroid[3 * e + k] < BLPF[k]) BLPF[k] = M.centroid[3 * e + k];
      }
      if (M.h[e] > hmax) hmax = M.h[e];
    }
  for (int k = 0; k < 3; ++k) {
    if (c->cfg.dom_max[k] > TRPR[k]) TRPR[k] = c->cfg.dom_max[k];
    if (c->cfg.dom_min[k] < BLPF[k]) BLPF[k] = c->cfg.dom_min[k];
  }
  for (int k = 0; k < 3; ++k) {
    TRPR[k] += hmax / 2;
    BLPF[k] -= hmax / 2;
  }
  long long ncells = 1;
  for (int k = 0; k < 3; ++k) {
    const double e = (TRPR[k] - BLPF[k]) / (c->cfg.cellfac * hmax);
    int cn = ((double)(std::ceil(e) - e) < (hmax / 10.0)) ? (int)std::ceil(e) : (int)std::floor(e);
    // 2D: one layer of cells along z (Source/Neighbour.cpp:56-75); its
    // height here is the box's, so every particle falls in it
    if (k == 2 && c->dim == 2) cn = 1;
    if (cn <= 0) throw Error(WF_EINVAL, "CellInitiate: zero cells along an axis");
    c->d.cellno[k] = cn;
    c->d.blpf[k] = BLPF[k];
    ncells *= cn;
  }
  if (ncells > (1LL << 30)) throw Error(WF_EINVAL, "CellInitiate: too many cells");
  for (int k = 0; k < 3; ++k) c->d.cellsize[k] = (TRPR[k] - BLPF[k]) / c->d.cellno[k];
  c->d.ncells = (int)ncells;
  // The SPH neighbour search's own grid.  The reference's cells are sized by
  // h_max over all particles, rigid-surface ones included; with large rigid
  // particles a cell then holds hundreds of SPH particles.  SPH pairs only
  // form within R = Cellfac x h_max over the SPH particles.  When every
  // reference cell is at least R wide, the reference finds every SPH pair
  // within R (clamped particles included: a pair within R lies in adjacent
  // clamped cells of any grid of cells at least R wide), so a finer grid of
  // cells at least R wide over the same box finds the same pairs.  It is
  // used when its cells are at most 2/3 of the reference's along some axis.
  double hs = 0.0;
  for (int64_t i = 0; i < c->n; ++i) hs = std::max(hs, h[i]);
  const double R = c->cfg.cellfac * hs;
  bool finer = R > 0.0;
  long long sn = 1;
  int scn[3];
  for (int k = 0; k < 3; ++k) {
    if (c->d.cellsize[k] < R) finer = false;
    scn[k] = std::max(1, (int)std::floor((TRPR[k] - BLPF[k]) / R));
    sn *= scn[k];
  }
  bool gain = false;
  for (int k = 0; k < 3; ++k) gain = gain || (TRPR[k] - BLPF[k]) / scn[k] <= c->d.cellsize[k] * (2.0 / 3.0);
  if (!(finer && gain && sn <= (1LL << 28))) {
    for (int k = 0; k < 3; ++k) scn[k] = c->d.cellno[k];
    sn = ncells;
  }
  for (int k = 0; k < 3; ++k) {
    c->d.scellno[k] = scn[k];
    c->d.scellsize[k] = (TRPR[k] - BLPF[k]) / scn[k];
  }
  c->d.sncells = (int)sn;
}

uint64_t spread3(uint64_t v) {
  v &= 0x1fffff;
  v = (v | v << 32) & 0x1f00000000ffffULL;
  v = (v | v << 16) & 0x1f0000ff0000ffULL;
  v = (v | v << 8) & 0x100f00f00f00f00fULL;
  v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
  v = (v | v << 2) & 0x1249249249249249ULL;
  return v;
}

int host_cell_coord(double x, double blpf, double cs, int cn) {
  double f = std::floor((x - blpf) / cs);
  if (!(f >= 0.0)) return 0;
  if (f >= (double)cn) return cn - 1;
  return (int)f;
}

// internal order: Morton order of the particles' initial cells (finer grid
// of half a cell so neighbouring lanes share cache lines), ties by index
void make_permutation(Ctx *c) {
  const std::vector<double> &x = c->hf[WF_F_X];
  std::vector<std::pair<uint64_t, int32_t>> key(c->n);
  for (int64_t i = 0; i < c->n; ++i) {
    uint64_t m = 0;
    for (int k = 0; k < 3; ++k) {
      const int cc = host_cell_coord(x[3 * i + k], c->d.blpf[k], c->d.cellsize[k] * 0.5, 2 * c->d.cellno[k]);
      m |= spread3((uint64_t)cc) << k;
    }
    key[i] = {m, (int32_t)i};
  }
  std::sort(key.begin(), key.end());
  c->perm.resize(c->n);
  c->inv.resize(c->n);
  for (int64_t k = 0; k < c->n; ++k) {
    c->perm[k] = key[k].second;
    c->inv[key[k].second] = (int32_t)k;
  }
}

DevMat to_dev(const wf_material &m) {
  DevMat d{};
  d.model = m.model;
  d.E = m.E;
  d.G = m.G;
  d.Cs = m.Cs;
  d.ref_density = m.ref_density;
  d.P0 = m.P0;
  d.alpha = m.alpha;
  d.beta = m.beta;
  // CalcStressStrain replaces a negative Ep by E on the plastic branch
  // (Source/Particle.cpp:1020), the only place Ep is read
  d.Ep = m.Ep < 0 ? 1. * m.E : m.Ep;
  d.hol_K = m.hol_K;
  d.hol_m = m.hol_m;
  d.hol_sy0 = m.hol_sy0;
  if (m.model == WF_HOLLOMON) {
    d.hol_eps0 = m.hol_sy0 / m.E;
    d.hol_eps1 = std::pow(m.hol_sy0 / m.hol_K, 1. / m.hol_m);
  }
  d.jc_A = m.jc_A;
  d.jc_B = m.jc_B;
  d.jc_n = m.jc_n;
  d.jc_C = m.jc_C;
  d.jc_eps0 = m.jc_eps0;
  d.jc_m = m.jc_m;
  d.jc_Tm = m.jc_Tm;
  d.jc_Tt = m.jc_Tt;
  d.k_T = m.k_T;
  d.cp_T = m.cp_T;
  d.th_exp = m.th_exp;
  d.ti = m.ti;
  d.ti_n = m.ti_n;
  d.ti_dist = m.ti_dist;
  return d;
}

void allocate(Ctx *c) {
  Dev &d = c->d;
  const size_t np = (size_t)c->npad;
  d.n = (int)c->n;
  d.dim = c->dim;
  d.npad = c->npad;
  d.max_nb = c->cfg.max_nb;
  d.cellfac = c->cfg.cellfac;
  d.pl_frac = c->cfg.pl_work_heat_frac;
  d.thermal = c->cfg.thermal_solver;
  d.pl_heating = c->cfg.pl_heating;
  d.mixed = c->cfg.precision == WF_MIXED;
  d.nblocks = (int)((c->n + kBlock - 1) / kBlock);
  for (int b = 0; b < 2; ++b) {
    d.xh[b] = c->alloc<double4>(np);
    d.rho[b] = c->alloc<double>(np);
    d.vol[b] = c->alloc<double>(np);
  }
  d.vm = c->alloc<double4>(np);
  d.pmass = c->alloc<double>(np);
  d.acc = c->alloc<double4>(np);
  d.disp = c->alloc<double4>(np);
  d.srec = c->alloc<StressRec>(np);
  d.S = c->alloc<double>(6 * np);
  d.strain = c->alloc<double>(6 * np);
  d.epl = c->alloc<double>(6 * np);
  d.eplinc = c->alloc<double>(6 * np);
  d.sr = c->alloc<double>(6 * np);
  d.rr = c->alloc<double>(3 * np);
  double **scal[] = {&d.pl, &d.dpl, &d.sigy, &d.effsr, &d.pres, &d.drho, &d.T, &d.dTdt,
                     &d.qpl, &d.psen, &d.qconv, &d.hconv, &d.tinf, &d.qsrc};
  for (double **p : scal) *p = c->alloc<double>(np);
  int **ints[] = {&d.id, &d.mat, &d.group, &d.thbc, &d.perm, &d.cell, &d.sorted, &d.nnb};
  for (int **p : ints) *p = c->alloc<int>(np);
  d.nbr = c->alloc<int>((size_t)nb_rows(c->cfg.max_nb) * np);
  d.cell_count = c->alloc<int>((size_t)d.sncells + 1);
  d.cell_start = c->alloc<int>((size_t)d.sncells + 1);
  {
    // chunks of cells per block of the cell-start scan
    const size_t nsch = ((size_t)d.sncells + kScanChunk - 1) / kScanChunk;
    const size_t nrch = ((size_t)d.ncells + kScanChunk - 1) / kScanChunk;
    d.srank = c->alloc<int>(np);
    d.cell_loc = c->alloc<int>((size_t)d.sncells + 1);
    d.ctot = c->alloc<int>(nsch);
    d.coff = c->alloc<int>(nsch);
    d.hred = c->alloc<unsigned>(nsch + nrch);
  }
  d.scell = c->alloc<int>(np);
  d.sxyz = c->alloc<float4>(np + 4);
  d.sinfo = c->alloc<int2>(np + 4);
  d.soff = c->alloc<int4>((size_t)d.sncells);
  d.sslice = c->alloc<unsigned char>(np);
  d.hblk = c->alloc<unsigned>((size_t)d.nblocks);
  c->dmats = c->alloc<DevMat>(kMaxMat);
  d.mats = c->dmats;
  c->damps = c->alloc<DevAmp>(kMaxAmp);
  d.amps = c->damps;
  d.flags = c->alloc<DevFlags>(1);
  d.partials = c->alloc<double>((size_t)d.nblocks * P_COUNT);
  d.fin = c->alloc<double>(P_COUNT);
  d.scalars = c->alloc<double>(S_COUNT);
  d.hist_cap = 1 << 16;
  d.react_hist = c->alloc<double>((size_t)d.hist_cap * kMaxZones * 3);
  if (d.mixed) {
    for (int b = 0; b < 2; ++b) {
      d.xhf[b] = c->alloc<float4>(np);
    }
    d.vmf = c->alloc<float4>(np);
    d.srecf = c->alloc<StressRecF>(np);
    d.prf = c->alloc<float>(np);
  }
  if (!c->meshes.empty()) {
    d.contact = 1;
    d.nmesh = (int)c->meshes.size();
    int nn = 0, ne = 0;
    for (int m = 0; m < d.nmesh; ++m) {
      const MeshHost &M = c->meshes[m];
      d.mesh_zone[m] = M.zone;
      d.mesh_node0[m] = nn;
      d.mesh_nnode[m] = (int)M.node.size() / 3;
      d.mesh_elem0[m] = ne;
      d.mesh_nelem[m] = (int)M.h.size();
      nn += d.mesh_nnode[m];
      ne += d.mesh_nelem[m];
    }
    d.nnode = nn;
    d.nrig = ne;
    d.node = c->alloc<double4>(nn);
    d.elem = c->alloc<int4>(ne);
    d.enorm = c->alloc<double4>(ne);
    d.rx = c->alloc<double4>(ne);
    d.rv = c->alloc<double4>(ne);
    d.rmesh = c->alloc<int>(ne);
    d.mv = c->alloc<double4>(kMaxMesh);
    d.mw = c->alloc<double4>(kMaxMesh);
    d.nodev = c->alloc<double4>(nn);
    d.rcell = c->alloc<int>(ne);
    d.rsorted = c->alloc<int>(ne);
    d.rcount = c->alloc<int>((size_t)d.ncells + 1);
    d.rstart = c->alloc<int>((size_t)d.ncells + 1);
    const size_t nrch = ((size_t)d.ncells + kScanChunk - 1) / kScanChunk;
    d.rrank = c->alloc<int>(ne);
    d.rloc = c->alloc<int>((size_t)d.ncells + 1);
    d.rtot = c->alloc<int>(nrch);
    d.roff = c->alloc<int>(nrch);
    d.id_orig = c->alloc<int>(np);
    d.ccnt = c->alloc<int>(np);
    d.cidx = c->alloc<int>((size_t)kMaxContact * np);
    d.cforce = c->alloc<double4>(np);
    d.cf_hist = c->alloc<double>((size_t)d.hist_cap * kMaxMesh * 3);
    d.fr_sta = c->fr_sta;
    d.fr_dyn = c->fr_dyn;
    d.dfac = c->dfac;
    d.contact_law = c->contact_law;
    d.pfac = c->pfac;
    d.fric_bound = c->fric_bound;
    d.fric_linear = c->fric_linear;
    d.fr_m = c->fr_m;
    d.fr_b = c->fr_b;
    if (d.contact_law == 0 || d.contact_law == 2) d.cstiff = c->alloc<double>(np);
    d.rT = c->alloc<double>(ne);
    d.heat_cond = c->heat_cond;
    d.heat_fric = c->heat_fric;
    d.hc = c->heat_hc;
    if (d.heat_cond || d.heat_fric) {
      d.qfric = c->alloc<double>(np);
      d.fhfl = c->alloc<double>(np);
      d.cshear = c->alloc<double>(np);
      d.qcont = c->alloc<double>(np);
    }
  }
  // Domain::model_damage: some particle's material has a damage model
  d.damage = 0;
  for (int64_t i = 0; i < c->n; ++i)
    if (c->mat_dam[c->hi[WF_I_MATERIAL][i]]) d.damage = 1;
  if (d.damage) {
    d.damD = c->alloc<double>(np);
    if (!d.id_orig) d.id_orig = c->alloc<int>(np);
  }
  d.ntiles = (int)(np / kBlock);
  d.stage_max = 0;
  d.lanes_mx = c->cfg.lanes ? c->cfg.lanes : kLanesMixed;
  d.lanes_f64 = c->cfg.lanes ? c->cfg.lanes : kLanesFP64;
  // the contact criterion (bit 2) acts only inside AdaptiveTimeStep, which
  // runs with the velocity or acceleration criterion
  d.auto_ts = (c->cfg.auto_ts & 3) ? c->cfg.auto_ts : 0;
  d.ti_on = 0;
  d.beta_on = 0;
  for (size_t k = 0; k < c->mats.size(); ++k) {
    if (c->mat_set[k] && c->mats[k].ti > 0.0) d.ti_on = 1;
    if (c->mat_set[k] && c->mats[k].beta != 0.0) d.beta_on = 1;
  }
  if (d.ti_on) {
    d.tir = c->alloc<double>(6 * np);
    if (d.mixed) d.tirf = c->alloc<float>(6 * np);
  }
  d.cfl = c->cfg.cfl;
  d.sqrt_h_a = c->cfg.sqrt_h_a;
  d.min_ts = c->cfg.min_ts;
  d.locking = c->cfg.locking ? 1 : 0;
  d.gc_on = c->cfg.grad_corr ? 1 : 0;
  if (d.gc_on) {
    // gradCorrM starts as the identity; the first step (or a rebuild) forms it
    d.gcm = c->alloc<double>(6 * np);
    std::vector<double> eye(6 * np, 0.0);
    for (size_t k = 0; k < 3 * np; ++k) eye[k] = 1.0;
    dput(c, d.gcm, eye);
  }
  if (d.mixed) {
    // stage_cnt = -1 marks a tile that gathers from HBM
    d.stage_cnt = c->alloc<int>((size_t)d.ntiles);
    HIPCHK(hipMemsetAsync(d.stage_cnt, 0xff, sizeof(int) * (size_t)d.ntiles, c->st));
  }
  if (d.mixed && c->cfg.stage_max >= 0) {
    d.stage_max = c->cfg.stage_max == 0 ? kStageMax : c->cfg.stage_max;
    d.stage = c->alloc<int>((size_t)d.ntiles * kStageMax);
    d.nbrl = c->alloc<uint16_t>((size_t)nb_rows(c->cfg.max_nb) * np);
    d.selfl = c->alloc<uint16_t>(np);
  }
  // per-group sums: each material index's particles in internal order
  d.ngroups = 0;
  if (c->gsum_zone != -2) {
    const std::vector<int32_t> &mat = c->hi[WF_I_MATERIAL];
    int ng = 0;
    for (int64_t k = 0; k < c->n; ++k) ng = std::max(ng, mat[c->perm[k]] + 1);
    std::vector<int> off(ng + 1, 0), idx(c->n);
    for (int64_t k = 0; k < c->n; ++k) off[mat[c->perm[k]] + 1]++;
    for (int g = 0; g < ng; ++g) off[g + 1] += off[g];
    std::vector<int> cur(off.begin(), off.end() - 1);
    for (int64_t k = 0; k < c->n; ++k) idx[cur[mat[c->perm[k]]]++] = (int)k;
    d.ngroups = ng;
    d.gzone = c->gsum_zone;
    d.ghist_cap = c->gsum_cap;
    d.goff = c->alloc<int>(off.size());
    d.gidx = c->alloc<int>(idx.size());
    dput(c, d.goff, off);
    dput(c, d.gidx, idx);
    d.gr_hist = c->alloc<double>((size_t)d.ghist_cap * ng * 3);
    if (d.contact) {
      d.gcf = c->alloc<double4>((size_t)kMaxMesh * np);
      d.gc_hist = c->alloc<double>((size_t)d.ghist_cap * ng * kMaxMesh * 3);
    }
  }
}

void upload_all(Ctx *c) {
  Dev &d = c->d;
  const int64_t n = c->n;
  const size_t np = (size_t)c->npad;
  auto &F = c->hf;
  auto field = [&](int f) -> const double * { return F[f].empty() ? nullptr : F[f].data(); };
  // positions, h, velocities, mass
  std::vector<double4> xh(np), vm(np), z4(np);
  for (int64_t k = 0; k < n; ++k) {
    const int o = c->perm[k];
    xh[k] = make_double4(F[WF_F_X][3 * o], F[WF_F_X][3 * o + 1], F[WF_F_X][3 * o + 2], F[WF_F_H][o]);
    const double *v = field(WF_F_V);
    vm[k] = make_double4(v ? v[3 * o] : 0.0, v ? v[3 * o + 1] : 0.0, v ? v[3 * o + 2] : 0.0, F[WF_F_MASS][o]);
  }
  dput(c, d.xh[0], xh);
  dput(c, d.xh[1], xh);
  dput(c, d.vm, vm);
  {
    std::vector<double> pm(vm.size());
    for (size_t k = 0; k < vm.size(); ++k) pm[k] = vm[k].w;
    dput(c, d.pmass, pm);
  }
  if (field(WF_F_A)) put_d4(c, d.acc, field(WF_F_A), 0, 3, 3);
  if (field(WF_F_DISP)) put_d4(c, d.disp, field(WF_F_DISP), 0, 3, 3);
  // density, volumes, stress records, pressure (InitialChecks,
  // Source/Domain.cpp:2097-2113: Pressure = EOS(Density))
  const std::vector<int32_t> &mt = c->hi[WF_I_MATERIAL];
  std::vector<double> rho(np, 0.0), vol(np, 0.0), pres(np, 0.0);
  std::vector<StressRec> rec(np);
  const double *sig = field(WF_F_SIGMA);
  for (int64_t k = 0; k < n; ++k) {
    const int o = c->perm[k];
    const double r = F[WF_F_DENSITY][o];
    const wf_material &M = c->mats[mt[o]];
    rho[k] = r;
    vol[k] = F[WF_F_MASS][o] / r;
    pres[k] = M.P0 + (M.Cs * M.Cs) * (r - M.ref_density);
    StressRec &R = rec[k];
    const int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
    for (int q = 0; q < 6; ++q) R.s[q] = sig ? sig[9 * (size_t)o + 3 * ia[q] + ib[q]] : 0.0;
    R.rho = r;
    R.irho2 = 1.0 / (r * r);
  }
  dput(c, d.rho[0], rho);
  dput(c, d.rho[1], rho);
  dput(c, d.vol[0], vol);
  dput(c, d.vol[1], vol);
  dput(c, d.pres, pres);
  dput(c, d.srec, rec);
  if (d.damD && field(WF_F_DAMAGE)) put_scalar(c, d.damD, field(WF_F_DAMAGE));
  if (d.damage && !d.contact) {
    // CalculateSurface resets ID to ID_orig, which SolveDiffUpdateLeapFrog
    // sets only with contact (Source/SolverLeapfrog.cpp:80-83); with damage
    // it is taken from ID here too.  The mean mass of the surface test runs
    // over the solid particles.
    std::vector<int> io(np, 0);
    for (int64_t k = 0; k < n; ++k) io[k] = c->hi[WF_I_ID][c->perm[k]];
    dput(c, d.id_orig, io);
    double tm = 0.0;
    for (int64_t o = 0; o < n; ++o) tm += F[WF_F_MASS][o];
    d.totmass = tm / (double)n;
  }
  if (d.cstiff) {
    // Particle::cont_stiff: SolveDiffUpdateLeapFrog never sets it; the
    // formula SolveDiffUpdateKickDrift applies before its loop
    // (Source/SolverKickDrift.cpp:79-84, Fraser 3-153), from the initial
    // density
    std::vector<double> ks(np, 0.0);
    for (int64_t k = 0; k < n; ++k) {
      const int o = c->perm[k];
      const wf_material &M = c->mats[mt[o]];
      const double r = F[WF_F_DENSITY][o];
      const double bulk = M.Cs * M.Cs * r;
      const double dS = std::pow(F[WF_F_MASS][o] / r, 0.33333);
      ks[k] = 9. * bulk * M.G / (3. * bulk + M.G) * dS;
    }
    dput(c, d.cstiff, ks);
  }
  if (d.contact) {
    std::vector<double4> node(d.nnode), en(d.nrig), rx(d.nrig), z(d.nrig);
    std::vector<int4> el(d.nrig);
    std::vector<int> rm(d.nrig);
    for (int m = 0; m < d.nmesh; ++m) {
      const MeshHost &M = c->meshes[m];
      const int n0 = d.mesh_node0[m], e0 = d.mesh_elem0[m];
      for (int q = 0; q < d.mesh_nnode[m]; ++q)
        node[n0 + q] = make_double4(M.node[3 * q], M.node[3 * q + 1], M.node[3 * q + 2], 0.0);
      for (int e = 0; e < d.mesh_nelem[m]; ++e) {
        el[e0 + e] = make_int4(n0 + M.el[3 * e], n0 + M.el[3 * e + 1], n0 + M.el[3 * e + 2], M.nfar[e]);
        en[e0 + e] = make_double4(M.normal[3 * e], M.normal[3 * e + 1], M.normal[3 * e + 2], M.pplane[e]);
        rx[e0 + e] = make_double4(M.centroid[3 * e], M.centroid[3 * e + 1], M.centroid[3 * e + 2], M.h[e]);
        z[e0 + e] = make_double4(0, 0, 0, 0);
        rm[e0 + e] = m;
      }
    }
    dput(c, d.node, node);
    dput(c, d.elem, el);
    dput(c, d.enorm, en);
    dput(c, d.rx, rx);
    dput(c, d.rv, z);
    dput(c, d.rmesh, rm);
    if (!c->rigid_T.empty()) {
      if ((int)c->rigid_T.size() != d.nrig) throw Error(WF_EINVAL, "rigid temperatures: one per rigid particle");
      dput(c, d.rT, c->rigid_T);
    }
    // ID_orig = ID before the first free-surface search
    // (Source/SolverLeapfrog.cpp:80-83); the mean mass of CalculateSurface
    // runs over every particle in the reference's order, rigid ones with 0
    std::vector<int> io(np, 0);
    for (int64_t k = 0; k < n; ++k) io[k] = c->hi[WF_I_ID][c->perm[k]];
    dput(c, d.id_orig, io);
    double tm = 0.0;
    for (int64_t o = 0; o < n; ++o) tm += F[WF_F_MASS][o];
    for (int r = 0; r < d.nrig; ++r) tm += 0.0;
    d.totmass = tm / (double)(n + d.nrig);
  }
  // tensors
  if (field(WF_F_SHEAR)) put_sym(c, d.S, field(WF_F_SHEAR));
  if (field(WF_F_STRAIN)) put_sym(c, d.strain, field(WF_F_STRAIN));
  if (field(WF_F_STRAIN_PL)) put_sym(c, d.epl, field(WF_F_STRAIN_PL));
  if (field(WF_F_STRAIN_PL_INCR)) put_sym(c, d.eplinc, field(WF_F_STRAIN_PL_INCR));
  if (field(WF_F_STRAIN_RATE)) put_sym(c, d.sr, field(WF_F_STRAIN_RATE));
  // scalars
  struct {
    int f;
    double *p;
  } sc[] = {{WF_F_PL_STRAIN, d.pl}, {WF_F_DPL_STRAIN, d.dpl}, {WF_F_SIGMAY, d.sigy}, {WF_F_EFF_SR, d.effsr},
            {WF_F_T, d.T},          {WF_F_DTDT, d.dTdt},     {WF_F_Q_PLHEAT, d.qpl}, {WF_F_Q_CONV, d.qconv},
            {WF_F_H_CONV, d.hconv}, {WF_F_T_INF, d.tinf},    {WF_F_Q_SOURCE, d.qsrc}, {WF_F_PS_ENERGY, d.psen}};
  for (auto &s : sc)
    if (field(s.f)) put_scalar(c, s.p, field(s.f));
  // ints
  std::vector<int> id(np, 0), mat(np, 0), grp(np, -1), thbc(np, 0), perm(np, 0);
  for (int64_t k = 0; k < n; ++k) {
    const int o = c->perm[k];
    id[k] = c->hi[WF_I_ID][o];
    mat[k] = c->hi[WF_I_MATERIAL][o];
    grp[k] = c->hi[WF_I_FREE][o] ? mat[k] : -1;
    thbc[k] = c->hi[WF_I_THERMAL_BC][o];
    perm[k] = o;
  }
  dput(c, d.id, id);
  dput(c, d.mat, mat);
  dput(c, d.group, grp);
  dput(c, d.thbc, thbc);
  dput(c, d.perm, perm);
  // per-material origin of the FP32 position shadows
  {
    std::vector<double> lo(3 * kMaxMat, 1e300), hi(3 * kMaxMat, -1e300);
    const double *X = F[WF_F_X].data();
    for (int64_t o = 0; o < n; ++o) {
      const int m = c->hi[WF_I_MATERIAL][o];
      for (int q = 0; q < 3; ++q) {
        lo[3 * m + q] = std::min(lo[3 * m + q], X[3 * o + q]);
        hi[3 * m + q] = std::max(hi[3 * m + q], X[3 * o + q]);
      }
    }
    c->mat_xc.assign(3 * kMaxMat, 0.0);
    for (int m = 0; m < kMaxMat; ++m)
      if (lo[3 * m] <= hi[3 * m])
        for (int q = 0; q < 3; ++q) c->mat_xc[3 * m + q] = 0.5 * (lo[3 * m + q] + hi[3 * m + q]);
  }
  std::vector<DevMat> dm(kMaxMat);
  for (size_t k = 0; k < c->mats.size() && k < (size_t)kMaxMat; ++k)
    if (c->mat_set[k]) dm[k] = to_dev(c->mats[k]);
  for (int m = 0; m < kMaxMat; ++m) {
    for (int q = 0; q < 3; ++q) dm[m].xc[q] = c->mat_xc[3 * m + q];
    dm[m].fr_sta = c->mat_fr[2 * m];
    dm[m].fr_dyn = c->mat_fr[2 * m + 1];
    dm[m].dam_model = c->mat_dam[m];
    for (int q = 0; q < 5; ++q) dm[m].dam[q] = c->mat_dpar[5 * m + q];
  }
  dput(c, c->dmats, dm);
  // device schedule state: the pre-loop ClearNbData leaves the data cleared
  DevFlags fl{};
  fl.cleared = 1;
  fl.ts_nb_inc = c->cfg.ts_nb_inc;
  fl.dt = fl.dt_prev = fl.dtint = c->dt;
  fl.min_force_ts = c->dt;   // Source/SolverLeapfrog.cpp:95
  const double *pl = field(WF_F_PL_STRAIN);
  if (pl)
    for (int64_t i = 0; i < n; ++i)
      if (pl[i] > 1.e-6) fl.yielded = 1;
  std::vector<DevFlags> fv(1, fl);
  dput(c, d.flags, fv);
}

// ---------------------------------------------------------------------------
// timing

hipEvent_t next_event(Ctx *c) {
  if (c->ev_next == c->ev_pool.size()) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    c->ev_pool.push_back(e);
  }
  return c->ev_pool[c->ev_next++];
}

struct Timed {
  Ctx *c;
  int k;
  hipEvent_t a{}, b{};
  Timed(Ctx *cc, int kk) : c(cc), k(kk) {
    if (c->timing) {
      a = next_event(c);
      b = next_event(c);
      HIPCHK(hipEventRecord(a, c->st));
    }
  }
  ~Timed() noexcept(false) {
    if (c->timing) {
      HIPCHK(hipEventRecord(b, c->st));
      c->recs.push_back({k, a, b});
    }
  }
};

void harvest_timing(Ctx *c) {
  if (c->recs.empty()) return;
  HIPCHK(hipStreamSynchronize(c->st));
  for (auto &r : c->recs) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
    c->tsum[r.k] += ms;
    c->tcnt[r.k] += 1;
  }
  c->recs.clear();
  c->ev_next = 0;
}

// BC and amplitude tables for the device; amplitudes are evaluated there
void refresh_bc(Ctx *c) {
  if (!c->bc_dirty) return;
  if ((int)c->amps.size() > kMaxAmp) throw Error(WF_EUNSUPPORTED, "more than 8 amplitude tables");
  std::vector<DevAmp> da(kMaxAmp);
  for (size_t k = 0; k < c->amps.size(); ++k) {
    const AmpHost &a = c->amps[k];
    if ((int)a.t.size() > kMaxAmpPts) throw Error(WF_EUNSUPPORTED, "amplitude with more than 16 points");
    da[k].id = a.id;
    da[k].n = (int)a.t.size();
    for (size_t q = 0; q < a.t.size(); ++q) {
      da[k].t[q] = a.t[q];
      da[k].v[q] = a.v[q];
    }
  }
  dput(c, c->damps, da);
  BcTable t{};
  for (const BcHost &b : c->bcs) {
    const int k = t.n++;
    t.zone[k] = b.zone;
    t.type[k] = b.type;
    t.value_type[k] = b.type == 0 ? b.value_type : 0;
    t.T[k] = b.T;
    t.amp_factor[k] = b.amp_factor;
    t.amp[k] = -1;
    // UserAcc applies the last amplitude with a matching id (Source/WeldForm.cpp:74-81)
    for (size_t q = 0; q < c->amps.size(); ++q)
      if (c->amps[q].id == b.amp_id) t.amp[k] = (int)q;
    for (int q = 0; q < 3; ++q) {
      t.v[k][q] = b.value[q];
      t.w[k][q] = b.w[q];
    }
  }
  t.nz = (int)c->rzones.size();
  for (int z = 0; z < t.nz; ++z) t.rzone[z] = c->rzones[z];
  c->bc = t;
  c->bc_dirty = false;
  c->drop_graphs();
}

void check_flags(Ctx *c) {
  auto f = dget(c, c->d.flags, 1)[0];
  if (f.overflow)
    throw Error(WF_ENBOVERFLOW, "a particle has more neighbours than max_nb (" + std::to_string(c->cfg.max_nb) +
                                    "); largest count " + std::to_string(f.max_count));
  if (f.nan) throw Error(WF_ENAN, "non-finite value in the particle state");
}

// adjoint tape of a replayed step: slot `slot` receives the state at the
// step's start and the step's intermediates (AdjSlot storage)
struct TapeHook {
  AdjSlot s;
  int slot;
  int force_lists;  // copy the lists even if the step does not search
  int *ver;
};

void d2d(Ctx *c, void *dst, const void *src, size_t bytes) {
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->st));
}

// the state the next step starts from: x, v, rho, S, Sy, pl
void tape_state(Ctx *c, const AdjSlot &s, int cur) {
  const Dev &d = c->d;
  const size_t np = (size_t)c->npad;
  d2d(c, s.xh, d.xh[cur], np * sizeof(double4));
  d2d(c, s.vm, d.vm, np * sizeof(double4));
  d2d(c, s.rho, d.rho[cur], np * sizeof(double));
  d2d(c, s.S, d.S, 6 * np * sizeof(double));
  d2d(c, s.sigy, d.sigy, np * sizeof(double));
  d2d(c, s.pl, d.pl, np * sizeof(double));
}

// One iteration of the Leapfrog / KickDrift loop, all on the stream.  The
// iteration is cut into parts where the reference calls its host hooks
// (Source/SolverLeapfrog.cpp:215-273, Source/SolverKickDrift.cpp:195-264):
//   P_SEARCH  neighbour schedule and search            -> GeneralBefore
//   P_ACCEL   CalcAccel + AccelReduction               -> GeneralAfter
//   P_KICK1   contact; Leapfrog: BCs and the first-step half kick;
//             KickDrift: the time step and v += a dt_prev / 2  -> GeneralAfter
//   P_KICK2   Leapfrog: the time step and v += a (dt_prev + dt) / 2;
//             KickDrift: density, drift, v += a dt / 2 -> GeneralAfter
//   P_REST    the rest of the step (Leapfrog: density and drift first)
// part < 0 enqueues the whole iteration (the captured graph); part >= 0
// enqueues one part for a caller that runs host hooks between the parts,
// with the reaction observed at the end of P_ACCEL, before any hook.
enum { P_SEARCH = 0, P_ACCEL, P_KICK1, P_KICK2, P_REST, P_PARTS };

void enqueue_step(Ctx *c, int scheme, int cur, const TapeHook *tp = nullptr, int part = -1) {
  Dev &d = c->d;
  Timed whole(c, T_STEP);
  const size_t np = (size_t)c->npad;
  const bool split = part >= 0;
  auto in = [&](int p) { return part < 0 || part == p; };
  // dt < 0: every kernel reads the step size from the device flags, which
  // the adaptive step updates between the two halves of the kick
  // (Source/SolverLeapfrog.cpp:241-262, Source/SolverKickDrift.cpp:208-226)
  const double dt = d.auto_ts ? -1.0 : c->dt;
  const int nxt = cur ^ 1;
  if (in(P_SEARCH)) {
    // StartAcceleration's plane-strain reset at the top of the step
    // (Source/SolverLeapfrog.cpp:126, Source/SolverKickDrift.cpp:119)
    if (d.dim == 2) launch_plane_strain(d, c->st);
    if (tp) tape_state(c, tp->s, cur);
    launch_schedule(d, c->st, 0);
    {
      Timed t(c, T_NBR);
      launch_neighbours(d, c->st, cur);
    }
    {
      Timed t(c, T_SURF);
      if (d.gc_on) launch_gradcorr(d, c->st, cur);
      if (d.contact) launch_surface(d, c->st, cur, 0);
      // with damage the free surface is refreshed every step
      // (Source/SolverLeapfrog.cpp:204-206)
      if (d.damage) launch_surface(d, c->st, cur, 2);
    }
    if (tp) launch_tape_lists(d, c->st, tp->s, tp->force_lists, tp->ver, tp->slot);
  }
  if (in(P_ACCEL)) {
    // StartAcceleration's tensile-instability stress (Source/Domain.cpp:1529-1563)
    if (d.ti_on) launch_tir(d, c->st, cur);
    {
      Timed t(c, T_ACCEL);
      launch_accel(d, c->st, cur);
    }
    if (d.ngroups > 0 && d.gzone >= 0) launch_group_sums(d, c->st, 0);
    if (tp && d.contact) d2d(c, tp->s.apre, d.acc, np * sizeof(double4));
    if (split) launch_kick(d, c->st, c->bc, scheme, dt, dt, K_OBSERVE);
  }
  if (in(P_KICK1)) {
    // contact acts after AccelReduction and the first GeneralAfter
    // (Source/SolverLeapfrog.cpp:228-240); free-surface particles carry ID 1,
    // so no BC zone applies to them
    // SolveDiffUpdateKickDrift calls Wang's law whatever contact_alg says
    // (Source/SolverKickDrift.cpp:211)
    if (d.contact) launch_contact(d, c->st, c->bc, cur, dt, scheme == WF_LEAPFROG ? d.contact_law : 1);
    if (d.contact && d.ngroups > 0) launch_group_sums(d, c->st, 1);
    if (tp && d.contact) {
      d2d(c, tp->s.node, d.node, (size_t)d.nnode * sizeof(double4));
      d2d(c, tp->s.enorm, d.enorm, (size_t)d.nrig * sizeof(double4));
      d2d(c, tp->s.rv, d.rv, (size_t)d.nrig * sizeof(double4));
    }
    const int obs = split ? 0 : 1;
    if (scheme == WF_LEAPFROG) {
      if (!split && !d.auto_ts) {
        launch_kick(d, c->st, c->bc, scheme, dt, dt, K_WHOLE);
      } else {
        launch_kick(d, c->st, c->bc, scheme, dt, dt, obs ? K_FIRST : K_FIRST_NOOBS);
      }
    } else {
      if (d.auto_ts) {
        launch_kick(d, c->st, c->bc, scheme, dt, dt, obs ? K_FIRST : K_FIRST_NOOBS);
        launch_time_step(d, c->st, cur);
        launch_kick(d, c->st, c->bc, scheme, dt, dt, K_SECOND);
      } else {
        launch_kick(d, c->st, c->bc, scheme, dt, dt, obs ? K_WHOLE : K_WHOLE_NOOBS);
      }
    }
  }
  if (in(P_KICK2)) {
    if (scheme == WF_LEAPFROG) {
      if (d.auto_ts) launch_time_step(d, c->st, cur);
      if (split || d.auto_ts) launch_kick(d, c->st, c->bc, scheme, dt, dt, K_SECOND);
    } else {
      {
        Timed t(c, T_DENS);
        launch_dens_drift(d, c->st, cur, dt, 1);
      }
      launch_kick2(d, c->st, dt);
      if (d.contact) launch_rigid_drift(d, c->st, dt);
    }
  }
  if (in(P_REST)) {
    if (scheme == WF_LEAPFROG) {
      {
        Timed t(c, T_DENS);
        launch_dens_drift(d, c->st, cur, dt, 1);
      }
      if (d.contact) launch_rigid_drift(d, c->st, dt);
    }
    {
      Timed t(c, T_RATES);
      launch_rates(d, c->st, nxt);
    }
    launch_stress(d, c->st, nxt, dt);
    if (c->cfg.thermal_solver) launch_thermal(d, c->st, nxt, dt);
    launch_finalize(d, c->st, dt, c->bc.nz);
    if (d.contact) launch_mesh_update(d, c->st, dt);
    if (tp) {
      d2d(c, tp->s.sr, d.sr, 6 * np * sizeof(double));
      d2d(c, tp->s.rr, d.rr, 3 * np * sizeof(double));
      d2d(c, tp->s.drho, d.drho, np * sizeof(double));
    }
  }
}

// the captured step of (scheme, buffer parity cur), instantiated on first use
hipGraphExec_t step_graph(Ctx *c, int scheme, int cur) {
  hipGraphExec_t &ex = c->gexec[scheme][cur];
  if (!ex) {
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
    enqueue_step(c, scheme, cur);
    HIPCHK(hipStreamEndCapture(c->st, &g));
    HIPCHK(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(g));
  }
  return ex;
}

void one_step(Ctx *c, int scheme) {
  refresh_bc(c);
  if (c->timing || c->cfg.no_graphs) {
    enqueue_step(c, scheme, c->cur);
    HIPCHK(hipGetLastError());
  } else {
    HIPCHK(hipGraphLaunch(step_graph(c, scheme, c->cur), c->st));
  }
  c->built = true;
  c->cur ^= 1;
  c->steps += 1;
}

// ---------------------------------------------------------------------------
// adjoint: snapshots of the whole device state, the tape of a replayed
// segment and the reverse sweep

void adj_check(Ctx *c) {
  if (!c->inited) throw Error(WF_ESTATE, "wf_init not called");
  if (c->cfg.precision != WF_FP64) throw Error(WF_EUNSUPPORTED, "the adjoint needs precision WF_FP64");
  if (c->cfg.auto_ts) throw Error(WF_EUNSUPPORTED, "the adjoint needs a fixed time step (auto_ts 0)");
  if (c->cfg.thermal_solver) throw Error(WF_EUNSUPPORTED, "the adjoint does not cover the thermal solver");
  if (c->d.ti_on) throw Error(WF_EUNSUPPORTED, "the adjoint does not cover the tensile-instability term");
  if (c->d.gc_on || c->d.locking)
    throw Error(WF_EUNSUPPORTED, "the adjoint covers the Nishimura sums without kernel-gradient correction");
  for (int m = 0; m < kMaxMat; ++m)
    if (c->mat_fr[2 * m] >= 0.0)
      throw Error(WF_EUNSUPPORTED, "the adjoint covers the context's friction, not per-material friction");
  if (c->d.contact && (c->contact_law != 1 || c->fric_linear))
    throw Error(WF_EUNSUPPORTED, "the adjoint covers Wang's contact law with constant friction");
  if (c->d.damage) throw Error(WF_EUNSUPPORTED, "the adjoint does not cover damage");
  if (c->d.dim == 2) throw Error(WF_EUNSUPPORTED, "the adjoint covers the 3D step");
}

inline size_t round256(size_t b) { return (b + 255) / 256 * 256; }

void snapshot(Ctx *c, size_t k) {
  char *dst = c->adj.ckpt[k];
  size_t off = 0;
  for (auto &a : c->allocs) {
    d2d(c, dst + off, a.first, a.second);
    off += round256(a.second);
  }
  c->adj.ckh[k] = {c->cur, c->steps};
}

void restore(Ctx *c, size_t k) {
  const char *src = c->adj.ckpt[k];
  size_t off = 0;
  for (auto &a : c->allocs) {
    d2d(c, a.first, src + off, a.second);
    off += round256(a.second);
  }
  c->cur = c->adj.ckh[k].first;
  c->steps = c->adj.ckh[k].second;
}

AdjSlot alloc_slot(Ctx *c, bool inter) {
  const Dev &d = c->d;
  const size_t np = (size_t)c->npad;
  AdjSlot s{};
  s.xh = c->adj_alloc<double4>(np);
  s.vm = c->adj_alloc<double4>(np);
  s.rho = c->adj_alloc<double>(np);
  s.S = c->adj_alloc<double>(6 * np);
  s.sigy = c->adj_alloc<double>(np);
  s.pl = c->adj_alloc<double>(np);
  if (!inter) return s;
  s.apre = c->adj_alloc<double4>(d.contact ? np : 1);
  s.drho = c->adj_alloc<double>(np);
  s.sr = c->adj_alloc<double>(6 * np);
  s.rr = c->adj_alloc<double>(3 * np);
  s.nbr = c->adj_alloc<int>((size_t)nb_rows(c->cfg.max_nb) * np);
  s.nnb = c->adj_alloc<int>(np);
  s.id = c->adj_alloc<int>(np);
  s.cidx = c->adj_alloc<int>(d.contact ? (size_t)kMaxContact * np : 1);
  s.ccnt = c->adj_alloc<int>(d.contact ? np : 1);
  s.node = c->adj_alloc<double4>(d.contact ? (size_t)d.nnode : 1);
  s.enorm = c->adj_alloc<double4>(d.contact ? (size_t)d.nrig : 1);
  s.rv = c->adj_alloc<double4>(d.contact ? (size_t)d.nrig : 1);
  return s;
}

void adj_prepare(Ctx *c, int k, size_t ncp) {
  auto &A = c->adj;
  size_t bytes = 0;
  for (auto &a : c->allocs) bytes += round256(a.second);
  A.snap_bytes = bytes;
  while (A.ckpt.size() < ncp) A.ckpt.push_back(c->adj_alloc<char>(bytes));
  A.ckh.resize(A.ckpt.size());
  if (A.seg != k) {
    A.slot.clear();
    for (int s = 0; s <= k; ++s) A.slot.push_back(alloc_slot(c, s < k));
    A.ver = c->adj_alloc<int>((size_t)k);
    A.seg = k;
  }
  if (!A.buf.xb) {
    const size_t np = (size_t)c->npad;
    AdjBufs &b = A.buf;
    b.xb = c->adj_alloc<double4>(np);
    b.vb = c->adj_alloc<double4>(np);
    b.ab = c->adj_alloc<double4>(np);
    b.vpb = c->adj_alloc<double4>(np);
    b.rhob = c->adj_alloc<double>(np);
    b.Sb = c->adj_alloc<double>(6 * np);
    b.syb = c->adj_alloc<double>(np);
    b.plb = c->adj_alloc<double>(np);
    b.rsb = c->adj_alloc<double4>(3 * np);
    b.tsig = c->adj_alloc<double4>(2 * np);
    b.db = c->adj_alloc<double>(np);
    b.thb = c->adj_alloc<double>((size_t)kAdjParams * np);
    b.part = c->adj_alloc<double>((size_t)kAdjParams * c->d.nblocks);
    A.grad = c->adj_alloc<double>(kAdjParams);
  }
}

int zone_slot(Ctx *c, int zone) {
  for (size_t z = 0; z < c->rzones.size(); ++z)
    if (c->rzones[z] == zone) return (int)z;
  return -1;
}

}  // namespace

// ===========================================================================
// exported C-ABI

extern "C" {

int wf_abi_version(void) { return WF_ABI_VERSION; }

const char *wf_last_error(const void *ctx) {
  if (ctx) return ((const Ctx *)ctx)->err.c_str();
  return g_err.c_str();
}

int wf_device_count(int32_t *count) {
  return guard(nullptr, [&] {
    if (!count) throw Error(WF_EINVAL, "null argument");
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    *count = n;
    return WF_OK;
  });
}

int wf_create(const wf_config *cfg, int64_t n, void **out) {
  return guard(nullptr, [&] {
    if (!cfg || !out) throw Error(WF_EINVAL, "null argument");
    if (n <= 0 || n > (1LL << 31) - 1024) throw Error(WF_EINVAL, "particle count out of range");
    if (cfg->max_nb < 1 || cfg->max_nb > kMaxNb) throw Error(WF_EINVAL, "max_nb must be in 1..128");
    if (cfg->precision != WF_FP64 && cfg->precision != WF_MIXED) throw Error(WF_EINVAL, "bad precision");
    if (!(cfg->cellfac > 0)) throw Error(WF_EINVAL, "cellfac must be positive");
    if (cfg->ts_nb_inc < 1) throw Error(WF_EINVAL, "ts_nb_inc must be >= 1");
    if (cfg->lanes != 0 && cfg->lanes != 1 && cfg->lanes != 2 && cfg->lanes != 4)
      throw Error(WF_EINVAL, "lanes must be 0, 1, 2 or 4");
    if (cfg->stage_max < -1 || cfg->stage_max > kStageMax)
      throw Error(WF_EINVAL, "stage_max must be -1, 0 or 1.." + std::to_string(kStageMax));
    if (cfg->auto_ts < 0 || cfg->auto_ts > 7) throw Error(WF_EINVAL, "auto_ts must be 0..7");
    if (cfg->auto_ts && !(cfg->cfl > 0 && cfg->sqrt_h_a > 0 && cfg->min_ts >= 0))
      throw Error(WF_EINVAL, "auto_ts needs cfl > 0, sqrt_h_a > 0 and min_ts >= 0");
    Ctx *c = new Ctx();
    c->cfg = *cfg;
    c->n = n;
    c->npad = (n + kBlock - 1) / kBlock * kBlock;
    c->mats.resize(kMaxMat);
    c->mat_set.assign(kMaxMat, 0);
    c->mat_fr.assign(2 * kMaxMat, -1.0);
    c->mat_dam.assign(kMaxMat, 0);
    c->mat_dpar.assign(5 * kMaxMat, 0.0);
    for (int f = 0; f < WF_NIFIELDS; ++f) c->hi[f].assign(n, 0);
    c->hi[WF_I_FREE].assign(n, 1);
    try {
      HIPCHK(hipSetDevice(cfg->device));
      HIPCHK(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
    } catch (...) {
      delete c;
      throw;
    }
    *out = c;
    return WF_OK;
  });
}

int wf_destroy(void *ctx) {
  if (!ctx) return WF_OK;
  Ctx *c = (Ctx *)ctx;
  if (c->st) (void)hipStreamSynchronize(c->st);
  delete c;
  return WF_OK;
}

int wf_set_material(void *ctx, int32_t index, const wf_material *m) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (!m || index < 0 || index >= kMaxMat) throw Error(WF_EINVAL, "material index out of range");
    if (m->model < WF_BILINEAR || m->model > WF_JOHNSON_COOK) throw Error(WF_EINVAL, "unknown material model");
    if (m->model == WF_HOLLOMON) {
      // Hollomon constructor check (Source/Material.cpp:48-57)
      const double eps0 = m->hol_sy0 / m->E, eps1 = std::pow(m->hol_sy0 / m->hol_K, 1. / m->hol_m);
      if (eps0 > eps1)
        throw Error(WF_EINVAL,
                    "Hollomon material bad definition, please correct Yield Stress, Elastic Modulus or Material "
                    "hardening constants.");
    }
    if (!(m->ti >= 0.0)) throw Error(WF_EINVAL, "ti must be >= 0");
    if (m->ti > 0.0 && !(m->ti_dist >= 0.0 && std::isfinite(m->ti_n)))
      throw Error(WF_EINVAL, "tensile instability needs ti_dist >= 0 and a finite ti_n");
    if (c->inited && m->ti > 0.0 && !c->d.tir)
      throw Error(WF_ESTATE, "tensile instability must be switched on before wf_init");
    c->mats[index] = *m;
    c->mat_set[index] = 1;
    if (c->inited) {
      std::vector<DevMat> one(1, to_dev(*m));
      for (int q = 0; q < 3; ++q) one[0].xc[q] = c->mat_xc[3 * index + q];
      one[0].fr_sta = c->mat_fr[2 * index];
      one[0].fr_dyn = c->mat_fr[2 * index + 1];
      one[0].dam_model = c->mat_dam[index];
      for (int q = 0; q < 5; ++q) one[0].dam[q] = c->mat_dpar[5 * index + q];
      HIPCHK(hipMemcpyAsync(c->dmats + index, one.data(), sizeof(DevMat), hipMemcpyHostToDevice, c->st));
      HIPCHK(hipStreamSynchronize(c->st));
      // the acceleration kernel is chosen by whether a material has beta != 0
      int beta_on = 0;
      for (size_t k = 0; k < c->mats.size(); ++k)
        if (c->mat_set[k] && c->mats[k].beta != 0.0) beta_on = 1;
      if (beta_on != c->d.beta_on) {
        c->d.beta_on = beta_on;
        c->drop_graphs();
      }
    }
    return WF_OK;
  });
}

int wf_set_field(void *ctx, int32_t field, const double *h) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (field < 0 || field >= WF_NFIELDS || !h) throw Error(WF_EINVAL, "bad field");
    if (field == WF_F_CONTFORCE || field == WF_F_GRADCORR || field == WF_F_ZWAB || field == WF_F_VXSPH ||
        field == WF_F_SUMDEN)
      throw Error(WF_EINVAL, "read-only field");
    const size_t cnt = (size_t)kFieldDim[field] * c->n;
    c->hf[field].assign(h, h + cnt);
    if (!c->inited) return WF_OK;
    Dev &d = c->d;
    switch (field) {
      case WF_F_X:
        put_d4(c, d.xh[c->cur], h, 0, 3, 3);
        break;
      case WF_F_H:
        put_d4(c, d.xh[0], h, 3, 1, 1);
        put_d4(c, d.xh[1], h, 3, 1, 1);
        break;
      case WF_F_V:
        put_d4(c, d.vm, h, 0, 3, 3);
        break;
      case WF_F_A:
        put_d4(c, d.acc, h, 0, 3, 3);
        break;
      case WF_F_DISP:
        put_d4(c, d.disp, h, 0, 3, 3);
        break;
      case WF_F_MASS:
      case WF_F_DENSITY:
      case WF_F_SIGMA: {
        // mass, density and stress feed vol = m/rho and the stress record
        if (field == WF_F_MASS) put_d4(c, d.vm, h, 3, 1, 1);
        auto vm = dget(c, d.vm, (size_t)c->npad);
        if (field == WF_F_MASS) {
          std::vector<double> pm(vm.size());
          for (size_t k = 0; k < vm.size(); ++k) pm[k] = vm[k].w;
          dput(c, d.pmass, pm);
        }
        std::vector<double> rho((size_t)c->npad);
        if (field == WF_F_DENSITY)
          for (int64_t k = 0; k < c->n; ++k) rho[k] = h[c->perm[k]];
        else
          rho = dget(c, d.rho[c->cur], (size_t)c->npad);
        auto rec = dget(c, d.srec, (size_t)c->npad);
        std::vector<double> vol((size_t)c->npad, 0.0);
        for (int64_t k = 0; k < c->n; ++k) {
          vol[k] = vm[k].w / rho[k];
          rec[k].rho = rho[k];
          rec[k].irho2 = 1.0 / (rho[k] * rho[k]);
          if (field == WF_F_SIGMA) {
            const double *m = h + 9 * (size_t)c->perm[k];
            const int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
            for (int q = 0; q < 6; ++q) rec[k].s[q] = m[3 * ia[q] + ib[q]];
          }
        }
        dput(c, d.rho[c->cur], rho);
        dput(c, d.vol[c->cur], vol);
        dput(c, d.srec, rec);
        break;
      }
      case WF_F_SHEAR: put_sym(c, d.S, h); break;
      case WF_F_STRAIN_RATE: put_sym(c, d.sr, h); break;
      case WF_F_ROT_RATE: {
        std::vector<double> t((size_t)3 * c->npad, 0.0);
        for (int64_t k = 0; k < c->n; ++k) {
          const double *m = h + 9 * (size_t)c->perm[k];
          t[k] = m[1];
          t[(size_t)c->npad + k] = m[2];
          t[(size_t)2 * c->npad + k] = m[5];
        }
        dput(c, d.rr, t);
        break;
      }
      case WF_F_STRAIN: put_sym(c, d.strain, h); break;
      case WF_F_STRAIN_PL: put_sym(c, d.epl, h); break;
      case WF_F_STRAIN_PL_INCR: put_sym(c, d.eplinc, h); break;
      case WF_F_PL_STRAIN: put_scalar(c, d.pl, h); break;
      case WF_F_DPL_STRAIN: put_scalar(c, d.dpl, h); break;
      case WF_F_SIGMAY: put_scalar(c, d.sigy, h); break;
      case WF_F_EFF_SR: put_scalar(c, d.effsr, h); break;
      case WF_F_PRESSURE: put_scalar(c, d.pres, h); break;
      case WF_F_T: put_scalar(c, d.T, h); break;
      case WF_F_DTDT: put_scalar(c, d.dTdt, h); break;
      case WF_F_DDENSITY: put_scalar(c, d.drho, h); break;
      case WF_F_Q_PLHEAT: put_scalar(c, d.qpl, h); break;
      case WF_F_Q_CONV: put_scalar(c, d.qconv, h); break;
      case WF_F_H_CONV: put_scalar(c, d.hconv, h); break;
      case WF_F_T_INF: put_scalar(c, d.tinf, h); break;
      case WF_F_Q_SOURCE: put_scalar(c, d.qsrc, h); break;
      case WF_F_PS_ENERGY: put_scalar(c, d.psen, h); break;
      case WF_F_DAMAGE:
        if (!d.damD) throw Error(WF_EINVAL, "no damage model set (wf_set_material_damage)");
        put_scalar(c, d.damD, h);
        break;
      default: throw Error(WF_EINVAL, "field not settable");
    }
    launch_shadows(d, c->st, c->cur);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->st));
    return WF_OK;
  });
}

int wf_get_field(void *ctx, int32_t field, double *h) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (field < 0 || field >= WF_NFIELDS || !h) throw Error(WF_EINVAL, "bad field");
    if (!c->inited) {
      if (c->hf[field].empty()) throw Error(WF_ESTATE, "field not set");
      std::memcpy(h, c->hf[field].data(), c->hf[field].size() * sizeof(double));
      return WF_OK;
    }
    Dev &d = c->d;
    switch (field) {
      case WF_F_X: get_d4(c, d.xh[c->cur], h, 0, 3, 3); break;
      case WF_F_H: get_d4(c, d.xh[c->cur], h, 3, 1, 1); break;
      case WF_F_V: get_d4(c, d.vm, h, 0, 3, 3); break;
      case WF_F_MASS: get_d4(c, d.vm, h, 3, 1, 1); break;
      case WF_F_A: get_d4(c, d.acc, h, 0, 3, 3); break;
      case WF_F_DISP: get_d4(c, d.disp, h, 0, 3, 3); break;
      case WF_F_DENSITY: get_scalar(c, d.rho[c->cur], h); break;
      case WF_F_SIGMA: {
        auto rec = dget(c, d.srec, (size_t)c->npad);
        for (int64_t k = 0; k < c->n; ++k) {
          double *m = h + 9 * (size_t)c->perm[k];
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) m[3 * a + b] = rec[k].s[sym_index(a, b)];
        }
        break;
      }
      case WF_F_SHEAR: get_sym(c, d.S, h); break;
      case WF_F_STRAIN_RATE: get_sym(c, d.sr, h); break;
      case WF_F_ROT_RATE: {
        auto t = dget(c, d.rr, (size_t)3 * c->npad);
        for (int64_t k = 0; k < c->n; ++k) {
          double *m = h + 9 * (size_t)c->perm[k];
          const double w01 = t[k], w02 = t[(size_t)c->npad + k], w12 = t[(size_t)2 * c->npad + k];
          const double f[9] = {0.0, w01, w02, -w01, 0.0, w12, -w02, -w12, 0.0};
          std::memcpy(m, f, sizeof f);
        }
        break;
      }
      case WF_F_STRAIN: get_sym(c, d.strain, h); break;
      case WF_F_STRAIN_PL: get_sym(c, d.epl, h); break;
      case WF_F_STRAIN_PL_INCR: get_sym(c, d.eplinc, h); break;
      case WF_F_PL_STRAIN: get_scalar(c, d.pl, h); break;
      case WF_F_DPL_STRAIN: get_scalar(c, d.dpl, h); break;
      case WF_F_SIGMAY: get_scalar(c, d.sigy, h); break;
      case WF_F_EFF_SR: get_scalar(c, d.effsr, h); break;
      case WF_F_PRESSURE: get_scalar(c, d.pres, h); break;
      case WF_F_T: get_scalar(c, d.T, h); break;
      case WF_F_DTDT: get_scalar(c, d.dTdt, h); break;
      case WF_F_DDENSITY: get_scalar(c, d.drho, h); break;
      case WF_F_Q_PLHEAT: get_scalar(c, d.qpl, h); break;
      case WF_F_Q_CONV: get_scalar(c, d.qconv, h); break;
      case WF_F_H_CONV: get_scalar(c, d.hconv, h); break;
      case WF_F_T_INF: get_scalar(c, d.tinf, h); break;
      case WF_F_Q_SOURCE: get_scalar(c, d.qsrc, h); break;
      case WF_F_PS_ENERGY: get_scalar(c, d.psen, h); break;
      case WF_F_DAMAGE:
        if (d.damD)
          get_scalar(c, d.damD, h);
        else
          std::memset(h, 0, sizeof(double) * (size_t)c->n);
        break;
      case WF_F_ZWAB:
      case WF_F_SUMDEN:
        if (d.zwab)
          get_scalar(c, field == WF_F_ZWAB ? d.zwab : d.sumden, h);
        else
          std::memset(h, 0, sizeof(double) * (size_t)c->n);
        break;
      case WF_F_VXSPH:
        if (d.vxsph)
          get_d4(c, d.vxsph, h, 0, 3, 3);
        else
          std::memset(h, 0, sizeof(double) * 3 * (size_t)c->n);
        break;
      case WF_F_CONTFORCE:
        if (d.contact)
          get_d4(c, d.cforce, h, 0, 3, 3);
        else
          std::memset(h, 0, sizeof(double) * 3 * (size_t)c->n);
        break;
      case WF_F_GRADCORR:
        if (d.gc_on) {
          get_sym(c, d.gcm, h);
        } else {
          for (int64_t k = 0; k < c->n; ++k)
            for (int q = 0; q < 9; ++q) h[9 * (size_t)k + q] = (q % 4 == 0) ? 1.0 : 0.0;
        }
        break;
      default: throw Error(WF_EINVAL, "bad field");
    }
    return WF_OK;
  });
}

// ---------------------------------------------------------------------------
// output tables: Domain::WriteXDMF's datasets and Domain::WriteCSV's rows
// (Source/Output.cpp:75-554), packed by wf_output.hip

static void *out_staging(Ctx *c, size_t bytes) {
  if (c->out_cap < bytes) {
    if (c->outbuf) {
      HIPCHK(hipStreamSynchronize(c->st));
      HIPCHK(hipFree(c->outbuf));
      c->outbuf = nullptr;
      c->out_cap = 0;
    }
    HIPCHK(hipMalloc(&c->outbuf, bytes));
    c->out_cap = bytes;
  }
  return c->outbuf;
}

int wf_pack_xdmf(void *ctx, float *f, int32_t *iv) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (!f || !iv) throw Error(WF_EINVAL, "null output buffer");
    if (!c->inited) throw Error(WF_ESTATE, "wf_pack_xdmf before wf_init");
    const size_t nf = (size_t)WF_XDMF_NF * (size_t)c->n, ni = (size_t)WF_XDMF_NI * (size_t)c->n;
    char *buf = (char *)out_staging(c, (nf + ni) * 4);
    float *df = (float *)buf;
    int *di = (int *)(buf + nf * 4);
    launch_pack_xdmf(c->d, c->st, c->cur, df, di);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(f, df, nf * 4, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(iv, di, ni * 4, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return WF_OK;
  });
}

int wf_pack_xdmf_noalloc(void *ctx, float *f, int32_t *iv) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (!f || !iv) throw Error(WF_EINVAL, "null output buffer");
    if (!c->inited) throw Error(WF_ESTATE, "wf_pack_xdmf_noalloc before wf_init");
    const size_t nf = (size_t)WF_XDMF_NF * (size_t)c->n, ni = (size_t)WF_XDMF_NI * (size_t)c->n;
    if ((nf + ni) * 4 > kOutScratchBytes)
      throw Error(WF_EINVAL, "wf_pack_xdmf_noalloc: " + std::to_string(c->n) + " particles need " +
                                 std::to_string((nf + ni) * 4) + " bytes of staging, more than the static scratch of " +
                                 std::to_string(kOutScratchBytes) + " bytes (at most " +
                                 std::to_string(wf_pack_xdmf_noalloc_max()) + " particles); use wf_pack_xdmf");
    HIPCHK(hipSetDevice(c->cfg.device));
    void *scr = nullptr;
    HIPCHK(out_scratch(c->cfg.device, &scr));
    float *df = (float *)scr;
    int *di = (int *)((char *)scr + nf * 4);
    launch_pack_xdmf(c->d, c->st, c->cur, df, di);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(f, df, nf * 4, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(iv, di, ni * 4, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return WF_OK;
  });
}

int64_t wf_pack_xdmf_noalloc_max(void) {
  return (int64_t)(kOutScratchBytes / (4 * (size_t)(WF_XDMF_NF + WF_XDMF_NI)));
}

int wf_pack_csv(void *ctx, double *rows) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (!rows) throw Error(WF_EINVAL, "null output buffer");
    if (!c->inited) throw Error(WF_ESTATE, "wf_pack_csv before wf_init");
    const size_t nr = (size_t)WF_CSV_NCOL * (size_t)c->n;
    double *dr = (double *)out_staging(c, nr * 8);
    launch_pack_csv(c->d, c->st, c->cur, dr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(rows, dr, nr * 8, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return WF_OK;
  });
}

int wf_set_int_field(void *ctx, int32_t field, const int32_t *h) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (field < 0 || field >= WF_NIFIELDS || field == WF_I_NB || field == WF_I_CONTNB || !h)
      throw Error(WF_EINVAL, "bad int field");
    if (c->inited) throw Error(WF_ESTATE, "int fields are fixed after wf_init");
    if (field == WF_I_MATERIAL)
      for (int64_t i = 0; i < c->n; ++i)
        if (h[i] < 0 || h[i] >= kMaxMat) throw Error(WF_EINVAL, "material index out of range");
    c->hi[field].assign(h, h + c->n);
    return WF_OK;
  });
}

int wf_get_int_field(void *ctx, int32_t field, int32_t *h) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (field < 0 || field >= WF_NIFIELDS || !h) throw Error(WF_EINVAL, "bad int field");
    if (field == WF_I_CONTNB && c->inited && !c->d.contact) {
      std::memset(h, 0, (size_t)c->n * sizeof(int32_t));
      return WF_OK;
    }
    if (field == WF_I_NB || field == WF_I_CONTNB ||
        (field == WF_I_ID && c->inited && (c->d.contact || c->d.damage))) {
      // neighbour and contact-pair counts; zone ids once free-surface
      // particles took ID 1 (contact, damage)
      if (!c->inited) throw Error(WF_ESTATE, "wf_init not called");
      const int *src = field == WF_I_NB ? c->d.nnb : (field == WF_I_CONTNB ? c->d.ccnt : c->d.id);
      auto t = dget(c, src, (size_t)c->npad);
      for (int64_t k = 0; k < c->n; ++k) h[c->perm[k]] = t[k];
      return WF_OK;
    }
    std::memcpy(h, c->hi[field].data(), (size_t)c->n * sizeof(int32_t));
    return WF_OK;
  });
}

int wf_add_amplitude(void *ctx, int32_t id, int32_t npts, const double *t, const double *v) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (npts < 2 || !t || !v) throw Error(WF_EINVAL, "amplitude needs at least 2 points");
    AmpHost a;
    a.id = id;
    a.t.assign(t, t + npts);
    a.v.assign(v, v + npts);
    c->amps.push_back(a);
    c->bc_dirty = true;
    return WF_OK;
  });
}

int wf_add_velocity_bc(void *ctx, int32_t zone, int32_t value_type, const double value[3], int32_t amp_id,
                       double amp_factor) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (!value || (value_type != 0 && value_type != 1)) throw Error(WF_EINVAL, "bad velocity BC");
    if ((int)c->bcs.size() >= kMaxBc) throw Error(WF_EUNSUPPORTED, "more than 8 boundary conditions");
    BcHost b{};
    b.zone = zone;
    b.type = 0;
    b.value_type = value_type;
    b.amp_id = amp_id;
    b.amp_factor = amp_factor;
    for (int k = 0; k < 3; ++k) b.value[k] = value[k];
    c->bcs.push_back(b);
    if (std::find(c->rzones.begin(), c->rzones.end(), zone) == c->rzones.end()) {
      if ((int)c->rzones.size() >= kMaxZones) throw Error(WF_EUNSUPPORTED, "more than 8 velocity-BC zones");
      c->rzones.push_back(zone);
    }
    c->bc_dirty = true;
    return WF_OK;
  });
}

int wf_set_bc_angular_velocity(void *ctx, int32_t zone, const double w[3]) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (!w || !std::isfinite(w[0]) || !std::isfinite(w[1]) || !std::isfinite(w[2]))
      throw Error(WF_EINVAL, "bad angular velocity");
    bool found = false;
    for (BcHost &b : c->bcs)
      if (b.type == 0 && b.zone == zone) {
        for (int k = 0; k < 3; ++k) b.w[k] = w[k];
        found = true;
      }
    if (!found) throw Error(WF_EINVAL, "no velocity BC on that zone");
    c->bc_dirty = true;
    return WF_OK;
  });
}

int wf_add_temperature_bc(void *ctx, int32_t zone, double T) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if ((int)c->bcs.size() >= kMaxBc) throw Error(WF_EUNSUPPORTED, "more than 8 boundary conditions");
    BcHost b{};
    b.zone = zone;
    b.type = 2;
    b.T = T;
    c->bcs.push_back(b);
    c->bc_dirty = true;
    return WF_OK;
  });
}

int wf_init(void *ctx, double dt) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (c->inited) throw Error(WF_ESTATE, "wf_init already called");
    if (!(dt > 0)) throw Error(WF_EINVAL, "dt must be positive");
    for (int f : {WF_F_X, WF_F_DENSITY, WF_F_MASS, WF_F_H})
      if (c->hf[f].empty()) throw Error(WF_ESTATE, "x, density, mass and h must be set before wf_init");
    for (int64_t i = 0; i < c->n; ++i)
      if (!c->mat_set[c->hi[WF_I_MATERIAL][i]]) throw Error(WF_ESTATE, "particle references an unset material");
    if (c->cfg.thermal_solver)
      for (int64_t i = 0; i < c->n; ++i)
        if (!(c->mats[c->hi[WF_I_MATERIAL][i]].cp_T > 0)) throw Error(WF_EINVAL, "thermal solver needs cp_T > 0");
    if (c->cfg.grad_corr)
      for (int64_t i = 0; i < c->n; ++i)
        if (!c->hi[WF_I_FREE][i])
          throw Error(WF_EUNSUPPORTED,
                      "the kernel-gradient correction with fixed particles (FSMPairs) is not implemented");
    if (c->cfg.precision == WF_MIXED) {
      // the FP32 gather records carry m/rho in place of h (see Src<true>)
      std::vector<double> hm(kMaxMat, -1.0);
      for (int64_t i = 0; i < c->n; ++i) {
        double &r = hm[c->hi[WF_I_MATERIAL][i]];
        if (r < 0) r = c->hf[WF_F_H][i];
        else if (r != c->hf[WF_F_H][i])
          throw Error(WF_EUNSUPPORTED, "mixed precision needs one smoothing length per material");
      }
    }
    if ((c->heat_cond || c->heat_fric) && c->meshes.empty())
      throw Error(WF_EINVAL, "contact heat (wf_set_contact_heat) needs a rigid mesh (wf_add_rigid_plane / _mesh)");
    if (c->dim == 2) {
      // plane strain covers the solid passes, conduction and convection;
      // the reference's 2D contact reads another particle's mass in its
      // dS2 (Source/Contact_Wang.cpp:259) and its 2D tensile-instability
      // branch is not modelled here
      if (!c->meshes.empty()) throw Error(WF_EUNSUPPORTED, "2D: rigid contact is not supported");
      for (int m = 0; m < kMaxMat; ++m)
        if (c->mat_set[m] && c->mats[m].ti > 0.0)
          throw Error(WF_EUNSUPPORTED, "2D: the tensile-instability term is not supported");
      const std::vector<double> &x = c->hf[WF_F_X];
      for (int64_t i = 1; i < c->n; ++i)
        if (x[3 * i + 2] != x[2]) throw Error(WF_EINVAL, "2D: every particle must have the same z");
    }
    {
      bool dam = false;
      for (int64_t i = 0; i < c->n; ++i) {
        const int m = c->hi[WF_I_MATERIAL][i];
        if (!c->mat_dam[m]) continue;
        dam = true;
        if (c->mats[m].model != WF_JOHNSON_COOK)
          throw Error(WF_EINVAL, "Johnson-Cook damage needs a Johnson-Cook material (T_t, T_m, eps_0)");
      }
      if (dam && c->cfg.precision != WF_FP64) throw Error(WF_EUNSUPPORTED, "damage needs precision WF_FP64");
      if (dam && !c->cfg.locking)
        throw Error(WF_EUNSUPPORTED, "damage needs the Locking sums (wf_config.locking 1, as the scene driver sets)");
    }
    if (c->heat_cond && c->rigid_T.empty())
      throw Error(WF_EINVAL,
                  "contact conduction needs the rigid particles' temperatures (wf_set_rigid_temperature)");
    HIPCHK(hipSetDevice(c->cfg.device));
    c->dt = dt;
    cell_initiate(c);
    make_permutation(c);
    allocate(c);
    upload_all(c);
    launch_shadows(c->d, c->st, c->cur);
    if (c->d.contact) {
      // free surface before the loop (Source/SolverLeapfrog.cpp:84-100)
      refresh_bc(c);
      launch_schedule(c->d, c->st, 1);
      launch_neighbours(c->d, c->st, 0);
      launch_surface(c->d, c->st, 0, 1);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->st));
    c->steps = 0;
    c->cur = 0;
    c->inited = true;
    return WF_OK;
  });
}

int wf_step(void *ctx, int32_t scheme, int64_t nsteps) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (!c->inited) throw Error(WF_ESTATE, "wf_init not called");
    if (scheme != WF_LEAPFROG && scheme != WF_KICKDRIFT) throw Error(WF_EINVAL, "unknown scheme");
    if (scheme == WF_KICKDRIFT && c->d.damage) throw Error(WF_EUNSUPPORTED, "damage is modelled by the Leapfrog step only");
    if (c->next_part != 0) throw Error(WF_ESTATE, "wf_step inside a step run by wf_step_part");
    for (int64_t s = 0; s < nsteps; ++s) one_step(c, scheme);
    return WF_OK;
  });
}

int wf_step_part(void *ctx, int32_t scheme, int32_t part) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (!c->inited) throw Error(WF_ESTATE, "wf_init not called");
    if (scheme != WF_LEAPFROG && scheme != WF_KICKDRIFT) throw Error(WF_EINVAL, "unknown scheme");
    if (part < 0 || part >= P_PARTS) throw Error(WF_EINVAL, "part must be 0..4");
    if (part != c->next_part)
      throw Error(WF_ESTATE, "wf_step_part: expected part " + std::to_string(c->next_part) + ", got " +
                                 std::to_string(part));
    if (part == 0) refresh_bc(c);
    enqueue_step(c, scheme, c->cur, nullptr, part);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->st));
    harvest_timing(c);
    c->next_part = (part + 1) % P_PARTS;
    if (c->next_part == 0) {
      c->built = true;
      c->cur ^= 1;
      c->steps += 1;
      check_flags(c);
    }
    return WF_OK;
  });
}

int wf_prepare(void *ctx, int32_t scheme) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (!c->inited) throw Error(WF_ESTATE, "wf_init not called");
    if (scheme != WF_LEAPFROG && scheme != WF_KICKDRIFT) throw Error(WF_EINVAL, "unknown scheme");
    refresh_bc(c);
    if (!c->cfg.no_graphs)
      for (int cur = 0; cur < 2; ++cur) step_graph(c, scheme, cur);
    return WF_OK;
  });
}

int wf_sync(void *ctx) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    HIPCHK(hipStreamSynchronize(c->st));
    if (c->inited) {
      harvest_timing(c);
      check_flags(c);
    }
    return WF_OK;
  });
}

int wf_get_scalars(void *ctx, wf_scalars *out) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (!out) throw Error(WF_EINVAL, "null argument");
    if (!c->inited) throw Error(WF_ESTATE, "wf_init not called");
    launch_max_pl(c->d, c->st);
    auto s = dget(c, c->d.scalars, S_COUNT);
    auto f = dget(c, c->d.flags, 1)[0];
    std::memset(out, 0, sizeof *out);
    out->time = f.time;
    out->dt = f.dt;
    out->steps = c->steps;
    out->nb_searches = f.nb_searches;
    out->npairs = f.directed_pairs / 2;
    out->max_nb_count = f.max_count;
    out->kin_energy_sum = s[S_KIN];
    out->int_energy_sum = s[S_INT];
    out->plastic_work = s[S_PLW];
    out->max_pl_strain = s[S_MAXPL];
    out->max_T = s[S_MAXT];
    out->min_T = s[S_MINT];
    if (c->d.contact) {
      out->contact_force_sum = s[S_CFSUM];
      out->contact_reaction_sum = s[S_CREACT];
      out->ext_forces_work = s[S_EXTW];
      out->contact_friction_work = s[S_FRW];
      out->accum_cont_heat_cond = s[S_CHC];
      out->min_force_ts = f.min_force_ts;
      auto cc = dget(c, c->d.ccnt, (size_t)c->npad);
      int64_t tot = 0;
      for (int64_t k = 0; k < c->n; ++k) tot += cc[k];
      out->contact_pairs = tot;
    }
    return WF_OK;
  });
}

int wf_get_reaction_history(void *ctx, int32_t zone, double *out3, int64_t cap, int64_t *count) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (!c->inited) throw Error(WF_ESTATE, "wf_init not called");
    auto it = std::find(c->rzones.begin(), c->rzones.end(), zone);
    if (it == c->rzones.end()) throw Error(WF_EINVAL, "zone has no velocity BC");
    const int z = (int)(it - c->rzones.begin());
    const long long hc = c->d.hist_cap;
    long long avail = std::min<long long>(c->steps, hc);
    long long m = std::min<long long>(avail, cap);
    auto h = dget(c, c->d.react_hist, (size_t)hc * kMaxZones * 3);
    const long long first = c->steps - m;
    for (long long s = 0; s < m; ++s) {
      const long long slot = (first + s) % hc;
      for (int k = 0; k < 3; ++k) out3[3 * s + k] = h[(slot * kMaxZones + z) * 3 + k];
    }
    if (count) *count = m;
    return WF_OK;
  });
}

// TriMesh::AxisPlaneMesh (3D), CalcNormals, CalcSpheres and
// Domain::AddTrimeshParticles (Source/Mesh.cpp:82-243, Source/Contact.cpp:10-37)
static void plane_normals(MeshHost &M) {
#pragma clang fp contract(off)
  const size_t ne = M.el.size() / 3;
  for (size_t e = 0; e < ne; ++e) {
    const double *n0 = &M.node[3 * M.el[3 * e]], *n1 = &M.node[3 * M.el[3 * e + 1]], *n2 = &M.node[3 * M.el[3 * e + 2]];
    double u[3], v[3];
    for (int k = 0; k < 3; ++k) {
      u[k] = n1[k] - n0[k];
      v[k] = n2[k] - n0[k];
    }
    const double w[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
    const double nw = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    for (int k = 0; k < 3; ++k) M.normal[3 * e + k] = w[k] / nw;
  }
}

static MeshHost make_plane(int zone, int axis, int flip, const double *p1, const double *p2, int dens, double hfac) {
#pragma clang fp contract(off)
  MeshHost M;
  M.zone = zone;
  const int ne = 2 * dens * dens;
  double p[3];
  for (int k = 0; k < 3; ++k) p[k] = p2[k] - p1[k];
  int dir[3];
  if (axis == 0) {
    dir[0] = 1;
    dir[1] = 2;
  } else if (axis == 1) {
    dir[0] = 0;
    dir[1] = 2;
  } else {
    dir[0] = 0;
    dir[1] = 1;
  }
  dir[2] = axis;
  const double x3 = p1[dir[2]];
  double x2 = p1[dir[1]];
  const double dl = p[dir[0]] / dens;
  for (int j = 0; j < dens + 1; ++j) {
    double x1 = p1[dir[0]];
    for (int i = 0; i < dens + 1; ++i) {
      double v[3];
      v[dir[0]] = x1;
      v[dir[1]] = x2;
      v[dir[2]] = x3;
      for (int k = 0; k < 3; ++k) M.node.push_back(v[k]);
      x1 += dl;
    }
    x2 += dl;
  }
  M.centroid.resize(3 * (size_t)ne);
  for (int j = 0; j < dens; ++j)
    for (int i = 0; i < dens; ++i) {
      const int n0 = (dens + 1) * j + i, n1 = n0 + 1, n2 = (dens + 1) * (j + 1) + i, n3 = n2 + 1;
      const int con[2][3] = {{n0, flip ? n1 : n2, flip ? n2 : n1}, {n1, flip ? n3 : n2, flip ? n2 : n3}};
      for (int e = 0; e < 2; ++e) {
        const size_t el = M.el.size() / 3;
        for (int k = 0; k < 3; ++k) M.el.push_back(con[e][k]);
        const double *a = &M.node[3 * con[e][0]], *b = &M.node[3 * con[e][1]], *c = &M.node[3 * con[e][2]];
        for (int k = 0; k < 3; ++k) M.centroid[3 * el + k] = (a[k] + b[k] + c[k]) / 3.;
      }
    }
  M.normal.resize(3 * (size_t)ne);
  plane_normals(M);
  if (flip)
    for (double &v : M.normal) v = -v;
  M.radius.resize(ne);
  M.nfar.resize(ne);
  M.pplane.resize(ne);
  M.h.resize(ne);
  const double hf = (double)(float)hfac;  // AddTrimeshParticles takes hfac as float
  for (int e = 0; e < ne; ++e) {
    double mx = 0.;
    for (int q = 0; q < 3; ++q) {
      double rv[3];
      for (int k = 0; k < 3; ++k) rv[k] = M.node[3 * M.el[3 * e + q] + k] - M.centroid[3 * e + k];
      const double nr = std::sqrt(rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2]);
      if (nr > mx) mx = nr;
      M.nfar[e] = q;  // CalcSpheres leaves the last node (Source/Mesh.cpp:212)
    }
    M.radius[e] = mx;
    const double *nf = &M.node[3 * M.el[3 * e + M.nfar[e]]];
    M.pplane[e] = nf[0] * M.normal[3 * e] + nf[1] * M.normal[3 * e + 1] + nf[2] * M.normal[3 * e + 2];
    M.h[e] = hf * M.radius[e];
  }
  return M;
}

int wf_add_rigid_plane(void *ctx, int32_t zone, int32_t axis, int32_t flip, const double p1[3], const double p2[3],
                       int32_t dens, double hfac) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (c->inited) throw Error(WF_ESTATE, "rigid planes must be added before wf_init");
    if (!p1 || !p2) throw Error(WF_EINVAL, "null argument");
    if (axis < 0 || axis > 2) throw Error(WF_EINVAL, "axis must be 0, 1 or 2");
    if (dens < 1 || dens > 1024) throw Error(WF_EINVAL, "dens must be in 1..1024");
    if (!(hfac > 0)) throw Error(WF_EINVAL, "hfac must be positive");
    if ((int)c->meshes.size() >= kMaxMesh) throw Error(WF_EUNSUPPORTED, "at most 4 rigid planes");
    if (zone == 1) throw Error(WF_EINVAL, "zone 1 is the free-surface id");
    const int d0 = axis == 0 ? 1 : 0;
    if (!(std::fabs(p2[d0] - p1[d0]) > 0)) throw Error(WF_EINVAL, "plane has no extent");
    c->meshes.push_back(make_plane(zone, axis, flip, p1, p2, dens, hfac));
    return (int)c->meshes.back().h.size();
  });
}

// TriMesh from explicit nodes and triangles (TriMesh(NastranReader &, bool),
// Source/Mesh.cpp:17-64, after the caller's flip / scale / translation):
// centroids (n0 + n1 + n2) / 3, normals from the winding (CalcNormals), then
// CalcSpheres (radius, nfar = the last node, plane coefficients) and
// AddTrimeshParticles' h = (float)hfac * radius
static MeshHost make_mesh(int zone, int64_t nnode, const double *nodes, int64_t nelem, const int32_t *el,
                          const double *normals, double hfac) {
#pragma clang fp contract(off)
  MeshHost M;
  M.zone = zone;
  M.node.assign(nodes, nodes + 3 * nnode);
  M.el.assign(el, el + 3 * nelem);
  M.centroid.resize(3 * (size_t)nelem);
  for (int64_t e = 0; e < nelem; ++e) {
    const double *a = &M.node[3 * (size_t)el[3 * e]], *b = &M.node[3 * (size_t)el[3 * e + 1]],
                 *c = &M.node[3 * (size_t)el[3 * e + 2]];
    for (int k = 0; k < 3; ++k) M.centroid[3 * e + k] = (a[k] + b[k] + c[k]) / 3.;
  }
  M.normal.resize(3 * (size_t)nelem);
  if (normals)
    M.normal.assign(normals, normals + 3 * nelem);
  else
    plane_normals(M);
  M.radius.resize(nelem);
  M.nfar.resize(nelem);
  M.pplane.resize(nelem);
  M.h.resize(nelem);
  const double hf = (double)(float)hfac;
  for (int64_t e = 0; e < nelem; ++e) {
    double mx = 0.;
    for (int q = 0; q < 3; ++q) {
      double rv[3];
      for (int k = 0; k < 3; ++k) rv[k] = M.node[3 * (size_t)M.el[3 * e + q] + k] - M.centroid[3 * e + k];
      const double nr = std::sqrt(rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2]);
      if (nr > mx) mx = nr;
      M.nfar[e] = q;
    }
    M.radius[e] = mx;
    const double *nf = &M.node[3 * (size_t)M.el[3 * e + M.nfar[e]]];
    M.pplane[e] = nf[0] * M.normal[3 * e] + nf[1] * M.normal[3 * e + 1] + nf[2] * M.normal[3 * e + 2];
    M.h[e] = hf * M.radius[e];
  }
  return M;
}

int wf_add_rigid_mesh(void *ctx, int32_t zone, int64_t nnode, const double *nodes, int64_t nelem,
                      const int32_t *elems, const double *normals, double hfac) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (c->inited) throw Error(WF_ESTATE, "rigid meshes must be added before wf_init");
    if (!nodes || !elems) throw Error(WF_EINVAL, "null argument");
    if (nnode < 3 || nelem < 1 || nnode > (1 << 24) || nelem > (1 << 24))
      throw Error(WF_EINVAL, "mesh needs 3..2^24 nodes and 1..2^24 triangles");
    if (!(hfac > 0)) throw Error(WF_EINVAL, "hfac must be positive");
    if ((int)c->meshes.size() >= kMaxMesh) throw Error(WF_EUNSUPPORTED, "at most 4 rigid meshes");
    if (zone == 1) throw Error(WF_EINVAL, "zone 1 is the free-surface id");
    for (int64_t k = 0; k < 3 * nelem; ++k)
      if (elems[k] < 0 || elems[k] >= nnode) throw Error(WF_EINVAL, "triangle references a node out of range");
    for (int64_t k = 0; k < 3 * nnode; ++k)
      if (!std::isfinite(nodes[k])) throw Error(WF_EINVAL, "non-finite node coordinate");
    if (normals)
      for (int64_t k = 0; k < 3 * nelem; ++k)
        if (!std::isfinite(normals[k])) throw Error(WF_EINVAL, "non-finite element normal");
    MeshHost M = make_mesh(zone, nnode, nodes, nelem, elems, normals, hfac);
    for (int64_t e = 0; e < nelem; ++e)
      if (!(M.radius[e] > 0) || !std::isfinite(M.normal[3 * e]))
        throw Error(WF_EINVAL, "degenerate triangle " + std::to_string(e));
    c->meshes.push_back(std::move(M));
    return (int)nelem;
  });
}

int wf_set_contact(void *ctx, double friction_sta, double friction_dyn, double damping) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (!(friction_sta >= 0) || !(friction_dyn >= 0) || !(damping >= 0)) throw Error(WF_EINVAL, "negative contact parameter");
    c->fr_sta = friction_sta;
    c->fr_dyn = friction_dyn;
    c->dfac = damping;
    if (c->inited) {
      // the kernels take the parameters by value: recapture the step graphs
      c->d.fr_sta = friction_sta;
      c->d.fr_dyn = friction_dyn;
      c->d.dfac = damping;
      c->drop_graphs();
    }
    return WF_OK;
  });
}

int wf_set_contact_law(void *ctx, int32_t law, double pfac, int32_t fric_bound) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (c->inited) throw Error(WF_ESTATE, "wf_set_contact_law after wf_init");
    if (law != 0 && law != 1 && law != 2 && law != 4)
      throw Error(WF_EINVAL, "contact law: 0 Fraser, 1 Wang, 2 Seo or 4 LSDyna");
    if (!(pfac > 0) || !std::isfinite(pfac)) throw Error(WF_EINVAL, "PFAC must be positive");
    c->contact_law = law;
    c->pfac = pfac;
    c->fric_bound = fric_bound ? 1 : 0;
    return WF_OK;
  });
}

int wf_set_dimension(void *ctx, int32_t dim) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (c->inited) throw Error(WF_ESTATE, "wf_set_dimension after wf_init");
    if (dim != 2 && dim != 3) throw Error(WF_EINVAL, "dimension must be 2 (plane strain) or 3");
    c->dim = dim;
    return WF_OK;
  });
}

int wf_set_friction_function(void *ctx, int32_t linear, double m, double b) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (c->inited) throw Error(WF_ESTATE, "wf_set_friction_function after wf_init");
    if (linear && !(std::isfinite(m) && std::isfinite(b))) throw Error(WF_EINVAL, "non-finite friction coefficients");
    c->fric_linear = linear ? 1 : 0;
    c->fr_m = linear ? m : 0.0;
    c->fr_b = linear ? b : 0.0;
    return WF_OK;
  });
}

int wf_set_material_damage(void *ctx, int32_t index, int32_t model, const double D[5]) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (c->inited) throw Error(WF_ESTATE, "wf_set_material_damage after wf_init");
    if (index < 0 || index >= kMaxMat) throw Error(WF_EINVAL, "material index out of range");
    if (model != 0 && model != 1) throw Error(WF_EINVAL, "damage model: 0 none, 1 Johnson-Cook");
    if (model && !D) throw Error(WF_EINVAL, "null damage parameters");
    if (model)
      for (int q = 0; q < 5; ++q)
        if (!std::isfinite(D[q])) throw Error(WF_EINVAL, "non-finite damage parameter");
    c->mat_dam[index] = model;
    for (int q = 0; q < 5; ++q) c->mat_dpar[5 * index + q] = model ? D[q] : 0.0;
    return WF_OK;
  });
}

int wf_set_contact_heat(void *ctx, int32_t conduction, int32_t friction, double hc) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (c->inited) throw Error(WF_ESTATE, "wf_set_contact_heat after wf_init");
    if (conduction && !(hc >= 0)) throw Error(WF_EINVAL, "negative contact conductance");
    c->heat_cond = conduction ? 1 : 0;
    c->heat_fric = friction ? 1 : 0;
    c->heat_hc = conduction ? hc : 0.0;
    return WF_OK;
  });
}

int wf_set_rigid_temperature(void *ctx, int64_t count, const double *T) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (c->inited) throw Error(WF_ESTATE, "wf_set_rigid_temperature after wf_init");
    if (count < 0 || (count > 0 && !T)) throw Error(WF_EINVAL, "null argument");
    int64_t nrig = 0;
    for (const MeshHost &M : c->meshes) nrig += (int64_t)M.h.size();
    if (count != nrig) throw Error(WF_EINVAL, "rigid temperatures: one per rigid particle of the meshes added");
    c->rigid_T.assign(T, T + count);
    return WF_OK;
  });
}

int wf_get_contact_force_history(void *ctx, int32_t mesh, double *out3, int64_t cap, int64_t *count) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (!c->inited) throw Error(WF_ESTATE, "wf_init not called");
    if (mesh < 0 || mesh >= c->d.nmesh) throw Error(WF_EINVAL, "no such rigid plane");
    const long long hc = c->d.hist_cap;
    const long long m = std::min<long long>(std::min<long long>(c->steps, hc), cap);
    auto h = dget(c, c->d.cf_hist, (size_t)hc * kMaxMesh * 3);
    const long long first = c->steps - m;
    for (long long s = 0; s < m; ++s) {
      const long long slot = (first + s) % hc;
      for (int k = 0; k < 3; ++k) out3[3 * s + k] = h[(slot * kMaxMesh + mesh) * 3 + k];
    }
    if (count) *count = m;
    return WF_OK;
  });
}

int wf_adj_forward(void *ctx, const wf_adj_config *cfg, const double *target, double *loss, double *history) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    adj_check(c);
    if (!cfg || !target || !loss) throw Error(WF_EINVAL, "null argument");
    const Dev &d = c->d;
    if (cfg->steps < 1 || cfg->steps > d.hist_cap) throw Error(WF_EINVAL, "window steps must be in 1..65536");
    if (cfg->checkpoint_every < 1) throw Error(WF_EINVAL, "checkpoint_every must be >= 1");
    if (cfg->loss_component < 0 || cfg->loss_component > 2) throw Error(WF_EINVAL, "loss_component must be 0..2");
    if (cfg->material < 0 || cfg->material >= kMaxMat) throw Error(WF_EINVAL, "material index out of range");
    if (cfg->loss_kind == WF_ADJ_PLANE) {
      if (cfg->loss_index < 0 || cfg->loss_index >= d.nmesh) throw Error(WF_EINVAL, "no such rigid plane");
    } else if (cfg->loss_kind == WF_ADJ_ZONE) {
      refresh_bc(c);
      if (zone_slot(c, cfg->loss_index) < 0) throw Error(WF_EINVAL, "no velocity BC on that zone");
    } else {
      throw Error(WF_EINVAL, "unknown loss kind");
    }
    HIPCHK(hipSetDevice(c->cfg.device));
    auto &A = c->adj;
    const int k = cfg->checkpoint_every;
    const long long steps = cfg->steps;
    const long long nseg = (steps + k - 1) / k;
    adj_prepare(c, k, (size_t)nseg + 1);
    A.cfg = *cfg;
    A.ready = false;
    A.have_grad = false;
    A.step0 = c->steps;
    refresh_bc(c);
    snapshot(c, 0);
    for (long long s = 0; s < steps; ++s) {
      one_step(c, WF_LEAPFROG);
      if ((s + 1) % k == 0 || s + 1 == steps) snapshot(c, (size_t)((s + 1 + k - 1) / k));
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->st));
    check_flags(c);
    A.hist.assign((size_t)steps, 0.0);
    A.target.assign(target, target + steps);
    if (cfg->loss_kind == WF_ADJ_PLANE) {
      auto h = dget(c, d.cf_hist, (size_t)d.hist_cap * kMaxMesh * 3);
      for (long long n = 0; n < steps; ++n) {
        const long long slot = (A.step0 + n) % d.hist_cap;
        A.hist[n] = h[(size_t)slot * kMaxMesh * 3 + cfg->loss_index * 3 + cfg->loss_component];
      }
    } else {
      const int z = zone_slot(c, cfg->loss_index);
      auto h = dget(c, d.react_hist, (size_t)d.hist_cap * kMaxZones * 3);
      for (long long n = 0; n < steps; ++n) {
        const long long slot = (A.step0 + n) % d.hist_cap;
        A.hist[n] = h[(size_t)slot * kMaxZones * 3 + z * 3 + cfg->loss_component];
      }
    }
    double L = 0.0;
    for (long long n = 0; n < steps; ++n) L += (A.hist[n] - A.target[n]) * (A.hist[n] - A.target[n]);
    A.loss = L;
    *loss = L;
    if (history) std::memcpy(history, A.hist.data(), (size_t)steps * sizeof(double));
    A.ready = true;
    A.window = true;
    return WF_OK;
  });
}

int wf_adj_backward(void *ctx, double *grad) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    adj_check(c);
    if (!grad) throw Error(WF_EINVAL, "null argument");
    auto &A = c->adj;
    if (!A.ready) throw Error(WF_ESTATE, "wf_adj_forward has not run (or the state was restored since)");
    HIPCHK(hipSetDevice(c->cfg.device));
    Dev &d = c->d;
    const int k = A.cfg.checkpoint_every;
    const long long steps = A.cfg.steps;
    const long long nseg = (steps + k - 1) / k;
    adj_prepare(c, k, (size_t)nseg + 1);
    const size_t np = (size_t)c->npad;
    AdjBufs &b = A.buf;
    HIPCHK(hipMemsetAsync(b.xb, 0, np * sizeof(double4), c->st));
    HIPCHK(hipMemsetAsync(b.vb, 0, np * sizeof(double4), c->st));
    HIPCHK(hipMemsetAsync(b.rhob, 0, np * sizeof(double), c->st));
    HIPCHK(hipMemsetAsync(b.Sb, 0, 6 * np * sizeof(double), c->st));
    HIPCHK(hipMemsetAsync(b.syb, 0, np * sizeof(double), c->st));
    HIPCHK(hipMemsetAsync(b.plb, 0, np * sizeof(double), c->st));
    HIPCHK(hipMemsetAsync(b.thb, 0, (size_t)kAdjParams * np * sizeof(double), c->st));
    refresh_bc(c);
    const int zslot = A.cfg.loss_kind == WF_ADJ_ZONE ? zone_slot(c, A.cfg.loss_index) : -1;
    for (long long sg = nseg - 1; sg >= 0; --sg) {
      restore(c, (size_t)sg);
      const long long w0 = sg * k;
      const int len = (int)std::min<long long>(k, steps - w0);
      HIPCHK(hipMemsetAsync(A.ver, 0, (size_t)k * sizeof(int), c->st));
      // replay the segment with the tape
      for (int s = 0; s < len; ++s) {
        TapeHook hk{A.slot[s], s, s == 0 ? 1 : 0, A.ver};
        enqueue_step(c, WF_LEAPFROG, c->cur, &hk);
        HIPCHK(hipGetLastError());
        c->cur ^= 1;
        c->steps += 1;
      }
      tape_state(c, A.slot[len], c->cur);
      const auto ver = dget(c, A.ver, (size_t)len);
      // reverse sweep over the segment
      for (int s = len - 1; s >= 0; --s) {
        AdjSlot s0 = A.slot[s];
        int v = s;
        while (v > 0 && !ver[v]) --v;
        s0.nbr = A.slot[v].nbr;
        s0.nnb = A.slot[v].nnb;
        s0.id = A.slot[v].id;
        s0.cidx = A.slot[v].cidx;
        s0.ccnt = A.slot[v].ccnt;
        const long long n = w0 + s;
        AdjSeeds seeds{};
        const double g = 2.0 * (A.hist[n] - A.target[n]);
        if (A.cfg.loss_kind == WF_ADJ_PLANE) seeds.cf[A.cfg.loss_index][A.cfg.loss_component] = g;
        else seeds.react[zslot][A.cfg.loss_component] = g;
        const double kick = (A.step0 + n == 0) ? 0.5 * c->dt : c->dt;
        launch_adj_step(d, c->st, s0, A.slot[s + 1], b, c->bc, seeds, c->dt, kick, A.cfg.material);
      }
      HIPCHK(hipGetLastError());
    }
    launch_adj_reduce(d, c->st, b, A.grad);
    auto gh = dget(c, A.grad, (size_t)kAdjParams);
    for (int p = 0; p < kAdjParams; ++p) grad[p] = gh[p];
    restore(c, (size_t)nseg);
    HIPCHK(hipStreamSynchronize(c->st));
    check_flags(c);
    A.have_grad = true;
    return WF_OK;
  });
}

int wf_adj_get_field(void *ctx, int32_t field, double *h) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (!h) throw Error(WF_EINVAL, "null argument");
    auto &A = c->adj;
    if (!A.have_grad) throw Error(WF_ESTATE, "wf_adj_backward has not run");
    switch (field) {
      case WF_F_X: get_d4(c, A.buf.xb, h, 0, 3, 3); break;
      case WF_F_V: get_d4(c, A.buf.vb, h, 0, 3, 3); break;
      case WF_F_DENSITY: get_scalar(c, A.buf.rhob, h); break;
      case WF_F_SHEAR: get_sym(c, A.buf.Sb, h); break;
      case WF_F_SIGMAY: get_scalar(c, A.buf.syb, h); break;
      case WF_F_PL_STRAIN: get_scalar(c, A.buf.plb, h); break;
      default: throw Error(WF_EINVAL, "no adjoint for that field");
    }
    return WF_OK;
  });
}

int wf_adj_restore(void *ctx) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    auto &A = c->adj;
    if (A.ckpt.empty() || !A.window) throw Error(WF_ESTATE, "no adjoint window to restore");
    HIPCHK(hipSetDevice(c->cfg.device));
    restore(c, 0);
    HIPCHK(hipStreamSynchronize(c->st));
    A.ready = false;
    A.have_grad = false;
    return WF_OK;
  });
}

int wf_get_rigid_field(void *ctx, int32_t field, double *h) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (!c->inited) throw Error(WF_ESTATE, "wf_init not called");
    if (!h) throw Error(WF_EINVAL, "null argument");
    const Dev &d = c->d;
    const double4 *src = field == WF_R_X ? d.rx : field == WF_R_V ? d.rv : field == WF_R_NORMAL ? d.enorm : nullptr;
    if (!src) throw Error(WF_EINVAL, "bad rigid field");
    if (d.nrig == 0) return WF_OK;
    auto t = dget(c, src, (size_t)d.nrig);
    for (int r = 0; r < d.nrig; ++r) {
      h[3 * r] = t[r].x;
      h[3 * r + 1] = t[r].y;
      h[3 * r + 2] = t[r].z;
    }
    return WF_OK;
  });
}

int wf_rebuild_neighbours(void *ctx) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (!c->inited) throw Error(WF_ESTATE, "wf_init not called");
    launch_schedule(c->d, c->st, 1);
    launch_neighbours(c->d, c->st, c->cur);
    if (c->d.gc_on) launch_gradcorr(c->d, c->st, c->cur);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->st));
    c->built = true;
    check_flags(c);
    return WF_OK;
  });
}

int wf_get_neighbour_pairs(void *ctx, int32_t *pairs2, int64_t cap, int64_t *count) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (!c->inited || !c->built) throw Error(WF_ESTATE, "no neighbour list built yet");
    auto nnb = dget(c, c->d.nnb, (size_t)c->npad);
    auto nbr = dget(c, c->d.nbr, (size_t)nb_rows(c->cfg.max_nb) * c->npad);
    std::vector<std::pair<int32_t, int32_t>> p;
    p.reserve((size_t)c->n * 32);
    for (int64_t i = 0; i < c->n; ++i)
      for (int k = 0; k < nnb[i]; ++k) {
        const int j = nbr[nb_slot(c->npad, k, i)];
        const int oi = c->perm[i], oj = c->perm[j];
        if (oi < oj) p.push_back({oi, oj});
      }
    std::sort(p.begin(), p.end());
    const int64_t m = std::min<int64_t>((int64_t)p.size(), cap);
    for (int64_t k = 0; k < m; ++k) {
      pairs2[2 * k] = p[k].first;
      pairs2[2 * k + 1] = p[k].second;
    }
    if (count) *count = (int64_t)p.size();
    return WF_OK;
  });
}

int wf_eval(void *ctx, int32_t what, double dt) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (!c->inited) throw Error(WF_ESTATE, "wf_init not called");
    if (!c->built && what != WF_EVAL_STRESS) throw Error(WF_ESTATE, "call wf_rebuild_neighbours first");
    switch (what) {
      case WF_EVAL_ACCEL:
        if (c->d.ti_on) launch_tir(c->d, c->st, c->cur);
        launch_accel(c->d, c->st, c->cur);
        break;
      case WF_EVAL_DENSITY: launch_dens_drift(c->d, c->st, c->cur, dt, 0); break;
      case WF_EVAL_RATES: launch_rates(c->d, c->st, c->cur); break;
      case WF_EVAL_STRESS: launch_stress(c->d, c->st, c->cur, dt); break;
      default: throw Error(WF_EINVAL, "unknown evaluation");
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->st));
    return WF_OK;
  });
}

int wf_calc_force2233(void *ctx, double xsph, int32_t shepard) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (!c->inited) throw Error(WF_ESTATE, "wf_init not called");
    if (!c->built) throw Error(WF_ESTATE, "call wf_rebuild_neighbours first");
    if (!std::isfinite(xsph)) throw Error(WF_EINVAL, "non-finite XSPH coefficient");
    if (c->d.ti_on) throw Error(WF_EUNSUPPORTED, "CalcForce2233 with the tensile-instability term");
    Dev &d = c->d;
    if (!d.zwab) {
      // the step graphs hold their own copy of the context and never read these
      d.zwab = c->alloc<double>((size_t)c->npad);
      d.sumden = c->alloc<double>((size_t)c->npad);
      d.vxsph = c->alloc<double4>((size_t)c->npad);
    }
    launch_force2233(d, c->st, c->cur, xsph, shepard ? 1 : 0);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->st));
    return WF_OK;
  });
}

int wf_set_timing(void *ctx, int32_t enable) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    harvest_timing(c);
    c->timing = enable != 0;
    for (int k = 0; k < T_COUNT; ++k) {
      c->tsum[k] = 0.0;
      c->tcnt[k] = 0;
    }
    return WF_OK;
  });
}

int wf_get_kernel_time(void *ctx, int32_t kernel, double *avg_ms, int64_t *launches) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (kernel < 0 || kernel >= T_COUNT) throw Error(WF_EINVAL, "bad kernel id");
    harvest_timing(c);
    if (avg_ms) *avg_ms = c->tcnt[kernel] ? c->tsum[kernel] / c->tcnt[kernel] : 0.0;
    if (launches) *launches = c->tcnt[kernel];
    return WF_OK;
  });
}

int wf_get_neighbour_stats(void *ctx, int64_t *directed_pairs, int32_t *max_count) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (!c->inited) throw Error(WF_ESTATE, "wf_init not called");
    auto f = dget(c, c->d.flags, 1)[0];
    if (directed_pairs) *directed_pairs = f.directed_pairs;
    if (max_count) *max_count = f.max_count;
    return WF_OK;
  });
}

int wf_get_stage_stats(void *ctx, int64_t *tiles, int64_t *staged_tiles, int64_t *staged_values) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (!c->inited) throw Error(WF_ESTATE, "wf_init not called");
    int64_t t = c->d.ntiles, st = 0, sv = 0;
    if (c->d.mixed && c->d.stage_cnt) {
      const auto cnt = dget(c, c->d.stage_cnt, (size_t)c->d.ntiles);
      for (int v : cnt)
        if (v >= 0) {
          ++st;
          sv += v;
        }
    }
    if (tiles) *tiles = t;
    if (staged_tiles) *staged_tiles = st;
    if (staged_values) *staged_values = sv;
    return WF_OK;
  });
}

int wf_set_material_friction(void *ctx, int32_t index, double friction_sta, double friction_dyn) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (index < 0 || index >= kMaxMat) throw Error(WF_EINVAL, "material index out of range");
    if (!(friction_sta >= 0) || !(friction_dyn >= 0)) throw Error(WF_EINVAL, "negative friction coefficient");
    c->mat_fr[2 * index] = friction_sta;
    c->mat_fr[2 * index + 1] = friction_dyn;
    if (c->inited) {
      // the friction sits in the device material, read by the contact kernel
      const size_t off = (size_t)index * sizeof(DevMat) + offsetof(DevMat, fr_sta);
      const double v[2] = {friction_sta, friction_dyn};
      static_assert(offsetof(DevMat, fr_dyn) == offsetof(DevMat, fr_sta) + sizeof(double), "friction pair");
      HIPCHK(hipMemcpyAsync((char *)c->dmats + off, v, sizeof v, hipMemcpyHostToDevice, c->st));
      HIPCHK(hipStreamSynchronize(c->st));
    }
    return WF_OK;
  });
}

int wf_set_group_sums(void *ctx, int32_t zone, int64_t cap) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (c->inited) throw Error(WF_ESTATE, "wf_set_group_sums must precede wf_init");
    if (zone < -1) throw Error(WF_EINVAL, "zone must be >= 0, or -1 for the contact forces only");
    if (cap < 1 || cap > (1 << 20)) throw Error(WF_EINVAL, "cap must be 1..2^20 steps");
    c->gsum_zone = zone;
    c->gsum_cap = cap;
    return WF_OK;
  });
}

int wf_get_group_history(void *ctx, int32_t what, int32_t group, int32_t mesh, double *out3, int64_t cap,
                         int64_t *count) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (!c->inited) throw Error(WF_ESTATE, "wf_init not called");
    const Dev &d = c->d;
    if (d.ngroups == 0) throw Error(WF_ESTATE, "group sums are off (wf_set_group_sums)");
    if (group < 0 || group >= d.ngroups) throw Error(WF_EINVAL, "group out of range");
    if (what == 0 && d.gzone < 0) throw Error(WF_EINVAL, "no zone was given for the group reactions");
    if (what == 1 && (!d.contact || mesh < 0 || mesh >= d.nmesh)) throw Error(WF_EINVAL, "no such rigid mesh");
    if (what != 0 && what != 1) throw Error(WF_EINVAL, "what must be 0 (reaction) or 1 (contact force)");
    if (!out3 && cap > 0) throw Error(WF_EINVAL, "null output buffer");
    const long long hc = d.ghist_cap;
    const long long m = std::min<long long>(std::min<long long>(c->steps, hc), cap);
    const size_t per = what == 0 ? 3 : (size_t)kMaxMesh * 3;
    auto h = dget(c, what == 0 ? d.gr_hist : d.gc_hist, (size_t)hc * d.ngroups * per);
    const long long first = c->steps - m;
    for (long long s = 0; s < m; ++s) {
      const long long slot = (first + s) % hc;
      const size_t base = ((size_t)slot * d.ngroups + group) * per + (what == 0 ? 0 : (size_t)mesh * 3);
      for (int k = 0; k < 3; ++k) out3[3 * s + k] = h[base + k];
    }
    if (count) *count = m;
    return WF_OK;
  });
}

int wf_get_cell_grids(void *ctx, int32_t reference[3], int32_t search[3]) {
  Ctx *c = (Ctx *)ctx;
  return guard(c, [&] {
    as_ctx(ctx);
    if (!c->inited) throw Error(WF_ESTATE, "wf_init not called");
    for (int k = 0; k < 3; ++k) {
      if (reference) reference[k] = c->d.cellno[k];
      if (search) search[k] = c->d.scellno[k];
    }
    return WF_OK;
  });
}

// AddCylinderLength lattice, 3D without ghost rows (Source/Domain.cpp:674-860)
static int half_part_count(double r, double R, int xinc) {
  int ypartcount = -1;
  if (xinc > 0) {
    ypartcount = 1;
    double yp = r;
    double xp = r + (double)(xinc - 1) * 2. * r;
    double rad = std::sqrt(yp * yp + xp * xp);
    while (rad <= R - r) {
      yp += 2. * r;
      rad = std::sqrt(yp * yp + xp * xp);
      ypartcount++;
    }
    ypartcount -= 1;
  }
  return ypartcount;
}

int64_t wf_gen_cylinder(const double V[3], double Rxy, double Lz, double r, double *out, int64_t cap) {
#pragma clang fp contract(off)
  if (!V || !(r > 0) || !(Rxy > 0) || !(Lz > 0)) return WF_EINVAL;
  int64_t cnt = 0;
  const int numpartxy = half_part_count(r, Rxy, 1);
  double zp = V[2] + r;
  while (zp <= (V[2] + Lz - r)) {
    double yp = V[1] - r - (2. * r * (numpartxy - 1));
    const int numypart = 2 * numpartxy;
    int yinc = numpartxy, yinc_sign = -1;
    for (int j = 0; j < numypart; j++) {
      const int numxpart = half_part_count(r, Rxy, yinc);
      double xp = V[0] - r - (2. * r * (numxpart - 1));
      for (int i = 0; i < 2 * numxpart; i++) {
        if (out && cnt < cap) {
          out[3 * cnt] = xp;
          out[3 * cnt + 1] = yp;
          out[3 * cnt + 2] = zp;
        }
        cnt++;
        xp += 2. * r;
      }
      yp += 2. * r;
      yinc += yinc_sign;
      if (yinc < 1) {
        yinc = 1;
        yinc_sign = 1;
      }
    }
    zp += 2. * r;
  }
  return cnt;
}

}  // extern "C"
