// wf_kernels.hip -- HIP kernels of the explicit SPH step for gfx950.
//
// Pair passes are gathers: one lane per particle walks its own neighbour
// row and accumulates in registers, so sums need no atomics and are
// deterministic for a fixed neighbour list.  Every directed pair is
// evaluated from both ends; the pair terms of the reference are odd or even
// under the swap (i,j) -> (j,i), so each end recovers the reference's
// contribution exactly (Source/InteractionAlt.cpp:12-640).
//
// Global reductions (reaction, energies, plastic work) are two-stage:
// fixed-order block partials, then one single-block finalize kernel.
#include <float.h>
#include <math.h>

#include "wf_device.h"

namespace wf {

// What a gather reads of particle j: the FP64 state, or its FP32 shadows.
// FP64:  x = (x, y, z, h), v = (vx, vy, vz, m), m/rho and rho separate.
// FP32:  x = (x, y, z, m/rho), v = (vx, vy, vz, m), rho = m / (m/rho); the
//        smoothing length is uniform within a material (checked at set-up),
//        so the pair h is the particle's own and needs no gathered load.
template <bool MX>
struct Src;
template <>
struct Src<false> {
  using R = double;
  using R4 = double4;
  __device__ static R4 x(const Dev &d, int b, int j) { return d.xh[b][j]; }
  __device__ static R4 v(const Dev &d, int j) { return d.vm[j]; }
  __device__ static R h_own(const Dev &d, int b, int i, const R4 &xi) { return xi.w; }
  __device__ static KH<R> kh(const KH<R> &ki, const R4 &xi, const R4 &xj) { return pair_h_t<R>(ki, xi.w, xj.w); }
  __device__ static R vol(const Dev &d, int b, int j, const R4 &xj) { return d.vol[b][j]; }
  __device__ static R rho(const Dev &d, int b, int j, const R4 &xj, const R4 &vj) { return d.rho[b][j]; }
  // t = sigma_j / rho_j^2 formed as the reference forms it per pair
  __device__ static void srec(const Dev &d, int j, R t[6], R &rho) {
    const StressRec r = d.srec[j];
    for (int k = 0; k < 6; ++k) t[k] = r.s[k] * r.irho2;
    rho = r.rho;
  }
};
template <>
struct Src<true> {
  using R = float;
  using R4 = float4;
  __device__ static R4 x(const Dev &d, int b, int j) { return d.xhf[b][j]; }
  __device__ static R4 v(const Dev &d, int j) { return d.vmf[j]; }
  __device__ static R h_own(const Dev &d, int b, int i, const R4 &xi) { return (float)d.xh[b][i].w; }
  __device__ static KH<R> kh(const KH<R> &ki, const R4 &xi, const R4 &xj) { return ki; }
  __device__ static R vol(const Dev &d, int b, int j, const R4 &xj) { return xj.w; }
  __device__ static R rho(const Dev &d, int b, int j, const R4 &xj, const R4 &vj) { return vj.w * rcp_t(xj.w); }
  __device__ static void srec(const Dev &d, int j, R t[6], R &rho) {
    const StressRecF r = d.srecf[j];
    for (int k = 0; k < 6; ++k) t[k] = r.t[k];
    rho = r.rho;
  }
};

// CheckRadius (Source/Neighbour.cpp:361-372), evaluated without contraction
// so it rounds exactly as the host predicate does.
__device__ inline bool in_radius(const double4 a, const double4 b, double cellfac) {
#pragma clang fp contract(off)
  const double h = (a.w + b.w) / 2;
  const double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  const double r = sqrt(dx * dx + dy * dy + dz * dz);
  return (r / h) <= cellfac;
}

// ListGenerate's cell coordinate (Source/Neighbour.cpp:185-225)
__device__ inline int cell_coord(double x, double blpf, double cs, int cn) {
  double f = floor((x - blpf) / cs);
  if (!(f >= 0.0)) return 0;  // also catches NaN
  if (f >= (double)cn) return cn - 1;
  return (int)f;
}

// ---------------------------------------------------------------------------
// neighbour schedule (Source/SolverLeapfrog.cpp:160-201, 347-356)

WF_KERNEL void k_schedule(Dev d, int force) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  DevFlags *f = d.flags;
  const int isfirst = f->step == 0;
  const int mx = f->yielded;
  int search = 0, periodic = 0;
  if (force) {
    search = periodic = 1;
  } else {
    // with damage: no search when yielding starts, and ts_i = 1 after the
    // first step, so the first step's search is the last
    // (Source/SolverLeapfrog.cpp:160-171)
    if (d.damage && !isfirst) f->ts_i = 1;
    if (!d.damage && mx && !f->isyielding) {
      search = 1;
      f->isyielding = 1;
      f->cleared = 0;
    } else if ((mx || isfirst) && f->ts_i == 0 && f->cleared) {
      search = periodic = 1;
      f->cleared = 0;
    }
  }
  f->step_yield = mx;
  f->do_search = search;
  // CalcGradCorrMatrix follows the periodic search and runs on the first
  // step, not after the first-yield search (Source/SolverLeapfrog.cpp:160-212)
  f->gc_due = d.gc_on && (periodic || isfirst);
  if (search) {
    f->nb_searches += 1;
    f->max_count = 0;
    f->directed_pairs = 0;
  }
}

// ---------------------------------------------------------------------------
// neighbour build: counting sort of the particles (and of the rigid contact
// particles) by reference cell, then one lane per particle scans the 27
// surrounding cells with the exact predicate.  A search step runs
// k_cell_hist, k_scan, k_scatter, k_cell_sort, k_build_nbr and, with LDS
// staging, k_stage; on other steps each of them returns at once.

__device__ inline int cell_of(const Dev &d, const double4 x) {
  const int cx = cell_coord(x.x, d.blpf[0], d.cellsize[0], d.cellno[0]);
  const int cy = cell_coord(x.y, d.blpf[1], d.cellsize[1], d.cellno[1]);
  const int cz = cell_coord(x.z, d.blpf[2], d.cellsize[2], d.cellno[2]);
  return (cx * d.cellno[1] + cy) * d.cellno[2] + cz;
}
// the particle's cell in the SPH search grid
__device__ inline int scell_of(const Dev &d, const double4 x) {
  const int cx = cell_coord(x.x, d.blpf[0], d.scellsize[0], d.scellno[0]);
  const int cy = cell_coord(x.y, d.blpf[1], d.scellsize[1], d.scellno[1]);
  const int cz = cell_coord(x.z, d.blpf[2], d.scellsize[2], d.scellno[2]);
  return (cx * d.scellno[1] + cy) * d.scellno[2] + cz;
}

// threads 0..n-1: the SPH particles; n..n+nrig-1: the rigid particles.
// The counts are zero on entry (allocation, then cleared by k_cell_sort
// once a search has consumed them).
// z-slice (quarter of the cell height) of a record with FP32 offset z from
// its cell's lower face: the floor of z / (cell height / 4), clamped into
// the cell (a particle clamped into a boundary cell, or one that rounds
// onto a face, counts in the nearest slice)
__device__ inline int z_slice(float z, float inv_hz) {
  const float q = z * inv_hz;
  return q < 1.f ? 0 : (q < 2.f ? 1 : (q < 3.f ? 2 : 3));
}

// The counter's old value is the particle's place among its cell's
// particles (in atomic order; k_cell_sort fixes the order).  Also records
// each SPH particle's z-slice in its search cell (offset
// rounded to FP32 exactly as k_cell_sort's records are) and, per block, the
// largest SPH h (its FP32 bits; h > 0, so the bits order as the values)
// for the candidate pruning of k_build_nbr.
WF_KERNEL void k_cell_hist(Dev d, int cur) {
  if (!d.flags->do_search) return;
  __shared__ unsigned hw[kBlock / 64];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const int lane = threadIdx.x & 63;
  unsigned hb = 0;
  int c = -1;
  if (i < d.n) {
    const double4 x = d.xh[cur][i];
    hb = __float_as_uint((float)x.w);
    d.cell[i] = cell_of(d, x);
    c = scell_of(d, x);
    d.scell[i] = c;
    const int cz = c % d.scellno[2];
    const double oz = d.blpf[2] + cz * d.scellsize[2];
    d.sslice[i] = (unsigned char)z_slice((float)(x.z - oz), 4.f / (float)d.scellsize[2]);
  }
  // one atomic per run of consecutive lanes in the same cell (the particles
  // are in Morton order, so a wavefront's lanes share few cells): the run's
  // first lane adds the run length, each lane takes the old count plus its
  // place in the run
  {
    const int cprev = __shfl_up(c, 1, 64);
    const unsigned long long heads = __ballot(lane == 0 || cprev != c);
    const unsigned long long upto = lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1);
    const int start = 63 - __clzll(heads & upto);
    const unsigned long long above = heads & ~upto;
    const int end = above ? __ffsll((long long)above) - 1 : 64;
    int base = 0;
    if (c >= 0 && lane == start) base = atomicAdd(&d.cell_count[c], end - start);
    base = __shfl(base, start, 64);
    if (c >= 0) d.srank[i] = base + (lane - start);
  }
  if (i >= d.n && i - d.n < d.nrig) {
    const int r = i - d.n;
    const int c = cell_of(d, d.rx[r]);
    d.rcell[r] = c;
    d.rrank[r] = atomicAdd(&d.rcount[c], 1);
  }
  for (int off = 32; off > 0; off >>= 1) hb = max(hb, (unsigned)__shfl_xor((int)hb, off, 64));
  if ((threadIdx.x & 63) == 0) hw[threadIdx.x >> 6] = hb;
  __syncthreads();
  if (threadIdx.x == 0 && blockIdx.x < d.nblocks) {
    unsigned m = hw[0];
    for (int w = 1; w < kBlock / 64; ++w) m = max(m, hw[w]);
    d.hblk[blockIdx.x] = m;
  }
}

// inclusive scan of one value per lane across a wavefront
__device__ inline int wave_scan_incl(int v) {
  const int lane = threadIdx.x & 63;
  for (int off = 1; off < 64; off <<= 1) {
    int t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  return v;
}

// block-wide exclusive scan (blockDim.x <= 1024); returns the block total
// in *tot
__device__ inline int block_scan_excl(int v, int *sh, int *tot) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int inc = wave_scan_incl(v);
  if (lane == 63) sh[w] = inc;
  __syncthreads();
  if (w == 0) {
    int s = lane < (int)(blockDim.x >> 6) ? sh[lane] : 0;
    int si = wave_scan_incl(s);
    if (lane < (int)(blockDim.x >> 6)) sh[lane] = si - s;
    if (lane == (int)(blockDim.x >> 6) - 1) sh[32] = si;
  }
  __syncthreads();
  int r = inc - v + sh[w];
  *tot = sh[32];
  __syncthreads();
  return r;
}

// Cell starts of the counting sort in two levels: each block scans one
// chunk of kScanChunk cells (four per thread, one block scan) into in-chunk
// offsets (loc) and the chunk's total (tot); the block that finishes last (a
// ticket counter, reset by it) scans the chunk totals into chunk offsets
// (off) and the grand total (start[ncells]).  A cell starts at loc[c] +
// off[c / kScanChunk]; k_cell_sort writes that into start[c].
constexpr int kScanThreads = 256;
constexpr int kScanPer = 4;
static_assert(kScanChunk == kScanThreads * kScanPer, "a chunk is one block's four cells per thread");
__host__ __device__ inline int scan_chunks(long long ncells) { return (int)((ncells + kScanChunk - 1) / kScanChunk); }

__device__ inline void scan_chunk(const int *cnt, int *loc, int *tot, int ncells, int b, int *sh) {
  const int c0 = b * kScanChunk + threadIdx.x * kScanPer;
  int v[kScanPer], sum = 0;
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    v[k] = c0 + k < ncells ? cnt[c0 + k] : 0;
    sum += v[k];
  }
  int t;
  int off = block_scan_excl(sum, sh, &t);
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    if (c0 + k < ncells) loc[c0 + k] = off;
    off += v[k];
  }
  if (threadIdx.x == 0) tot[b] = t;
}

// exclusive scan of the chunk totals written by the other blocks of the
// launch (read past the L1 cache) into off[0..nch), the sum into *end
__device__ inline void scan_totals(const int *tot, int *off, int nch, int *end, int *sh) {
  int carry = 0;
  for (int base = 0; base < nch; base += kScanThreads) {
    const int q = base + threadIdx.x;
    const int v = q < nch ? __hip_atomic_load(&tot[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
    int t;
    const int e = block_scan_excl(v, sh, &t);
    if (q < nch) off[q] = carry + e;
    carry += t;
  }
  if (threadIdx.x == 0) *end = carry;
}

// blocks [0, ns): chunks of the search grid's counts; [ns, ns + nr): chunks
// of the rigid particles' counts (reference grid).  Each block also reduces
// a share of k_cell_hist's per-block largest h; the last block combines the
// shares into the search's largest h.
__global__ __launch_bounds__(kScanThreads) void k_scan(Dev d) {
  if (!d.flags->do_search) return;
  __shared__ int sh[34];
  __shared__ int last;
  const int ns = scan_chunks(d.sncells), nr = d.nrig ? scan_chunks(d.ncells) : 0, nb = ns + nr;
  const int b = blockIdx.x;
  if (b < ns) scan_chunk(d.cell_count, d.cell_loc, d.ctot, d.sncells, b, sh);
  else scan_chunk(d.rcount, d.rloc, d.rtot, d.ncells, b - ns, sh);
  const int per = (d.nblocks + nb - 1) / nb;
  unsigned hb = 0;
  for (int q = b * per + threadIdx.x; q < min(d.nblocks, (b + 1) * per); q += kScanThreads) hb = max(hb, d.hblk[q]);
  for (int o = 32; o > 0; o >>= 1) hb = max(hb, (unsigned)__shfl_xor((int)hb, o, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = (int)hb;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned m = 0;
    for (int w = 0; w < kScanThreads / 64; ++w) m = max(m, (unsigned)sh[w]);
    d.hred[b] = m;
    __threadfence();
    last = atomicAdd(&d.flags->scan_ticket, 1u) == (unsigned)(nb - 1);
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  scan_totals(d.ctot, d.coff, ns, d.cell_start + d.sncells, sh);
  if (nr) scan_totals(d.rtot, d.roff, nr, d.rstart + d.ncells, sh);
  hb = 0;
  for (int q = threadIdx.x; q < nb; q += kScanThreads)
    hb = max(hb, __hip_atomic_load(&d.hred[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  for (int o = 32; o > 0; o >>= 1) hb = max(hb, (unsigned)__shfl_xor((int)hb, o, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = (int)hb;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned m = 0;
    for (int w = 0; w < kScanThreads / 64; ++w) m = max(m, (unsigned)sh[w]);
    d.flags->hmax_bits = m;
    d.flags->scan_ticket = 0;
  }
}

// each particle to its place: its cell's start plus its in-cell rank
WF_KERNEL void k_scatter(Dev d) {
  if (!d.flags->do_search) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < d.n) {
    const int c = d.scell[i];
    d.sorted[d.cell_loc[c] + d.coff[c / kScanChunk] + d.srank[i]] = i;
  } else if (i - d.n < d.nrig) {
    const int r = i - d.n, c = d.rcell[r];
    d.rsorted[d.rloc[c] + d.roff[c / kScanChunk] + d.rrank[r]] = r;
  }
}

// where search cell c's records start (its in-chunk offset plus its chunk's)
// and end
__device__ inline int cell_first(const Dev &d, int c) { return d.cell_loc[c] + d.coff[c / kScanChunk]; }
__device__ inline int cell_end(const Dev &d, int c) {
  return c + 1 < d.sncells ? cell_first(d, c + 1) : d.cell_start[d.sncells];
}
// lower corner of search cell c
__device__ inline double3 cell_corner(const Dev &d, int c) {
  const int nz = d.scellno[2], ny = d.scellno[1];
  const int cz = c % nz, cy = (c / nz) % ny, cx = c / (nz * ny);
  return make_double3(d.blpf[0] + cx * d.scellsize[0], d.blpf[1] + cy * d.scellsize[1],
                      d.blpf[2] + cz * d.scellsize[2]);
}
// the candidate record of particle j at position p, offsets from corner o
__device__ inline void put_record(const Dev &d, int cur, int p, int j, const double3 &o) {
  const double4 x = d.xh[cur][j];
  d.sxyz[p] = make_float4((float)(x.x - o.x), (float)(x.y - o.y), (float)(x.z - o.z), (float)x.w);
  d.sinfo[p] = make_int2(j, d.group[j]);
}

// cell c (or none, c < 0) of 2-16 particles by the 16-lane segment of the
// wavefront that lane belongs to: each lane holds one particle and counts
// the segment's smaller keys; every lane of the wavefront calls this
__device__ inline void sort_small_cell(const Dev &d, int cur, int c, int lane) {
  const int seg = lane >> 4, sl = lane & 15;
  int s = 0, e = 0;
  if (c >= 0) {
    s = cell_first(d, c);
    e = cell_end(d, c);
  }
  const bool act = sl < e - s;
  int j = 0, zs = 0;
  if (act) {
    j = d.sorted[s + sl];
    zs = d.sslice[j];
  }
  const long long key = act ? ((long long)zs << 32) | (long long)j : 0x7fffffffffffffffLL;
  int rank = 0;
#pragma unroll
  for (int q = 0; q < 16; ++q) rank += __shfl(key, (seg << 4) + q, 64) < key ? 1 : 0;
  const int sh = seg << 4;
  const unsigned b0 = (unsigned)((__ballot(act && zs == 0) >> sh) & 0xffffu);
  const unsigned b1 = (unsigned)((__ballot(act && zs <= 1) >> sh) & 0xffffu);
  const unsigned b2 = (unsigned)((__ballot(act && zs <= 2) >> sh) & 0xffffu);
  if (act) {
    d.sorted[s + rank] = j;
    put_record(d, cur, s + rank, j, cell_corner(d, c));
  }
  if (c >= 0 && sl == 0) d.soff[c] = make_int4(s + __popc(b0), s + __popc(b1), s + __popc(b2), e);
}

// one cell of two or more particles, by the 64 lanes of a wavefront (lane =
// its lane)
__device__ inline void sort_cell(const Dev &d, int cur, int c, int lane) {
  const int s = cell_first(d, c), e = cell_end(d, c);
  const int cnt = e - s;
  const double3 o = cell_corner(d, c);
  if (cnt <= 64) {
    int j = 0, sl = 0;
    if (lane < cnt) {
      j = d.sorted[s + lane];
      sl = d.sslice[j];
    }
    const long long key = lane < cnt ? ((long long)sl << 32) | (long long)j : 0x7fffffffffffffffLL;
    int rank = 0;
    for (int l = 0; l < cnt; ++l) rank += __shfl(key, l, 64) < key ? 1 : 0;
    const unsigned long long b0 = __ballot(lane < cnt && sl == 0);
    const unsigned long long b1 = __ballot(lane < cnt && sl <= 1);
    const unsigned long long b2 = __ballot(lane < cnt && sl <= 2);
    if (lane < cnt) {
      d.sorted[s + rank] = j;
      put_record(d, cur, s + rank, j, o);
    }
    if (lane == 0)
      d.soff[c] = make_int4(s + __popcll(b0), s + __popcll(b1), s + __popcll(b2), e);
    return;
  }
  if (lane != 0) return;
  // key (slice, index) of each entry, held in sinfo while sorting
  for (int p = s; p < e; ++p) {
    const int j = d.sorted[p];
    d.sinfo[p] = make_int2(j, d.sslice[j]);
  }
  for (int a = s + 1; a < e; ++a) {
    const int2 v = d.sinfo[a];
    int b = a - 1;
    for (; b >= s; --b) {
      const int2 u = d.sinfo[b];
      if (u.y < v.y || (u.y == v.y && u.x < v.x)) break;
      d.sinfo[b + 1] = u;
    }
    d.sinfo[b + 1] = v;
  }
  int st[3] = {e, e, e};
  for (int p = e - 1; p >= s; --p) {
    const int2 k = d.sinfo[p];
    if (k.y >= 1) st[0] = p;
    if (k.y >= 2) st[1] = p;
    if (k.y >= 3) st[2] = p;
    d.sorted[p] = k.x;
    put_record(d, cur, p, k.x, o);
  }
  d.soff[c] = make_int4(st[0], st[1], st[2], e);
}

// per cell: write the cell's start, clear the counters for the next search,
// order the cell's particles by z-slice, then by index (cells hold ~10-30
// particles; the order, and with it every neighbour row, becomes
// independent of the atomic order), record where slices 1-3 start, and write
// the candidate records (FP32 offsets from the cell's lower corner, h,
// index, group) in cell order.  A wavefront takes cells in groups of
// kSortGroup: one lane per cell writes the start, clears the counter and
// finishes a cell of at most one particle; then the wavefront sorts the
// group's cells of 2-16 particles four at a time (sort_small_cell) and the
// larger ones one after another (sort_cell), so grids of mostly empty
// cells cost little.  The wavefronts, and the threads writing the
// rigid grid's starts, stride over the cells (a modest grid, so a step
// without a search exits early at little cost).
constexpr int kSortGroup = 16;
WF_KERNEL void k_cell_sort(Dev d, int cur) {
  if (!d.flags->do_search) return;
  if (d.nrig)
    for (int c = blockIdx.x * kBlock + threadIdx.x; c < d.ncells; c += gridDim.x * kBlock) {
      d.rstart[c] = d.rloc[c] + d.roff[c / kScanChunk];
      d.rcount[c] = 0;
    }
  const int lane = threadIdx.x & 63;
  const int nw = gridDim.x * (kBlock / 64);
  for (int base = (blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)) * kSortGroup; base < d.sncells;
       base += nw * kSortGroup) {
    const int c = base + lane;
    const bool mine = lane < kSortGroup && c < d.sncells;
    int cnt = 0;
    if (mine) {
      const int s = cell_first(d, c), e = cell_end(d, c);
      cnt = e - s;
      d.cell_start[c] = s;
      d.cell_count[c] = 0;
      if (cnt <= 1) {
        int4 so = make_int4(s, s, s, e);
        if (cnt == 1) {
          const int j = d.sorted[s], sl = d.sslice[j];
          put_record(d, cur, s, j, cell_corner(d, c));
          so = make_int4(s + (sl <= 0 ? 1 : 0), s + (sl <= 1 ? 1 : 0), s + (sl <= 2 ? 1 : 0), e);
        }
        d.soff[c] = so;
      }
    }
    // cells of 2-16 particles four at a time, one per 16-lane segment;
    // larger cells by the whole wavefront
    unsigned long long small = __ballot(mine && cnt >= 2 && cnt <= 16);
    unsigned long long big = __ballot(mine && cnt > 16);
    while (small) {
      int cq[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        cq[q] = small ? base + __ffsll((long long)small) - 1 : -1;
        small &= small - 1;
      }
      const int seg = lane >> 4;
      sort_small_cell(d, cur, seg == 0 ? cq[0] : (seg == 1 ? cq[1] : (seg == 2 ? cq[2] : cq[3])), lane);
    }
    while (big) {
      const int l = __ffsll((long long)big) - 1;
      big &= big - 1;
      sort_cell(d, cur, base + l, lane);
    }
  }
}

// One lane per particle (lanes in cell order) over the 27 cells around the
// particle's own, in cell order (each cell's particles by z-slice, then by
// index), so a row lists its neighbours in a fixed order (cells of the SPH
// search grid).  The 27 cells are 9 columns of up to three consecutive cell
// ids, whose candidates are one contiguous range of the cell-ordered
// records; of it only the slices within reach of the particle are walked
// (columns out of reach are skipped), four candidates at a time, their
// loads issued together.
// The radius test runs in FP32 on offsets from the cell corners (|error| <
// 1e-6 relative at the radius) and is trusted outside a band of 1e-4
// around it; inside the band, and for a particle lying far outside the grid
// (clamped into a boundary cell, where its FP32 offset loses digits), the
// exact FP64 predicate CheckRadius decides.
WF_KERNEL void k_build_nbr(Dev d, int cur) {
  if (!d.flags->do_search) return;
  __shared__ double sh[kBlock / 64];
  // lanes take the particles in cell order: a cell's particles walk the same
  // candidate ranges in step, so their loads are shared and their loops
  // agree
  const int ps = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  const int i = ps < d.n ? d.sorted[ps] : ps;
  const long long np = d.npad;
  int4 *row = reinterpret_cast<int4 *>(d.nbr) + i;
  int4 pend = make_int4(i, i, i, i);  // the row's last four entries (the group being filled)
  int4 *wp = row;                     // where the group being filled goes
  int cnt = 0;
  if (i < d.n && d.group[i] >= 0) {
    const int g = d.group[i];
    const double4 xi = d.xh[cur][i];
    const int c = d.scell[i];
    const int nz = d.scellno[2], ny = d.scellno[1], nx = d.scellno[0];
    const int cz = c % nz, cy = (c / nz) % ny, cx = c / (nz * ny);
    const float csx = (float)d.scellsize[0], csy = (float)d.scellsize[1], csz = (float)d.scellsize[2];
    const float oxi = (float)(xi.x - (d.blpf[0] + cx * d.scellsize[0]));
    const float oyi = (float)(xi.y - (d.blpf[1] + cy * d.scellsize[1]));
    const float ozi = (float)(xi.z - (d.blpf[2] + cz * d.scellsize[2]));
    const float hi = (float)xi.w, hcf = 0.5f * (float)d.cellfac;
    const bool exact_all = fabsf(oxi) > 64.f * csx || fabsf(oyi) > 64.f * csy || fabsf(ozi) > 64.f * csz;
    const int az0 = cz > 0 ? cz - 1 : 0, az1 = cz + 1 < nz ? cz + 1 : nz - 1, m = az1 - az0;
    // every neighbour lies within ri of particle i: the pair radius
    // Cellfac (h_i + h_j) / 2 with h_j at most the largest h of the search,
    // widened by 1e-4 (FP32 rounding of these offsets is ~1e-7 of a cell)
    const float ri = hcf * (hi + __uint_as_float(d.flags->hmax_bits)) * 1.0001f, ri2 = ri * ri;
    const float inv_hz = 4.f / csz;  // the z-slices of k_cell_sort
    const int gmax = 4 * m + 3;      // last slice of the column
    // z of i from the lower faces of cells az0, az0 + 1, az0 + 2
    const float bz0 = (float)(cz - az0) * csz + ozi, bz1 = (float)(cz - az0 - 1) * csz + ozi,
                bz2 = (float)(cz - az0 - 2) * csz + ozi;
    for (int col = 0; col < 9; ++col) {
      const int dcx = col / 3 - 1, dcy = col % 3 - 1;
      const int ax = cx + dcx, ay = cy + dcy;
      if (ax < 0 || ax >= nx || ay < 0 || ay >= ny) continue;
      const float bx = (float)(cx - ax) * csx + oxi, by = (float)(cy - ay) * csy + oyi;
      const int cc0 = (ax * ny + ay) * nz + az0;
      // candidates of cells az0, az0 + 1, az0 + 2 start at s0, s1, s2
      const int s0 = d.cell_start[cc0], s1 = d.cell_start[cc0 + 1];
      const int s2 = m >= 1 ? d.cell_start[cc0 + 2] : s1;
      const int se = d.cell_start[cc0 + m + 1];
      // the column's records within reach: the column lies at least rho from
      // i across x and y (a face of i's own cell bounds it, clamped records
      // included), so only |z - z_i| <= sqrt(ri^2 - rho^2) can pair, which
      // is slices glo..ghi of the column (z from the bottom of cell az0)
      int lo = s0, hi_end = se;
      if (!exact_all) {
        const float gx = fmaxf(dcx < 0 ? oxi : (dcx > 0 ? csx - oxi : 0.f), 0.f);
        const float gy = fmaxf(dcy < 0 ? oyi : (dcy > 0 ? csy - oyi : 0.f), 0.f);
        const float rho2 = gx * gx + gy * gy;
        if (rho2 > ri2) continue;
        const float w = sqrtf(ri2 - rho2);
        const int glo = min(max((int)floorf((bz0 - w) * inv_hz), 0), gmax);
        const int ghi = min(max((int)floorf((bz0 + w) * inv_hz), 0), gmax);
        const int alo = glo >> 2, klo = glo & 3, ahi = ghi >> 2, khi = ghi & 3;
        const int4 ol = d.soff[cc0 + alo], oh = d.soff[cc0 + ahi];
        lo = klo == 0 ? (alo == 0 ? s0 : (alo == 1 ? s1 : s2)) : (klo == 1 ? ol.x : (klo == 2 ? ol.y : ol.z));
        hi_end = khi == 3 ? oh.w : (khi == 0 ? oh.x : (khi == 1 ? oh.y : oh.z));
      }
      // (the record arrays hold 4 spare entries, so a batch may read past
      // the range's end; those candidates are not valid)
      const int2 *fp = d.sinfo + lo;
      const float4 *qp = d.sxyz + lo;
      for (int pb = lo; pb < hi_end; pb += 4, fp += 4, qp += 4) {
        int2 f[4];
        float4 q[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          f[u] = fp[u];
          q[u] = qp[u];
        }
        // compiler-only barrier: the four candidates' loads stay issued
        // together (not sunk into the accept branches); emits no instruction
        asm volatile("" ::: "memory");
        bool in[4], near[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int p = pb + u;
          const bool valid = (p < hi_end) & (f[u].x != i) & (f[u].y == g);
          const float bz = p >= s2 ? bz2 : (p >= s1 ? bz1 : bz0);
          const float dx = bx - q[u].x, dy = by - q[u].y, dz = bz - q[u].z;
          const float r2 = dx * dx + dy * dy + dz * dz;
          const float t = hcf * (hi + q[u].w);
          const float t2 = t * t;
          in[u] = valid & (r2 < t2);
          near[u] = valid & (exact_all | (fabsf(r2 - t2) <= 1e-4f * t2));
        }
        // the exact predicate decides the (rare) candidates in the band
        if (near[0] | near[1] | near[2] | near[3]) {
#pragma unroll
          for (int u = 0; u < 4; ++u)
            if (near[u]) in[u] = in_radius(xi, d.xh[cur][f[u].x], d.cellfac);
        }
        // append without branches: the group being filled shifts left and
        // the entry enters at w; a full group is stored (entries past max_nb
        // are only counted)
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const bool take = in[u] && cnt < d.max_nb;
          pend.x = take ? pend.y : pend.x;
          pend.y = take ? pend.z : pend.y;
          pend.z = take ? pend.w : pend.z;
          pend.w = take ? f[u].x : pend.w;
          cnt += in[u] ? 1 : 0;
          if (take & ((cnt & 3) == 0)) {
            *wp = pend;
            wp += np;
          }
        }
      }
    }
  }
  if (i < d.n) {
    const int kept = cnt < d.max_nb ? cnt : d.max_nb;
    d.nnb[i] = kept;
    if (cnt > d.max_nb) d.flags->overflow = 1;
    // the pair passes read rows in groups of four: the rest of the last
    // group holds the particle itself (a pair term of exact zeros)
    // (the k = kept & 3 pending entries sit in the last k places of pend)
    if (kept & 3) {
      const int k = kept & 3;
      const int4 last = k == 1 ? make_int4(pend.w, i, i, i)
                               : (k == 2 ? make_int4(pend.z, pend.w, i, i) : make_int4(pend.y, pend.z, pend.w, i));
      *wp = last;
    }
  }
  // deterministic totals via block reduction + integer atomics
  const double bt = block_reduce<0>((double)cnt, sh);
  const double bm = block_reduce<1>((double)cnt, sh);
  if (threadIdx.x == 0) {
    atomicAdd((unsigned long long *)&d.flags->directed_pairs, (unsigned long long)bt);
    atomicMax(&d.flags->max_count, (int)bm);
  }
}

// LDS staging lists of the mixed-precision pair passes, per tile of kBlock
// consecutive particles after every neighbour build: the tile's particles
// and all their neighbour entries go into an LDS hash set (open
// addressing); its distinct values (the staged set, ~6 % of the ~7000
// entries: each neighbour appears in many rows) are sorted (bitonic, in
// registers and lane shuffles) so a pass loads the set in long contiguous
// runs; each value's hash
// slot then records its position in the set, and every row entry is
// rewritten as that position.  The hash table's layout depends on insertion
// order, the sorted set does not.  Tiles whose set exceeds stage_max keep
// stage_cnt = -1 and gather from HBM.  Rows keep their order, so staged and
// HBM gathers sum in the same order.
constexpr int kStageThreads = 256;
constexpr int kHashSlots = 2048;  // >= 2 kStageMax: load <= 1/2 for sets of <= kStageMax values
static_assert(kHashSlots >= 2 * kStageMax && (kHashSlots & (kHashSlots - 1)) == 0, "hash table size");
static_assert(kStagePow2 >= kStageMax && (kStagePow2 & (kStagePow2 - 1)) == 0, "sort padding");
static_assert(kStageThreads == 2 * kBlock, "two threads per particle of a tile");
__device__ inline unsigned stage_hash(int v) { return ((unsigned)v * 2654435761u) >> 21; }  // 11 bits
static_assert(kStagePow2 == 4 * kStageThreads, "the set sort holds four values per thread");

// Bitonic sort (ascending) of the first P values (P a power of two <=
// kStagePow2) of the block, element e = 4 tid + r held in v[r]; elements at
// e >= P must hold values >= every value below P.  Compare-exchanges within
// a thread (j = 1, 2) stay in registers, those within a wavefront
// (4 <= j < 256) exchange through lane shuffles, and only j >= 256 goes
// through LDS (buf: kStagePow2 ints), two barriers each.  Every thread of
// the block calls this with the same P.
__device__ inline void stage_sort(int (&v)[4], int P, int *buf) {
  const int tid = threadIdx.x;
  for (int k = 2; k <= P; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      if (j >= 4) {
        int o[4];
        if (j >= 256) {
          reinterpret_cast<int4 *>(buf)[tid] = make_int4(v[0], v[1], v[2], v[3]);
          __syncthreads();
          const int4 w = reinterpret_cast<const int4 *>(buf)[tid ^ (j >> 2)];
          __syncthreads();
          o[0] = w.x;
          o[1] = w.y;
          o[2] = w.z;
          o[3] = w.w;
        } else {
#pragma unroll
          for (int r = 0; r < 4; ++r) o[r] = __shfl_xor(v[r], j >> 2, 64);
        }
        // the element at the lower index of its pair keeps the minimum when
        // the run (e & k) sorts ascending
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int e = 4 * tid + r;
          const bool keep_min = ((e & j) == 0) == ((e & k) == 0);
          v[r] = keep_min ? min(v[r], o[r]) : max(v[r], o[r]);
        }
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          if (r & j) continue;
          const int e = 4 * tid + r;
          const int a = v[r], b = v[r ^ j];
          const bool asc = (e & k) == 0;
          v[r] = asc ? min(a, b) : max(a, b);
          v[r ^ j] = asc ? max(a, b) : min(a, b);
        }
      }
    }
  }
}

__global__ __launch_bounds__(kStageThreads) void k_stage(Dev d) {
  if (!d.flags->do_search) return;
  __shared__ int uniq[kStagePow2];
  __shared__ int ht[kHashSlots];
  __shared__ unsigned short hp[kHashSlots];
  __shared__ int nset, full;
  __shared__ int sh[34];
  const int tid = threadIdx.x, li = tid % kBlock, half = tid / kBlock;
  const long long np = d.npad;
  constexpr unsigned mask = kHashSlots - 1;
  for (int t = blockIdx.x; t < d.ntiles; t += gridDim.x) {
    const int i = t * kBlock + li;
    const int c = (i < d.n && d.group[i] >= 0) ? d.nnb[i] : 0;
    for (int p = tid; p < kHashSlots; p += kStageThreads) ht[p] = -1;
    if (tid == 0) {
      nset = 0;
      full = 0;
    }
    __syncthreads();
    // a value is counted by the thread whose CAS claims its slot; probing
    // stops once the set has more than stage_max values (the tile then
    // gathers from HBM), and every probe sequence ends within kHashSlots.
    // A group's four home-slot CAS are issued together (they are
    // independent; a lane's atomics on one slot apply in order, so the
    // repeated padding entries find their own value); only collisions
    // probe on, one value at a time.
    auto insert4 = [&](const int4 v4) {
      if (*(volatile int *)&full) return;
      const int vv[4] = {v4.x, v4.y, v4.z, v4.w};
      unsigned s[4];
      int old[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) s[u] = stage_hash(vv[u]) & mask;
#pragma unroll
      for (int u = 0; u < 4; ++u) old[u] = atomicCAS(&ht[s[u]], -1, vv[u]);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        int o = old[u];
        unsigned sl = s[u];
        for (int probe = 1; probe < kHashSlots && o != -1 && o != vv[u]; ++probe) {
          if (*(volatile int *)&full) break;
          sl = (sl + 1) & mask;
          o = atomicCAS(&ht[sl], -1, vv[u]);
        }
        if (o == -1 && atomicAdd(&nset, 1) >= d.stage_max) full = 1;
      }
    };
    // a row's groups of four entries (the last padded with i itself), the
    // two halves taking alternate groups, four groups' loads in flight; the
    // groups read past a row's end lie inside the row allocation
    const int4 *row = reinterpret_cast<const int4 *>(d.nbr) + i;
    const int ng = (c + 3) >> 2;
    if (i < d.n) {
      if (half == 0) insert4(make_int4(i, i, i, i));
      for (int g0 = half; g0 < ng; g0 += 8) {
        int4 v[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) v[b] = row[(long long)(g0 + 2 * b) * np];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          if (g0 + 2 * b >= ng) break;
          insert4(v[b]);
        }
      }
    }
    __syncthreads();
    if (full || nset > d.stage_max) {
      if (tid == 0) d.stage_cnt[t] = -1;
      __syncthreads();
      continue;
    }
    // compact the table: thread tid takes slots [8 tid, 8 tid + 8), one
    // block scan places its values (the order is irrelevant: sorted next)
    constexpr int kSlotsPer = kHashSlots / kStageThreads;
    static_assert(kSlotsPer == 8, "two int4 loads of slots per thread");
    const int4 h0 = reinterpret_cast<const int4 *>(ht)[2 * tid];
    const int4 h1 = reinterpret_cast<const int4 *>(ht)[2 * tid + 1];
    const int hv[8] = {h0.x, h0.y, h0.z, h0.w, h1.x, h1.y, h1.z, h1.w};
    int f = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) f += hv[q] >= 0 ? 1 : 0;
    int U;
    int e = block_scan_excl(f, sh, &U);
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (hv[q] >= 0) uniq[e++] = hv[q];
    int P = 1;
    while (P < U) P <<= 1;
    __syncthreads();
    int sv[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) sv[r] = 4 * tid + r < U ? uniq[4 * tid + r] : 0x7fffffff;
    __syncthreads();
    stage_sort(sv, P, uniq);
    // each value's slot records its position in the sorted set
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int p = 4 * tid + r;
      if (p < U) {
        const int v = sv[r];
        unsigned s = stage_hash(v) & mask;
        while (ht[s] != v) s = (s + 1) & mask;
        hp[s] = (unsigned short)p;
        d.stage[(long long)t * kStageMax + p] = v;
      }
    }
    __syncthreads();
    auto pos = [&](int v) -> unsigned short {
      unsigned s = stage_hash(v) & mask;
      while (ht[s] != v) s = (s + 1) & mask;
      return hp[s];
    };
    if (i < d.n) {
      if (half == 0) d.selfl[i] = pos(i);
      // the padding entries hold i, so they become the particle's own position
      ushort4 *rowl = reinterpret_cast<ushort4 *>(d.nbrl) + i;
      for (int g0 = half; g0 < ng; g0 += 8) {
        int4 v[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) v[b] = row[(long long)(g0 + 2 * b) * np];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          if (g0 + 2 * b >= ng) break;
          rowl[(long long)(g0 + 2 * b) * np] = make_ushort4(pos(v[b].x), pos(v[b].y), pos(v[b].z), pos(v[b].w));
        }
      }
    }
    if (tid == 0) d.stage_cnt[t] = U;
    __syncthreads();
  }
}

// Domain::CalcGradCorrMatrix (Source/Domain.cpp:1832-1940):
// L_i = (-sum_j V_j GK_ij x_ij (x) x_ij)^-1 over i's neighbour row (FP64
// gather, V_j = m_j / rho_j, x_ij = x_i - x_j; in 2D the sum's zz entry is
// set to 1 before the inversion), symmetric; the identity
// where |det| < 1e-10 (InvLog, External/matvec.h:673-693) and for fixed
// particles.  Runs when the schedule set gc_due.
WF_KERNEL void k_gradcorr(Dev d, int cur) {
  if (!d.flags->gc_due) return;
  const long long np = d.npad;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < d.n; i += gridDim.x * kBlock) {
    double L[6] = {1.0, 1.0, 1.0, 0.0, 0.0, 0.0};
    if (d.group[i] >= 0) {
      const double4 xi = d.xh[cur][i];
      const KH<double> ki = kern_h_t<double>(xi.w, d.dim);
      double m[6] = {0, 0, 0, 0, 0, 0};
      const int cnt = d.nnb[i];
      for (int k = 0; k < cnt; ++k) {
        const int j = d.nbr[nb_slot(np, k, i)];
        const double4 xj = d.xh[cur][j];
        const double dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
        const double GK = gk_t<double>(dx * dx + dy * dy + dz * dz, pair_h_t<double>(ki, xi.w, xj.w));
        const double w = d.vol[cur][j] * GK;
        m[XX] -= w * dx * dx;
        m[YY] -= w * dy * dy;
        m[ZZ] -= w * dz * dz;
        m[XY] -= w * dx * dy;
        m[XZ] -= w * dx * dz;
        m[YZ] -= w * dy * dz;
      }
      if (d.dim == 2) m[ZZ] = 1.0;  // the 2D sum's zz entry (Source/Domain.cpp:1911-1915)
      const double c00 = m[YY] * m[ZZ] - m[YZ] * m[YZ], c01 = m[XZ] * m[YZ] - m[XY] * m[ZZ];
      const double c02 = m[XY] * m[YZ] - m[XZ] * m[YY];
      const double det = m[XX] * c00 + m[XY] * c01 + m[XZ] * c02;
      if (fabs(det) >= 1.0e-10) {
        L[XX] = c00 / det;
        L[XY] = c01 / det;
        L[XZ] = c02 / det;
        L[YY] = (m[XX] * m[ZZ] - m[XZ] * m[XZ]) / det;
        L[YZ] = (m[XZ] * m[XY] - m[XX] * m[YZ]) / det;
        L[ZZ] = (m[XX] * m[YY] - m[XY] * m[XY]) / det;
      }
    }
    for (int k = 0; k < 6; ++k) d.gcm[k * np + i] = L[k];
  }
}

void launch_gradcorr(const Dev &d, hipStream_t st, int cur) {
  k_gradcorr<<<nblk_small(d.n), kBlock, 0, st>>>(d, cur);
}

// plane strain: StartAcceleration resets each particle's z velocity at the
// top of every step (Source/Domain.cpp:1592); the FP32 shadow with it
WF_KERNEL void k_plane_strain(Dev d) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= d.n) return;
  d.vm[i].z = 0.0;
  if (d.mixed) d.vmf[i].z = 0.0f;
}
void launch_plane_strain(const Dev &d, hipStream_t st) { k_plane_strain<<<d.nblocks, kBlock, 0, st>>>(d); }

void launch_schedule(const Dev &d, hipStream_t st, int force) {
  k_schedule<<<1, 64, 0, st>>>(d, force);
}

void launch_neighbours(const Dev &d, hipStream_t st, int cur) {
  k_cell_hist<<<nblk((long long)d.n + d.nrig), kBlock, 0, st>>>(d, cur);
  k_scan<<<scan_chunks(d.sncells) + (d.nrig ? scan_chunks(d.ncells) : 0), kScanThreads, 0, st>>>(d);
  k_scatter<<<nblk((long long)d.n + d.nrig), kBlock, 0, st>>>(d);
  // (at most 8192 blocks: an idle launch of that size costs about as much
  // as the other search kernels', and 16k wavefronts keep the cells' loads
  // in flight)
  {
    const int g = nblk((long long)(d.sncells + kSortGroup - 1) / kSortGroup * 64);
    k_cell_sort<<<g < 8192 ? g : 8192, kBlock, 0, st>>>(d, cur);
  }
  k_build_nbr<<<nblk(d.n), kBlock, 0, st>>>(d, cur);
  if (d.stage_max > 0) k_stage<<<d.ntiles < 2048 ? d.ntiles : 2048, kStageThreads, 0, st>>>(d);
}

// ---------------------------------------------------------------------------
// pass 1: acceleration (CalcAccel + AccelReduction, GradientType 0, TI = 0)

// Where a pass reads neighbour j: straight from HBM (Src<MX>, row of global
// indices in nbr) or from the tile's LDS copy (row of staged positions in
// nbrl).  Both give the same values, so results do not depend on staging.
template <bool MX>
struct HbmJ {
  using S = Src<MX>;
  using R = typename S::R;
  using R4 = typename S::R4;
  // FP64 gathers: a batch of 4 in flight (8 costs more registers than it
  // gains); FP32: the whole batch
  static constexpr int kUnroll = MX ? kNbBatch : 4;
  const Dev &d;
  const int buf;
  const int i;
  // entries k0..k0+7 (k0 a multiple of 8) of particle i's row
  __device__ void fetch(int k0, int out[kNbBatch]) const {
    const int4 *row = reinterpret_cast<const int4 *>(d.nbr) + i;
    const int4 a = row[(long long)(k0 >> 2) * d.npad];
    const int4 b = row[(long long)((k0 >> 2) + 1) * d.npad];
    out[0] = a.x, out[1] = a.y, out[2] = a.z, out[3] = a.w;
    out[4] = b.x, out[5] = b.y, out[6] = b.z, out[7] = b.w;
  }
  __device__ int self() const { return i; }
  __device__ R4 x(int j) const { return S::x(d, buf, j); }
  __device__ R4 v(int j) const { return S::v(d, j); }
  __device__ void srec(int j, R t[6], R &rho) const { S::srec(d, j, t, rho); }
  // tensile-instability stress of particle j
  __device__ void tir(int j, R t[6]) const {
    for (int k = 0; k < 6; ++k) {
      if constexpr (MX) t[k] = d.tirf[(long long)k * d.npad + j];
      else t[k] = d.tir[(long long)k * d.npad + j];
    }
  }
  // P_j / rho_j^2 of the kinetic-energy pair term
  __device__ R pr(int j, R rhoj, R P0, R cs2, R rho0) const {
    if constexpr (MX) {
      return d.prf[j];
    } else {
      const R irj = rcp_t(rhoj);
      return (P0 + cs2 * (rhoj - rho0)) * irj * irj;
    }
  }
  __device__ bool lower(int i, int j) const { return i < j; }
};
// Calls f(j) for the entries of a neighbour row of cnt entries that belong
// to lane `sub` of the particle's `lanes` lanes: batches of kNbBatch entries
// sub, sub + lanes, sub + 2 lanes, ... in row order, each batch's indices
// fetched while the previous batch is processed.  Entries past cnt in the
// last batch call f with j = the particle itself: a particle paired with
// itself has x_ij = v_ij = 0, so every pair term is an exact zero and adding
// it changes no sum.
template <class G, class F>
__device__ inline void for_each_nb(const G &g, int cnt, int sub, int lanes, F &&f) {
  int cur[kNbBatch], nxt[kNbBatch];
  const int step = lanes * kNbBatch;
  int base = sub * kNbBatch;
  g.fetch(base, cur);
  const int self = g.self();
  for (; base < cnt; base += step) {
    g.fetch(base + step, nxt);
#pragma unroll G::kUnroll
    for (int u = 0; u < kNbBatch; ++u) f(base + u < cnt ? cur[u] : self);
#pragma unroll
    for (int u = 0; u < kNbBatch; ++u) cur[u] = nxt[u];
  }
}

// copy the tile's staged set into LDS (all NT threads of the block call
// these).  A thread's share (entries p = tid + r NT) is loaded in three
// rounds, each with all of the share's loads in flight: the set's particle
// indices, then the records (from particle 0 for entries past the set's
// end, so no load is conditional), then the LDS stores of the set's
// entries.  Entries past the set's end are stored to the spare slot
// kStageMax (the LDS arrays hold kStageMax + 1 records), so no store is
// conditional either and nothing moves a load down to its store.
// density / rates: (x, m/rho), (v, m) and, for rates, P / rho^2
template <int NT>
__device__ inline void stage_xv(const Dev &d, int t, int U, const float4 *gx, float4 *lx, float4 *lv, float *lp) {
  constexpr int R = (kStageMax + NT - 1) / NT;
  const int *st = d.stage + (long long)t * kStageMax;
  int j[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int p = (int)threadIdx.x + r * NT;
    j[r] = p < U ? st[p] : 0;
  }
  float4 x[R], v[R];
  float pr[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    x[r] = gx[j[r]];
    v[r] = d.vmf[j[r]];
    if (lp) pr[r] = d.prf[j[r]];
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int p = (int)threadIdx.x + r * NT;
    const int q = p < U ? p : kStageMax;
    lx[q] = x[r];
    lv[q] = v[r];
    if (lp) lp[q] = pr[r];
  }
  __syncthreads();
}
// acceleration: (x, rho), (v, m), sigma / rho^2 as (t0..t3), (t4, t5)
template <int NT>
__device__ inline void stage_accel(const Dev &d, int t, int U, const float4 *gx, float4 *lx, float4 *lv, float4 *lt,
                                   float2 *lt2) {
  constexpr int R = (kStageMax + NT - 1) / NT;
  const int *st = d.stage + (long long)t * kStageMax;
  const float4 *gs = reinterpret_cast<const float4 *>(d.srecf);  // (t0..t3), (t4, t5, rho, pad)
  int j[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int p = (int)threadIdx.x + r * NT;
    j[r] = p < U ? st[p] : 0;
  }
  float4 x[R], v[R], s0[R], s1[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    x[r] = gx[j[r]];
    v[r] = d.vmf[j[r]];
    s0[r] = gs[2 * j[r]];
    s1[r] = gs[2 * j[r] + 1];
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int p = (int)threadIdx.x + r * NT;
    const int q = p < U ? p : kStageMax;
    lx[q] = make_float4(x[r].x, x[r].y, x[r].z, s1[r].z);
    lv[q] = v[r];
    lt[q] = s0[r];
    lt2[q] = make_float2(s1[r].x, s1[r].y);
  }
  __syncthreads();
}

// cubic-spline kernel shape w(q) = W(q h) pi h^3 (Source/Functions.cpp:25-37)
template <typename R>
__device__ inline R w_shape(R q) {
  const R t = (R)2 - q;
  return q < (R)1 ? (R)1 - (R)1.5 * q * q + (R)0.75 * q * q * q : (q < (R)2 ? (R)0.25 * t * t * t : (R)0);
}

// TI: with the tensile-instability term (W(r)/W(d0))^n (TIR_i + TIR_j) of
// CalcAccel (Source/InteractionAlt.cpp:104-110); pairs form within one
// material, so n and d0 are the material's
// particle i's gradCorrM (symmetric, 6 components): the identity when the
// correction is off
template <bool GC, typename R>
__device__ inline void own_corr(const Dev &d, int i, R (&L)[6]) {
  L[XX] = L[YY] = L[ZZ] = (R)1;
  L[XY] = L[XZ] = L[YZ] = (R)0;
  if constexpr (GC)
    for (int k = 0; k < 6; ++k) L[k] = (R)d.gcm[(long long)k * d.npad + i];
}
// e = L x_ij: the direction the kernel gradient GK e takes under the
// correction, (GK L) x_ij of Source/InteractionAlt.cpp:118-133
template <bool GC, typename R>
__device__ inline void corr_dir(const R (&L)[6], R dx, R dy, R dz, R &ex, R &ey, R &ez) {
  if constexpr (GC) {
    ex = L[XX] * dx + L[XY] * dy + L[XZ] * dz;
    ey = L[XY] * dx + L[YY] * dy + L[YZ] * dz;
    ez = L[XZ] * dx + L[YZ] * dy + L[ZZ] * dz;
  } else {
    ex = dx, ey = dy, ez = dz;
  }
}

// GC: with the kernel-gradient correction (particle i's own gradCorrM
// applied to x_ij in the pair term)
template <bool MX, bool TI, bool GC, class G>
__device__ inline void accel_lane(const Dev &d, int cur, int i, const G &g, int sub, int lanes,
                                  typename Src<MX>::R (&acc)[3]) {
  using S = Src<MX>;
  using R = typename S::R;
  R ax = 0, ay = 0, az = 0;
  if (d.group[i] >= 0) {
    R Li[6];
    own_corr<GC>(d, i, Li);
    const DevMat &M = d.mats[d.mat[i]];
    const R alpha = (R)M.alpha, beta = (R)M.beta, cs = (R)M.Cs;
    const bool av = (M.alpha != 0.0 || M.beta != 0.0);
    const auto xi = S::x(d, cur, i);
    const auto vi = S::v(d, i);
    R ti[6], rhoi;
    S::srec(d, i, ti, rhoi);
    const KH<R> ki = kern_h_t<R>(S::h_own(d, cur, i, xi), d.dim);
    const R acs = alpha * cs;
    R tri[6] = {0, 0, 0, 0, 0, 0};
    const R tin = (R)M.ti_n, tid = (R)M.ti_dist;
    if constexpr (TI) g.tir(i, tri);
    for_each_nb(g, d.nnb[i], sub, lanes, [&](int j) {
      const auto xj = g.x(j);
      const auto vj = g.v(j);
      R tj[6], rhoj;
      g.srec(j, tj, rhoj);
      const KH<R> kh = S::kh(ki, xi, xj);
      const R dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
      const R r2 = dx * dx + dy * dy + dz * dz;
      const R GK = gk_t<R>(r2, kh);
      // Monaghan viscosity for approaching pairs (selected, not branched):
      // mu = h vdot / D1, Pi = mu (alpha cs + beta mu) / D2 with a single
      // reciprocal of D1 D2 (D1 = r^2 + 0.01 h^2, D2 = (rho_i + rho_j) / 2)
      const R vdot = (vi.x - vj.x) * dx + (vi.y - vj.y) * dy + (vi.z - vj.z) * dz;
      const R D1 = r2 + (R)0.01 * kh.h * kh.h, D2 = (R)0.5 * (rhoi + rhoj);
      const R hv = kh.h * vdot * rcp_t(D1 * D2);
      const R pv = hv * (acs + beta * (hv * D2));
      const R pi = (av && vdot < (R)0) ? pv : (R)0;
      const R Mxx = ti[XX] + tj[XX] + pi;
      const R Myy = ti[YY] + tj[YY] + pi;
      const R Mzz = ti[ZZ] + tj[ZZ] + pi;
      R Mxy = ti[XY] + tj[XY];
      R Mxz = ti[XZ] + tj[XZ];
      R Myz = ti[YZ] + tj[YZ];
      R Mxx2 = Mxx, Myy2 = Myy, Mzz2 = Mzz;
      if constexpr (TI) {
        // a padded self entry has r2 = 0: q = 0 keeps the factor finite
        const R q = r2 > (R)0 ? r2 * rsqrt_t(r2) * kh.ih : (R)0;
        const R s = pow(w_shape<R>(q) / w_shape<R>(tid * kh.ih), tin);
        R trj[6];
        g.tir(j, trj);
        Mxx2 += s * (tri[XX] + trj[XX]);
        Myy2 += s * (tri[YY] + trj[YY]);
        Mzz2 += s * (tri[ZZ] + trj[ZZ]);
        Mxy += s * (tri[XY] + trj[XY]);
        Mxz += s * (tri[XZ] + trj[XZ]);
        Myz += s * (tri[YZ] + trj[YZ]);
      }
      R w = vj.w * GK;
      if constexpr (!MX) {
        // damage weight of the pair (Source/InteractionAlt.cpp:54-60)
        if (d.damage) {
          const R df = dam_weight(d.damD[i], d.damD[j]);
          w = df > (R)0 ? (df * vj.w) * GK : (R)0;
        }
      }
      R ex, ey, ez;
      corr_dir<GC>(Li, dx, dy, dz, ex, ey, ez);
      ax += w * (ex * Mxx2 + ey * Mxy + ez * Mxz);
      ay += w * (ex * Mxy + ey * Myy2 + ez * Myz);
      az += w * (ex * Mxz + ey * Myz + ez * Mzz2);
    });
  }
  acc[0] = ax;
  acc[1] = ay;
  acc[2] = az;
}

// ---------------------------------------------------------------------------
// Mixed-precision pair passes (no tensile instability).  A lane walks its
// particle's row in groups of four entries (rows are padded to a multiple
// of four with the particle itself, whose pair terms are exact zeros):
// lane `sub` of LANES takes groups sub, sub + LANES, ..., the row words two
// groups ahead in flight.  All four neighbours' records of a group are
// loaded before any is used and the pair arithmetic has no branch, so a
// group's loads are in flight together.  The records come from the tile's
// LDS copy (row entries are positions in the staged set, nbrl) or, for a
// tile whose set did not fit, straight from HBM (row entries are particle
// indices, nbr); both give the same values in the same order.

struct LdsRows {
  using G = ushort4;
  const ushort4 *row;
  long long np;
  int own;  // the particle's position in the staged set
  __device__ G group(int g) const { return row[(long long)g * np]; }
  __device__ G self() const {
    const unsigned short o = (unsigned short)own;
    return make_ushort4(o, o, o, o);
  }
};
struct HbmRows {
  using G = int4;
  const int4 *row;
  long long np;
  int own;  // the particle's index
  __device__ G group(int g) const { return row[(long long)g * np]; }
  __device__ G self() const { return make_int4(own, own, own, own); }
};
template <class G>
__device__ inline int entry(const G &c, int u) {
  return u == 0 ? (int)c.x : (u == 1 ? (int)c.y : (u == 2 ? (int)c.z : (int)c.w));
}

// Lane `sub` of LANES takes the group pairs (2k, 2k + 1) with k = sub,
// sub + LANES, ...: a row of ng groups costs its lanes ng groups in all
// (the second group of the last pair is a self group when ng is odd).  An
// iteration processes one pair, the next pair's row words loaded at its top
// so they are in flight while this pair is processed.  The iteration is one
// branch-free block (a branch between the groups lets the compiler sink the
// prefetch down to its use): a second group past the row's end is replaced
// by `self`, a group of four self entries, whose pair terms are zeros.  The
// groups beyond a row's end that the prefetch touches (at most 2 LANES + 1)
// lie inside the allocation (nb_rows keeps kRowSpare spare entries) and are
// never used.
template <int LANES, class Rows, class F>
__device__ inline void walk_groups(const Rows &rows, int cnt, int sub, const typename Rows::G &self,
                                   F &&group_fn) {
  static_assert(4 * (2 * LANES + 2) <= kRowSpare, "prefetch stays in the row allocation");
  const int ng = (cnt + 3) >> 2;
  int g = 2 * sub;
  typename Rows::G a = rows.group(g);
  typename Rows::G b = rows.group(g + 1);
  for (; g < ng; g += 2 * LANES) {
    const typename Rows::G a2 = rows.group(g + 2 * LANES);
    const typename Rows::G b2 = rows.group(g + 2 * LANES + 1);
    // compiler-only barrier: the prefetch stays issued here (it may not be
    // re-formed as a load next to its use); emits no instruction
    asm volatile("" ::: "memory");
    group_fn(a);
    // the second group's loads are not hoisted over the first group's
    // arithmetic (that would hold both groups' records in registers)
    __builtin_amdgcn_sched_barrier(0);
    group_fn(g + 1 < ng ? b : self);
    // the register moves of the hand-over stay at the end of the iteration
    // (moved earlier they would wait for the prefetch)
    __builtin_amdgcn_sched_barrier(0);
    a = a2;
    b = b2;
  }
}

// particle i's kernel constants for the FP32 gradient factor
struct GkF {
  float ih, gin_a, gin_b, gout;  // 1/h, 2.25 c5, -3 c5, -0.75 c5 h
};
__device__ inline GkF gk_consts(float h, int dim) {
  const KH<float> k = kern_h_t<float>(h, dim);
  return GkF{k.ih, 2.25f * k.c5, -3.0f * k.c5, -0.75f * k.c5 * h};
}
// GK of the cubic spline (Source/Functions.cpp:74-87) for r2 = |x_ij|^2:
// c5 (2.25 q - 3) for q < 1, -0.75 c5 h (2 - q)^2 / r for 1 <= q < 2, 0
// beyond, selected without a branch.  r2 = 0 (a padded self entry, or two
// coincident particles) gives q = 0 and GK = -3 c5, the reference's q = 0
// value, with every operand finite.
__device__ inline float gk_f(float r2, const GkF &k) {
  const float ri = __builtin_amdgcn_rsqf(fmaxf(r2, FLT_MIN));
  const float q = (r2 * ri) * k.ih;
  const float gin = fmaf(q, k.gin_a, k.gin_b);
  const float t = fmaxf(2.0f - q, 0.0f);
  const float gout = (t * t) * (ri * k.gout);
  return q < 1.0f ? gin : gout;
}

// acceleration records of neighbour j: X = (x, y, z, rho), V = (v, m),
// T = sigma / rho^2 as (xx, yy, zz, xy), T2 = (xz, yz)
struct AccelLds {
  const float4 *lx, *lv, *lt;
  const float2 *lt2;
  __device__ void load(int l, float4 &X, float4 &V, float4 &T, float2 &T2) const {
    X = lx[l];
    V = lv[l];
    T = lt[l];
    T2 = lt2[l];
  }
};
struct AccelHbm {
  const float4 *gx, *gv;
  const float4 *gs;  // StressRecF as two float4: (t0..t3), (t4, t5, rho, pad)
  __device__ void load(int j, float4 &X, float4 &V, float4 &T, float2 &T2) const {
    const float4 x = gx[j], s1 = gs[2 * j + 1];
    V = gv[j];
    T = gs[2 * j];
    X = make_float4(x.x, x.y, x.z, s1.z);
    T2 = make_float2(s1.x, s1.y);
  }
};

// CalcAccel + AccelReduction in FP32 (Source/InteractionAlt.cpp:12-260,
// GradientType 0): a_i = sum_j m_j GK (sigma_i/rho_i^2 + sigma_j/rho_j^2 +
// Pi_ij) e_ij with e = x_ij (or gradCorrM_i x_ij).  The particle's own term
// sigma_i/rho_i^2 (sum_j m_j GK e_ij) is applied once after the row.
// BETA: some material has a quadratic viscosity term (beta != 0); without
// it Pi = 2 alpha c h vdot / ((r^2 + 0.01 h^2) D)
template <int LANES, bool GC, bool BETA, class Rows, class Rec>
__device__ inline void accel_mx_lane(const Dev &d, int cur, int i, int sub, const Rows &rows, const Rec &rec,
                                     float (&acc)[3]) {
  float ax = 0, ay = 0, az = 0, sx = 0, sy = 0, sz = 0;
  float Li[6];
  own_corr<GC>(d, i, Li);
  const DevMat &M = d.mats[d.mat[i]];
  const float4 xi = d.xhf[cur][i], vi = d.vmf[i];
  const StressRecF ri = d.srecf[i];
  const float h = (float)d.xh[cur][i].w;
  const GkF k = gk_consts(h, d.dim);
  const float h2c = 0.01f * h * h, hh = 2.0f * h;
  const float acs = (float)(M.alpha * M.Cs), bh = 0.5f * (float)M.beta, hha = hh * acs;
  walk_groups<LANES>(rows, d.nnb[i], sub, rows.self(), [&](const typename Rows::G &c) {
    float4 X[4], V[4], T[4];
    float2 T2[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) rec.load(entry(c, u), X[u], V[u], T[u], T2[u]);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const float dx = xi.x - X[u].x, dy = xi.y - X[u].y, dz = xi.z - X[u].z;
      const float r2 = dx * dx + dy * dy + dz * dz;
      const float GK = gk_f(r2, k);
      // Monaghan viscosity for approaching pairs: with D = rho_i + rho_j,
      // mu = h vdot / (r^2 + 0.01 h^2), Pi = mu (alpha c + beta mu) / (D/2)
      const float vdot = (vi.x - V[u].x) * dx + (vi.y - V[u].y) * dy + (vi.z - V[u].z) * dz;
      const float D = ri.rho + X[u].w;
      const float rd = __builtin_amdgcn_rcpf((r2 + h2c) * D);
      float pv;
      if constexpr (BETA) {
        const float hv = (hh * vdot) * rd;
        pv = hv * fmaf(bh, hv * D, acs);
      } else {
        pv = (hha * vdot) * rd;
      }
      const float pi = vdot < 0.0f ? pv : 0.0f;
      const float w = V[u].w * GK;
      float ex, ey, ez;
      corr_dir<GC>(Li, dx, dy, dz, ex, ey, ez);
      const float ux = fmaf(T[u].x, ex, fmaf(T[u].w, ey, fmaf(T2[u].x, ez, pi * ex)));
      const float uy = fmaf(T[u].w, ex, fmaf(T[u].y, ey, fmaf(T2[u].y, ez, pi * ey)));
      const float uz = fmaf(T2[u].x, ex, fmaf(T2[u].y, ey, fmaf(T[u].z, ez, pi * ez)));
      ax = fmaf(w, ux, ax);
      ay = fmaf(w, uy, ay);
      az = fmaf(w, uz, az);
      sx = fmaf(w, ex, sx);
      sy = fmaf(w, ey, sy);
      sz = fmaf(w, ez, sz);
    }
  });
  acc[0] = ax + (ri.t[XX] * sx + ri.t[XY] * sy + ri.t[XZ] * sz);
  acc[1] = ay + (ri.t[XY] * sx + ri.t[YY] * sy + ri.t[YZ] * sz);
  acc[2] = az + (ri.t[XZ] * sx + ri.t[YZ] * sy + ri.t[ZZ] * sz);
}

// FP64: HBM gathers, LANES threads per particle
template <int LANES>
__global__ __launch_bounds__(kBlock *LANES) void k_accel_f64(Dev d, int cur) {
  __shared__ double red[(LANES > 1 ? LANES - 1 : 1) * 3 * kBlock];
  const int t = xcd_tile(blockIdx.x, gridDim.x), sub = threadIdx.x / kBlock;
  const int i = t * kBlock + (int)(threadIdx.x % kBlock);
  double a[3] = {0, 0, 0};
  if (i < d.n) {
    const HbmJ<false> g{d, cur, i};
    if (d.gc_on) {
      if (d.ti_on) accel_lane<false, true, true>(d, cur, i, g, sub, LANES, a);
      else accel_lane<false, false, true>(d, cur, i, g, sub, LANES, a);
    } else {
      if (d.ti_on) accel_lane<false, true, false>(d, cur, i, g, sub, LANES, a);
      else accel_lane<false, false, false>(d, cur, i, g, sub, LANES, a);
    }
  }
  combine_lanes<LANES>(a, red);
  // plane strain: the pair forces have no z part (Source/InteractionAlt.cpp:
  // 184; the corrected-gradient terms keep theirs)
  if (d.dim == 2 && !d.gc_on) a[2] = 0.0;
  if (i < d.n && sub == 0) d.acc[i] = make_double4(a[0], a[1], a[2], 0.0);
}

// mixed precision: LDS-staged gathers where the tile is staged, else HBM
template <int LANES, bool GC, bool BETA>
__device__ inline void k_accel_mx_body(const Dev &d, int cur) {
  __shared__ float4 lx[kStageMax + 1], lv[kStageMax + 1], lt[kStageMax + 1];
  __shared__ float2 lt2[kStageMax + 1];
  __shared__ float red[(LANES > 1 ? LANES - 1 : 1) * 3 * kBlock];
  const int t = xcd_tile(blockIdx.x, gridDim.x), sub = threadIdx.x / kBlock;
  const int i = t * kBlock + (int)(threadIdx.x % kBlock);
  const int U = d.stage_cnt[t];
  float a[3] = {0, 0, 0};
  if (U < 0) {
    if (i < d.n && d.group[i] >= 0) {
      const HbmRows rows{reinterpret_cast<const int4 *>(d.nbr) + i, d.npad, i};
      const AccelHbm rec{d.xhf[cur], d.vmf, reinterpret_cast<const float4 *>(d.srecf)};
      accel_mx_lane<LANES, GC, BETA>(d, cur, i, sub, rows, rec, a);
    }
  } else {
    stage_accel<kBlock * LANES>(d, t, U, d.xhf[cur], lx, lv, lt, lt2);
    if (i < d.n && d.group[i] >= 0) {
      const LdsRows rows{reinterpret_cast<const ushort4 *>(d.nbrl) + i, d.npad, (int)d.selfl[i]};
      const AccelLds rec{lx, lv, lt, lt2};
      accel_mx_lane<LANES, GC, BETA>(d, cur, i, sub, rows, rec, a);
    }
  }
  combine_lanes<LANES>(a, red);
  if (d.dim == 2 && !d.gc_on) a[2] = 0.0f;  // plane strain, as the FP64 pass
  if (i < d.n && sub == 0) d.acc[i] = make_double4((double)a[0], (double)a[1], (double)a[2], 0.0);
}

// the kernels with and without the quadratic viscosity term.  The staged
// records fill ~51 KB of LDS, so three blocks share a CU: with two or more
// lanes per particle the kernels are held to the registers of three waves
// per SIMD
template <int LANES, bool GC>
__global__ __launch_bounds__(kBlock *LANES) __attribute__((amdgpu_waves_per_eu(LANES >= 2 ? 3 : 1))) void k_accel_mx_b(
    Dev d, int cur) {
  k_accel_mx_body<LANES, GC, true>(d, cur);
}
template <int LANES, bool GC>
__global__ __launch_bounds__(kBlock *LANES) __attribute__((amdgpu_waves_per_eu(LANES >= 2 ? 3 : 1))) void k_accel_mx_nb(
    Dev d, int cur) {
  k_accel_mx_body<LANES, GC, false>(d, cur);
}

// mixed precision with tensile instability: HBM gathers (the staged
// records do not carry the TI stress)
template <int LANES>
__global__ __launch_bounds__(kBlock *LANES) void k_accel_mx_ti(Dev d, int cur) {
  __shared__ float red[(LANES > 1 ? LANES - 1 : 1) * 3 * kBlock];
  const int t = xcd_tile(blockIdx.x, gridDim.x), sub = threadIdx.x / kBlock;
  const int i = t * kBlock + (int)(threadIdx.x % kBlock);
  float a[3] = {0, 0, 0};
  if (i < d.n) {
    const HbmJ<true> g{d, cur, i};
    if (d.gc_on) accel_lane<true, true, true>(d, cur, i, g, sub, LANES, a);
    else accel_lane<true, true, false>(d, cur, i, g, sub, LANES, a);
  }
  combine_lanes<LANES>(a, red);
  if (d.dim == 2 && !d.gc_on) a[2] = 0.0f;  // plane strain, as the FP64 pass
  if (i < d.n && sub == 0) d.acc[i] = make_double4((double)a[0], (double)a[1], (double)a[2], 0.0);
}

#define WF_LANES_SWITCH(lanes, KERNEL, ...)                                    \
  switch (lanes) {                                                            \
    case 1: KERNEL<1><<<nblk(d.n), kBlock * 1, 0, st>>>(__VA_ARGS__); break;  \
    case 2: KERNEL<2><<<nblk(d.n), kBlock * 2, 0, st>>>(__VA_ARGS__); break;  \
    default: KERNEL<4><<<nblk(d.n), kBlock * 4, 0, st>>>(__VA_ARGS__); break; \
  }
// the mixed-precision passes are also instantiated per gradient correction
#define WF_LANES_GC_SWITCH(lanes, gc, KERNEL, ...)                   \
  if (gc) {                                                         \
    switch (lanes) {                                                \
      case 1: KERNEL<1, true><<<nblk(d.n), kBlock * 1, 0, st>>>(__VA_ARGS__); break;  \
      case 2: KERNEL<2, true><<<nblk(d.n), kBlock * 2, 0, st>>>(__VA_ARGS__); break;  \
      default: KERNEL<4, true><<<nblk(d.n), kBlock * 4, 0, st>>>(__VA_ARGS__); break; \
    }                                                               \
  } else {                                                          \
    switch (lanes) {                                                \
      case 1: KERNEL<1, false><<<nblk(d.n), kBlock * 1, 0, st>>>(__VA_ARGS__); break;  \
      case 2: KERNEL<2, false><<<nblk(d.n), kBlock * 2, 0, st>>>(__VA_ARGS__); break;  \
      default: KERNEL<4, false><<<nblk(d.n), kBlock * 4, 0, st>>>(__VA_ARGS__); break; \
    }                                                               \
  }

// StartAcceleration's tensile-instability stress, 3D branch
// (Source/Domain.cpp:1529-1563): principal stresses s_k > 0 become
// -TI s_k / rho^2, the others 0, rotated back with the eigenvectors
// ((Vec Val) Vec^T).  The eigen-decomposition is cyclic Jacobi (Golub & Van
// Loan, Alg. 8.5.3) in FP64; the result is a spectral function of sigma, so
// it does not depend on the eigenvectors' signs or order.
WF_KERNEL void k_tir(Dev d, int cur) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= d.n) return;
  const long long np = d.npad;
  double out[6] = {0, 0, 0, 0, 0, 0};
  if (d.group[i] >= 0 && d.mats[d.mat[i]].ti > 0.0) {
    const StressRec r = d.srec[i];
    double A[3][3] = {{r.s[XX], r.s[XY], r.s[XZ]}, {r.s[XY], r.s[YY], r.s[YZ]}, {r.s[XZ], r.s[YZ], r.s[ZZ]}};
    double V[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int sweep = 0; sweep < 64; ++sweep) {
      const double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
      const double dia = A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2];
      if (off == 0.0 || off <= 1e-40 * dia) break;
      for (int p = 0; p < 2; ++p)
        for (int q = p + 1; q < 3; ++q) {
          if (A[p][q] == 0.0) continue;
          const double th = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
          const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
          const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
          for (int k = 0; k < 3; ++k) {
            const double akp = A[k][p], akq = A[k][q];
            A[k][p] = c * akp - s * akq;
            A[k][q] = s * akp + c * akq;
          }
          for (int k = 0; k < 3; ++k) {
            const double apk = A[p][k], aqk = A[q][k];
            A[p][k] = c * apk - s * aqk;
            A[q][k] = s * apk + c * aqk;
          }
          A[p][q] = A[q][p] = 0.0;
          for (int k = 0; k < 3; ++k) {
            const double vkp = V[k][p], vkq = V[k][q];
            V[k][p] = c * vkp - s * vkq;
            V[k][q] = s * vkp + c * vkq;
          }
        }
    }
    const double pc = d.mats[d.mat[i]].ti / (r.rho * r.rho);
    double w[3];
    for (int k = 0; k < 3; ++k) w[k] = A[k][k] > 0 ? -pc * A[k][k] : 0.0;
    const int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
    for (int m = 0; m < 6; ++m) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += (V[ia[m]][k] * w[k]) * V[ib[m]][k];
      out[m] = s;
    }
  }
  for (int m = 0; m < 6; ++m) d.tir[(long long)m * np + i] = out[m];
  if (d.mixed)
    for (int m = 0; m < 6; ++m) d.tirf[(long long)m * np + i] = (float)out[m];
}

void launch_tir(const Dev &d, hipStream_t st, int cur) { k_tir<<<nblk(d.n), kBlock, 0, st>>>(d, cur); }

void launch_accel(const Dev &d, hipStream_t st, int cur) {
  if (d.mixed && d.ti_on) {
    WF_LANES_SWITCH(d.lanes_mx, k_accel_mx_ti, d, cur)
  } else if (d.mixed && d.beta_on) {
    WF_LANES_GC_SWITCH(d.lanes_mx, d.gc_on, k_accel_mx_b, d, cur)
  } else if (d.mixed) {
    WF_LANES_GC_SWITCH(d.lanes_mx, d.gc_on, k_accel_mx_nb, d, cur)
  } else {
    WF_LANES_SWITCH(d.lanes_f64, k_accel_f64, d, cur)
  }
}

// ---------------------------------------------------------------------------
// velocity kicks with the UserAcc BC semantics (Source/WeldForm.cpp:51-92)
// and the reaction observer at the first GeneralAfter of the step.
//   Leapfrog  (Source/SolverLeapfrog.cpp:228-273):
//       [first step] v -= a*0.5*dt ;  v += a*0.5*(dt_prev + dt)
//   KickDrift (Source/SolverKickDrift.cpp:208-231): v += a*0.5*dt_prev


WF_KERNEL void k_kick(Dev d, BcTable bc, int scheme, double dt_prev, double dt, int kphase) {
  __shared__ double sh[kBlock / 64];
  __shared__ double bcv[kMaxBc][3];
  __shared__ int bcon[kMaxBc];
  const bool observe = kphase == K_WHOLE || kphase == K_FIRST || kphase == K_OBSERVE;
  const int phase = kphase == K_WHOLE_NOOBS ? K_WHOLE : (kphase == K_FIRST_NOOBS ? K_FIRST : kphase);
  const int isfirst = d.flags->step == 0;
  if (dt < 0) {
    dt = d.flags->dt;
    dt_prev = d.flags->dt_prev;
  }
  if (threadIdx.x < bc.n) {
    // velocity of each BC at the step's time (UserAcc, Source/WeldForm.cpp:67-82)
    const int b = threadIdx.x;
    int on = 1;
    double f = 1.0;
    if (bc.type[b] == 0 && bc.value_type[b] == 1) {
      if (bc.amp[b] < 0) on = 0;
      else f = bc.amp_factor[b] * amp_value(d.amps[bc.amp[b]], d.flags->time);
    }
    for (int k = 0; k < 3; ++k) bcv[b][k] = bc.value_type[b] == 1 ? f * bc.v[b][k] : bc.v[b][k];
    bcon[b] = on;
  }
  __syncthreads();
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double4 a = make_double4(0, 0, 0, 0);
  double m = 0.0;
  int id = -0x7fffffff;
  if (i < d.n) {
    a = d.acc[i];
    const double4 vm = d.vm[i];
    m = vm.w;
    id = d.id[i];
  }
  if (i < d.n && phase != K_OBSERVE) {
    const double4 vm = d.vm[i];
    double v[3] = {vm.x, vm.y, vm.z};
    double av[3] = {a.x, a.y, a.z};
    if (!isfinite(a.x) || !isfinite(a.y) || !isfinite(a.z)) d.flags->nan = 1;
    bool fixed = false;
    for (int b = 0; b < bc.n; ++b) {
      if (bc.zone[b] != id || !bcon[b]) continue;
      if (bc.type[b] == 0) {
        fixed = true;
        for (int k = 0; k < 3; ++k) {
          av[k] = 0.0;
          v[k] = bcv[b][k];
        }
      } else if (bc.type[b] == 2) {
        d.T[i] = bc.T[b];
      }
    }
    if (!fixed) {
      if (scheme == WF_LEAPFROG) {
        if (isfirst && phase != K_SECOND)
          for (int k = 0; k < 3; ++k) v[k] -= av[k] * 0.5 * dt;
        if (phase != K_FIRST)
          for (int k = 0; k < 3; ++k) v[k] += av[k] * 0.5 * (dt_prev + dt);
      } else if (phase != K_FIRST) {
        for (int k = 0; k < 3; ++k) v[k] += av[k] * 0.5 * dt_prev;
      }
    }
    d.vm[i] = make_double4(v[0], v[1], v[2], vm.w);
    if (d.mixed) d.vmf[i] = make_float4((float)v[0], (float)v[1], (float)v[2], (float)vm.w);
    d.acc[i] = make_double4(av[0], av[1], av[2], 0.0);
  }
  for (int z = 0; z < bc.nz && observe; ++z) {
    const bool in = (id == bc.rzone[z]);
    const double fx = block_reduce<0>(in ? m * a.x : 0.0, sh);
    const double fy = block_reduce<0>(in ? m * a.y : 0.0, sh);
    const double fz = block_reduce<0>(in ? m * a.z : 0.0, sh);
    if (threadIdx.x == 0) {
      part(d, P_REACT + 3 * z, blockIdx.x) = fx;
      part(d, P_REACT + 3 * z + 1, blockIdx.x) = fy;
      part(d, P_REACT + 3 * z + 2, blockIdx.x) = fz;
    }
  }
}

// KickDrift's second half kick, v += a*dt/2 (Source/SolverKickDrift.cpp:261-265)
WF_KERNEL void k_kick2(Dev d, double dt) {
  if (dt < 0) dt = d.flags->dt;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= d.n) return;
  const double4 a = d.acc[i];
  double4 v = d.vm[i];
  v.x += a.x * dt / 2.;
  v.y += a.y * dt / 2.;
  v.z += a.z * dt / 2.;
  d.vm[i] = v;
  if (d.mixed) d.vmf[i] = make_float4((float)v.x, (float)v.y, (float)v.z, (float)v.w);
}

void launch_kick(const Dev &d, hipStream_t st, const BcTable &bc, int scheme, double dt_prev, double dt, int phase) {
  k_kick<<<d.nblocks, kBlock, 0, st>>>(d, bc, scheme, dt_prev, dt, phase);
}

// CheckMinTSVel / CheckMinTSAccel (Source/Domain.cpp:265-319): per particle
// the smallest neighbour distance over CFL (Cs + |v|), or sqrt_h_a
// sqrt(h / |a|) (the acceleration criterion, evaluated last, wins when both
// are on); block minima, then AdaptiveTimeStep (:206-239) in one thread.
WF_KERNEL void k_ts_check(Dev d, int cur) {
#pragma clang fp contract(off)
  __shared__ double sh[kBlock / 64];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double test = INFINITY;
  if (i < d.n) {
    const double4 xi = d.xh[cur][i];
    if (d.auto_ts & 2) {
      const double4 a = d.acc[i];
      test = d.sqrt_h_a * sqrt(xi.w / sqrt(a.x * a.x + a.y * a.y + a.z * a.z));
    } else {
      double mn = 1000.;
      const int cnt = d.nnb[i];
      for (int k = 0; k < cnt; ++k) {
        const double4 xj = d.xh[cur][d.nbr[nb_slot(d.npad, k, i)]];
        const double dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
        const double dd = sqrt(dx * dx + dy * dy + dz * dz);
        if (dd < mn) mn = dd;
      }
      const double4 v = d.vm[i];
      test = d.cfl * mn / (d.mats[d.mat[i]].Cs + sqrt(v.x * v.x + v.y * v.y + v.z * v.z));
    }
  }
  const double r = block_reduce<2>(test, sh);
  if (threadIdx.x == 0) part(d, P_DTMIN, blockIdx.x) = r;
}

__global__ __launch_bounds__(1024) void k_ts_adapt(Dev d) {
  __shared__ double sh[16];
  double v = INFINITY;
  for (int b = threadIdx.x; b < d.nblocks; b += blockDim.x) v = fmin(v, part(d, P_DTMIN, b));
  const double r = block_reduce<2>(v, sh);
  // the contact criterion: the least dt_fext of this step's pairs
  const bool cont_ts = d.contact && (d.auto_ts & 4);
  double fts = INFINITY;
  if (cont_ts) {
    for (int b = threadIdx.x; b < d.nblocks; b += blockDim.x) fts = fmin(fts, part(d, P_FTS, b));
    fts = block_reduce<2>(fts, sh);
  }
  if (threadIdx.x != 0) return;
  DevFlags *f = d.flags;
  if (fts < INFINITY) f->min_force_ts = fts;
  const double dint = f->dtint, dmin = fmin(dint, r);
  double dt = f->dt;
  f->dt_prev = dt;
  if (dint > dmin) {
    if (dt < dmin) dt = 2.0 * dt * dmin / (dt + dmin);
    else dt = dmin;
  } else {
    if (dint != dt) dt = 2.0 * dt * dint / (dt + dint);
    else dt = dint;
  }
  if (cont_ts && f->min_force_ts < dt) dt = f->min_force_ts;
  if (dt < d.min_ts) dt = d.min_ts;
  f->dt = dt;
}

void launch_time_step(const Dev &d, hipStream_t st, int cur) {
  k_ts_check<<<d.nblocks, kBlock, 0, st>>>(d, cur);
  k_ts_adapt<<<1, 1024, 0, st>>>(d);
}

void launch_kick2(const Dev &d, hipStream_t st, double dt) { k_kick2<<<nblk(d.n), kBlock, 0, st>>>(d, dt); }

// ---------------------------------------------------------------------------
// pass 2: density rate + density update + drift
// (CalcDensInc + DensReduction, Source/InteractionAlt.cpp:483-640;
//  Source/SolverLeapfrog.cpp:280-301).  Reads x_n, v_{n+1/2}, m/rho_n and
// writes rho_{n+1}, m/rho_{n+1}, x_{n+1} into the other buffer.

template <bool MX, bool GC, class G>
__device__ inline typename Src<MX>::R dens_lane(const Dev &d, int cur, int i, const G &g, int sub, int lanes) {
  using S = Src<MX>;
  using R = typename S::R;
  R s = 0;
  if (d.group[i] >= 0) {
    R Li[6];
    own_corr<GC>(d, i, Li);
    const auto xi = S::x(d, cur, i);
    const auto vi = S::v(d, i);
    const KH<R> ki = kern_h_t<R>(S::h_own(d, cur, i, xi), d.dim);
    for_each_nb(g, d.nnb[i], sub, lanes, [&](int j) {
      const auto xj = g.x(j);
      const auto vj = g.v(j);
      const R volj = MX ? (R)xj.w : S::vol(d, cur, j, xj);
      const R dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
      const R GK = gk_t<R>(dx * dx + dy * dy + dz * dz, S::kh(ki, xi, xj));
      R ex, ey, ez;
      corr_dir<GC>(Li, dx, dy, dz, ex, ey, ez);
      if constexpr (MX) {
        s += (volj * GK) * ((vi.x - vj.x) * ex + (vi.y - vj.y) * ey + (vi.z - vj.z) * ez);
      } else {
        const R t = (vi.x - vj.x) * (GK * ex) + (vi.y - vj.y) * (GK * ey) + (vi.z - vj.z) * (GK * ez);
        R wv = volj;
        // damage weight of the pair (Source/InteractionAlt.cpp:558-572)
        if (d.damage) {
          const R df = dam_weight(d.damD[i], d.damD[j]);
          wv = df > (R)0 ? df * volj : (R)0;
        }
        s += wv * t;
      }
    });
  }
  return s;
}

// the particle's own FP64 state the update reads
struct DensOwn {
  double rho;
  double4 x, v, u;
};
__device__ inline DensOwn dens_own(const Dev &d, int cur, int i) {
  return DensOwn{d.rho[cur][i], d.xh[cur][i], d.vm[i], d.disp[i]};
}

template <bool MX>
__device__ inline void dens_update(const Dev &d, int cur, int i, double s, double dt, int do_update,
                                   const DensOwn &o) {
  if (dt < 0) dt = d.flags->dt;
  const double rhoi = o.rho;
  const double drho = s * rhoi;
  d.drho[i] = drho;
  if (!do_update) return;
  const int nxt = cur ^ 1;
  const double4 xd = o.x;
  const double4 vd = o.v;
  const double rn = rhoi + dt * drho;
  const double vn = vd.w / rn;
  d.rho[nxt][i] = rn;
  d.vol[nxt][i] = vn;
  const double ux = vd.x * dt, uy = vd.y * dt, uz = vd.z * dt;
  double4 u = o.u;
  u.x += ux;
  u.y += uy;
  u.z += uz;
  d.disp[i] = u;
  const double4 xn = make_double4(xd.x + ux, xd.y + uy, xd.z + uz, xd.w);
  d.xh[nxt][i] = xn;
  if (MX) {
    const double *o = d.mats[d.mat[i]].xc;
    d.xhf[nxt][i] = make_float4((float)(xn.x - o[0]), (float)(xn.y - o[1]), (float)(xn.z - o[2]), (float)vn);
    const DevMat &M = d.mats[d.mat[i]];
    d.prf[i] = (float)((M.P0 + M.Cs * M.Cs * (rn - M.ref_density)) / (rn * rn));
  }
}

template <int LANES>
__global__ __launch_bounds__(kBlock *LANES) void k_dens_f64(Dev d, int cur, double dt, int do_update) {
  __shared__ double red[(LANES > 1 ? LANES - 1 : 1) * kBlock];
  const int t = xcd_tile(blockIdx.x, gridDim.x), sub = threadIdx.x / kBlock;
  const int i = t * kBlock + (int)(threadIdx.x % kBlock);
  double s[1] = {0};
  if (i < d.n) {
    const HbmJ<false> g{d, cur, i};
    s[0] = d.gc_on ? dens_lane<false, true>(d, cur, i, g, sub, LANES) : dens_lane<false, false>(d, cur, i, g, sub, LANES);
  }
  combine_lanes<LANES>(s, red);
  if (i < d.n && sub == 0) dens_update<false>(d, cur, i, s[0], dt, do_update, dens_own(d, cur, i));
}

// records of the density and rate passes: X = (x, y, z, m/rho), V = (v, m)
// and, for the rates, P / rho^2
struct XvLds {
  const float4 *lx, *lv;
  const float *lp;
  __device__ void load(int l, float4 &X, float4 &V) const {
    X = lx[l];
    V = lv[l];
  }
  __device__ float pr(int l) const { return lp[l]; }
};
struct XvHbm {
  const float4 *gx, *gv;
  const float *gp;
  __device__ void load(int j, float4 &X, float4 &V) const {
    X = gx[j];
    V = gv[j];
  }
  __device__ float pr(int j) const { return gp[j]; }
};

// CalcDensInc + DensReduction in FP32 (Source/InteractionAlt.cpp:483-640):
// sum_j (m_j / rho_j) v_ij . (GK e_ij); the caller multiplies by rho_i
template <int LANES, bool GC, class Rows, class Rec>
__device__ inline float dens_mx_lane(const Dev &d, int cur, int i, int sub, const Rows &rows, const Rec &rec) {
  float s = 0;
  float Li[6];
  own_corr<GC>(d, i, Li);
  const float4 xi = d.xhf[cur][i], vi = d.vmf[i];
  const GkF k = gk_consts((float)d.xh[cur][i].w, d.dim);
  walk_groups<LANES>(rows, d.nnb[i], sub, rows.self(), [&](const typename Rows::G &c) {
    float4 X[4], V[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) rec.load(entry(c, u), X[u], V[u]);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const float dx = xi.x - X[u].x, dy = xi.y - X[u].y, dz = xi.z - X[u].z;
      const float GK = gk_f(dx * dx + dy * dy + dz * dz, k);
      float ex, ey, ez;
      corr_dir<GC>(Li, dx, dy, dz, ex, ey, ez);
      const float t = (vi.x - V[u].x) * ex + (vi.y - V[u].y) * ey + (vi.z - V[u].z) * ez;
      s = fmaf(X[u].w * GK, t, s);
    }
  });
  return s;
}

template <int LANES, bool GC>
__global__ __launch_bounds__(kBlock *LANES) void k_dens_mx(Dev d, int cur, double dt, int do_update) {
  __shared__ float4 lx[kStageMax + 1], lv[kStageMax + 1];
  __shared__ float red[(LANES > 1 ? LANES - 1 : 1) * kBlock];
  const int t = xcd_tile(blockIdx.x, gridDim.x), sub = threadIdx.x / kBlock;
  const int i = t * kBlock + (int)(threadIdx.x % kBlock);
  const int U = d.stage_cnt[t];
  // the update's FP64 inputs are loaded now and arrive during the row walk
  DensOwn own{};
  if (i < d.n && sub == 0) own = dens_own(d, cur, i);
  float s[1] = {0};
  if (U < 0) {
    if (i < d.n && d.group[i] >= 0) {
      const HbmRows rows{reinterpret_cast<const int4 *>(d.nbr) + i, d.npad, i};
      s[0] = dens_mx_lane<LANES, GC>(d, cur, i, sub, rows, XvHbm{d.xhf[cur], d.vmf, nullptr});
    }
  } else {
    stage_xv<kBlock * LANES>(d, t, U, d.xhf[cur], lx, lv, nullptr);
    if (i < d.n && d.group[i] >= 0) {
      const LdsRows rows{reinterpret_cast<const ushort4 *>(d.nbrl) + i, d.npad, (int)d.selfl[i]};
      s[0] = dens_mx_lane<LANES, GC>(d, cur, i, sub, rows, XvLds{lx, lv, nullptr});
    }
  }
  combine_lanes<LANES>(s, red);
  if (i < d.n && sub == 0) dens_update<true>(d, cur, i, (double)s[0], dt, do_update, own);
}

void launch_dens_drift(const Dev &d, hipStream_t st, int cur, double dt, int do_update) {
  if (d.mixed) {
    WF_LANES_GC_SWITCH(d.lanes_mx, d.gc_on, k_dens_mx, d, cur, dt, do_update)
  } else {
    WF_LANES_SWITCH(d.lanes_f64, k_dens_f64, d, cur, dt, do_update)
  }
}

// ---------------------------------------------------------------------------
// pass 3: strain-rate / rotation-rate gather (CalcRateTensors +
// RateTensorsReduction, Source/InteractionAlt.cpp:263-479) with the kinetic
// energy pair term (Source/Domain.cpp:1752-1808), then the per-particle
// stress update (Particle::CalcStressStrain, Source/Particle.cpp:960-1067)
// with plastic work heat (Source/Particle.cpp:922-939), the temperature
// update of ThermalCalcs (Source/Thermal.cpp:726-734) and the internal
// energy term (Source/Domain.cpp:1810-1829).  The gather and the update are
// separate kernels so the gather keeps a small register footprint.

// one particle's rate-tensor gather: v[0..5] strain rate, v[6..8] rotation
// rate (w01, w02, w12), v[9] the kinetic-energy pair sum
// LK: the Locking sums' weight, m_j / rho_j rounded to float (the FP32
// path's weight already is)
template <bool MX, bool GC, bool LK, class G>
__device__ inline void rates_lane(const Dev &d, int buf, int i, const G &g, int sub, int lanes,
                                  typename Src<MX>::R (&v)[10]) {
  using Sr = Src<MX>;
  using R = typename Sr::R;
  R sr[6] = {0, 0, 0, 0, 0, 0}, w[3] = {0, 0, 0};
  R ekin = 0;
  if (d.group[i] >= 0) {
    R Li[6];
    own_corr<GC>(d, i, Li);
    const DevMat &M = d.mats[d.mat[i]];
    const auto xi = Sr::x(d, buf, i);
    const auto vi = Sr::v(d, i);
    const R P0 = (R)M.P0, cs2 = (R)(M.Cs * M.Cs), rho0 = (R)M.ref_density;
    const R pri = g.pr(MX ? g.self() : i, MX ? (R)0 : Sr::rho(d, buf, i, xi, vi), P0, cs2, rho0);
    const R mi = vi.w;
    const KH<R> ki = kern_h_t<R>(Sr::h_own(d, buf, i, xi), d.dim);
    for_each_nb(g, d.nnb[i], sub, lanes, [&](int j) {
      const auto xj = g.x(j);
      const auto vj = g.v(j);
      R volj = MX ? (R)xj.w : Sr::vol(d, buf, j, xj);
      if constexpr (LK && !MX) volj = (R)(float)volj;
      const R dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
      const R GK = gk_t<R>(dx * dx + dy * dy + dz * dz, Sr::kh(ki, xi, xj));
      const R vx = vi.x - vj.x, vy = vi.y - vj.y, vz = vi.z - vj.z;
      const R c = volj * ((R)-0.5 * GK);
      R ex, ey, ez;
      corr_dir<GC>(Li, dx, dy, dz, ex, ey, ez);
      sr[XX] += c * ((R)2 * vx * ex);
      sr[YY] += c * ((R)2 * vy * ey);
      sr[ZZ] += c * ((R)2 * vz * ez);
      sr[XY] += c * (vx * ey + vy * ex);
      sr[XZ] += c * (vx * ez + vz * ex);
      sr[YZ] += c * (vy * ez + vz * ey);
      w[0] += c * (vx * ey - vy * ex);
      w[1] += c * (vx * ez - vz * ex);
      w[2] += c * (vy * ez - vz * ey);
      // kinetic-energy pair term (Source/Domain.cpp:1776-1792): the pair
      // is oriented lower-index first; its sum over particles vanishes
      const R prj = g.pr(j, MX ? (R)0 : Sr::rho(d, buf, j, xj, vj), P0, cs2, rho0);
      const bool lo = g.lower(i, j);
      const R msec = lo ? vj.w : mi;
      const R term = (R)0.5 * msec * (pri + prj) * (vx * (GK * dx) + vy * (GK * dy) + vz * (GK * dz));
      ekin += lo ? term : -term;
    });
  }
  for (int k = 0; k < 6; ++k) v[k] = sr[k];
  for (int k = 0; k < 3; ++k) v[6 + k] = w[k];
  v[9] = ekin;
}

template <typename R>
__device__ inline double rates_store(const Dev &d, int i, const R (&v)[10]) {
  const long long np = d.npad;
  for (int k = 0; k < 6; ++k) d.sr[k * np + i] = (double)v[k];
  for (int k = 0; k < 3; ++k) d.rr[k * np + i] = (double)v[6 + k];
  return (double)v[9];
}

template <int LANES>
__global__ __launch_bounds__(kBlock *LANES) void k_rates_f64(Dev d, int buf) {
  __shared__ double red[(LANES > 1 ? LANES - 1 : 1) * 10 * kBlock];
  __shared__ double sh[kBlock * LANES / 64];
  const int t = xcd_tile(blockIdx.x, gridDim.x), sub = threadIdx.x / kBlock;
  const int i = t * kBlock + (int)(threadIdx.x % kBlock);
  double v[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (i < d.n) {
    const HbmJ<false> g{d, buf, i};
    if (d.gc_on) {
      if (d.locking) rates_lane<false, true, true>(d, buf, i, g, sub, LANES, v);
      else rates_lane<false, true, false>(d, buf, i, g, sub, LANES, v);
    } else {
      if (d.locking) rates_lane<false, false, true>(d, buf, i, g, sub, LANES, v);
      else rates_lane<false, false, false>(d, buf, i, g, sub, LANES, v);
    }
  }
  combine_lanes<LANES>(v, red);
  const double e_kin = (i < d.n && sub == 0) ? rates_store(d, i, v) : 0.0;
  const double r1 = block_reduce<0>(e_kin, sh);
  if (threadIdx.x == 0) part(d, P_KIN, blockIdx.x) = r1;
}

// CalcRateTensors + RateTensorsReduction in FP32
// (Source/InteractionAlt.cpp:263-479) with the kinetic-energy pair term
// (Source/Domain.cpp:1776-1792).  With c_j = (m_j / rho_j) GK and
// (a, b, g) = c_j e, the sums kept per lane are sum v_x a, v_y b, v_z g
// (diagonal) and the six cross products sum v_x b, v_y a, v_x g, v_z a,
// v_y g, v_z b; strain rate and rotation are their halved sums and
// differences (eps_xy = -(P + Q)/2, w_01 = -(P - Q)/2).  The kinetic term
// carries the lower-index particle's partner mass with a plus sign and the
// higher's with a minus sign; its factor 1/2 is applied at the end.
// v[0..5] strain rate, v[6..8] rotation (w01, w02, w12), v[9] the kinetic sum.
template <int LANES, bool GC, bool STAGED, class Rows, class Rec>
__device__ inline void rates_mx_lane(const Dev &d, int buf, int i, int own, int sub, const Rows &rows,
                                     const Rec &rec, float (&v)[10]) {
  float dxx = 0, dyy = 0, dzz = 0, pxy = 0, qxy = 0, pxz = 0, qxz = 0, pyz = 0, qyz = 0, ek = 0;
  float Li[6];
  own_corr<GC>(d, i, Li);
  const float4 xi = d.xhf[buf][i], vi = d.vmf[i];
  const float pri = d.prf[i], mneg = -vi.w;
  const GkF k = gk_consts((float)d.xh[buf][i].w, d.dim);
  walk_groups<LANES>(rows, d.nnb[i], sub, rows.self(), [&](const typename Rows::G &c) {
    float4 X[4], V[4];
    float P[4];
    int l[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      l[u] = entry(c, u);
      rec.load(l[u], X[u], V[u]);
      P[u] = rec.pr(l[u]);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const float dx = xi.x - X[u].x, dy = xi.y - X[u].y, dz = xi.z - X[u].z;
      const float GK = gk_f(dx * dx + dy * dy + dz * dz, k);
      const float vx = vi.x - V[u].x, vy = vi.y - V[u].y, vz = vi.z - V[u].z;
      const float cj = X[u].w * GK;
      float ex, ey, ez;
      corr_dir<GC>(Li, dx, dy, dz, ex, ey, ez);
      const float a = cj * ex, b = cj * ey, g = cj * ez;
      dxx = fmaf(vx, a, dxx);
      dyy = fmaf(vy, b, dyy);
      dzz = fmaf(vz, g, dzz);
      pxy = fmaf(vx, b, pxy);
      qxy = fmaf(vy, a, qxy);
      pxz = fmaf(vx, g, pxz);
      qxz = fmaf(vz, a, qxz);
      pyz = fmaf(vy, g, pyz);
      qyz = fmaf(vz, b, qyz);
      // staged positions follow particle index order
      const bool lo = STAGED ? own < l[u] : i < l[u];
      const float ms = lo ? V[u].w : mneg;
      ek = fmaf(ms * (pri + P[u]), GK * (vx * dx + vy * dy + vz * dz), ek);
    }
  });
  v[XX] = -dxx;
  v[YY] = -dyy;
  v[ZZ] = -dzz;
  v[XY] = -0.5f * (pxy + qxy);
  v[XZ] = -0.5f * (pxz + qxz);
  v[YZ] = -0.5f * (pyz + qyz);
  v[6] = -0.5f * (pxy - qxy);
  v[7] = -0.5f * (pxz - qxz);
  v[8] = -0.5f * (pyz - qyz);
  v[9] = 0.5f * ek;
}

template <int LANES, bool GC>
__global__ __launch_bounds__(kBlock *LANES) void k_rates_mx(Dev d, int buf) {
  __shared__ float4 lx[kStageMax + 1], lv[kStageMax + 1];
  __shared__ float lp[kStageMax + 1];
  __shared__ float red[(LANES > 1 ? LANES - 1 : 1) * 10 * kBlock];
  __shared__ double sh[kBlock * LANES / 64];
  const int t = xcd_tile(blockIdx.x, gridDim.x), sub = threadIdx.x / kBlock;
  const int i = t * kBlock + (int)(threadIdx.x % kBlock);
  const int U = d.stage_cnt[t];
  float v[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (U < 0) {
    if (i < d.n && d.group[i] >= 0) {
      const HbmRows rows{reinterpret_cast<const int4 *>(d.nbr) + i, d.npad, i};
      rates_mx_lane<LANES, GC, false>(d, buf, i, i, sub, rows, XvHbm{d.xhf[buf], d.vmf, d.prf}, v);
    }
  } else {
    stage_xv<kBlock * LANES>(d, t, U, d.xhf[buf], lx, lv, lp);
    if (i < d.n && d.group[i] >= 0) {
      const LdsRows rows{reinterpret_cast<const ushort4 *>(d.nbrl) + i, d.npad, (int)d.selfl[i]};
      rates_mx_lane<LANES, GC, true>(d, buf, i, d.selfl[i], sub, rows, XvLds{lx, lv, lp}, v);
    }
  }
  combine_lanes<LANES>(v, red);
  const double e_kin = (i < d.n && sub == 0) ? rates_store(d, i, v) : 0.0;
  const double r1 = block_reduce<0>(e_kin, sh);
  if (threadIdx.x == 0) part(d, P_KIN, blockIdx.x) = r1;
}

template <bool MX>
WF_KERNEL void k_stress(Dev d, int buf, double dt) {
  if (dt < 0) dt = d.flags->dt;
  __shared__ double sh[kBlock / 64];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const long long np = d.npad;
  double e_int = 0.0, e_plw = 0.0, tmax = -INFINITY, tmin = INFINITY;
  if (i < d.n) {
    const int g = d.group[i];
    double SR[6], W[3];
    for (int k = 0; k < 6; ++k) SR[k] = d.sr[k * np + i];
    for (int k = 0; k < 3; ++k) W[k] = d.rr[k * np + i];
    if (g >= 0) {
      const DevMat &M = d.mats[d.mat[i]];
      const double rho = d.rho[buf][i];
      const double P = M.P0 + (M.Cs * M.Cs) * (rho - M.ref_density);
      double S[6];
      for (int k = 0; k < 6; ++k) S[k] = d.S[k * np + i];
      // Jaumann terms S*W^T + W*S with W antisymmetric
      const double Sf[3][3] = {{S[XX], S[XY], S[XZ]}, {S[XY], S[YY], S[YZ]}, {S[XZ], S[YZ], S[ZZ]}};
      const double Wf[3][3] = {{0.0, W[0], W[1]}, {-W[0], 0.0, W[2]}, {-W[1], -W[2], 0.0}};
      const int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
      const double c3 = 1.0 / 3.0 * (SR[XX] + SR[YY] + SR[ZZ]);
      const double G2 = 2.0 * M.G;
      double Sn[6];
      for (int k = 0; k < 6; ++k) {
        const int a = ia[k], b = ib[k];
        double jt = 0.0;
        for (int q = 0; q < 3; ++q) jt += Sf[a][q] * Wf[b][q] + Wf[a][q] * Sf[q][b];
        const double dev = k < 3 ? SR[k] - c3 : SR[k];
        Sn[k] = dt * (G2 * dev + jt) + S[k];
      }
      const double effsr = sqrt(0.5 * ((SR[XX] - SR[YY]) * (SR[XX] - SR[YY]) + (SR[YY] - SR[ZZ]) * (SR[YY] - SR[ZZ]) +
                                       (SR[ZZ] - SR[XX]) * (SR[ZZ] - SR[XX])) +
                                3.0 * (SR[XY] * SR[XY] + SR[YZ] * SR[YZ] + SR[XZ] * SR[XZ]));
      const double J2 = 0.5 * (Sn[XX] * Sn[XX] + 2.0 * Sn[XY] * Sn[XY] + 2.0 * Sn[XZ] * Sn[XZ] + Sn[YY] * Sn[YY] +
                               2.0 * Sn[YZ] * Sn[YZ] + Sn[ZZ] * Sn[ZZ]);
      const double sig_trial = sqrt(3.0 * J2);
      double Sy = d.sigy[i];
      double pl = d.pl[i];
      const double ratio = Sy / sig_trial;
      const double fac = (1.0 < ratio) ? 1.0 : ratio;
      for (int k = 0; k < 6; ++k) Sn[k] *= fac;
      const double T = d.T[i];
      if (M.model == WF_HOLLOMON) Sy = hol_yield(M, pl);
      else if (M.model == WF_JOHNSON_COOK) Sy = jc_yield(M, pl, effsr, T);
      double dep = 0.0;
      if (sig_trial > Sy) {
        dep = (sig_trial - Sy) / (3. * M.G + M.Ep);
        pl += dep;
        d.dpl[i] = dep;
        if (M.model == WF_BILINEAR) Sy += dep * M.Ep;
      }
      double sig[6];
      for (int k = 0; k < 6; ++k) sig[k] = k < 3 ? -P + Sn[k] : Sn[k];
      if (dep > 0.0) {
        const double f = dep / Sy;
        const double inc[6] = {f * (sig[XX] - 0.5 * (sig[YY] + sig[ZZ])), f * (sig[YY] - 0.5 * (sig[XX] + sig[ZZ])),
                               f * (sig[ZZ] - 0.5 * (sig[XX] + sig[YY])), 1.5 * f * sig[XY],
                               1.5 * f * sig[XZ],                          1.5 * f * sig[YZ]};
        for (int k = 0; k < 6; ++k) {
          d.eplinc[k * np + i] = inc[k];
          d.epl[k * np + i] += inc[k];
        }
      }
      for (int k = 0; k < 6; ++k) {
        d.S[k * np + i] = Sn[k];
        d.strain[k * np + i] += dt * SR[k];
      }
      d.effsr[i] = effsr;
      d.sigy[i] = Sy;
      d.pl[i] = pl;
      d.pres[i] = P;
      // Domain::CalcDamage, Johnson-Cook criterion (Source/Damage.cpp:9-46),
      // after the stress update: the (never reset) delta_pl_strain over the
      // fracture strain at the plastic rate, the triaxiality -P / Sigma_eq
      // (Particle::CalculateEquivalentStress, Source/Particle.cpp:623-633)
      // and T (0 without the thermal solver: the reference's is an
      // uninitialised private copy of a local set to 0)
      if (d.damage && M.dam_model == 1) {
        const double dpl = d.dpl[i];
        if (dpl > 0.0) {
          const double J2e = 0.5 * (Sn[XX] * Sn[XX] + 2.0 * Sn[XY] * Sn[XY] + 2.0 * Sn[XZ] * Sn[XZ] + Sn[YY] * Sn[YY] +
                                    2.0 * Sn[YZ] * Sn[YZ] + Sn[ZZ] * Sn[ZZ]);
          const double seq = sqrt(3.0 * J2e);
          if (seq > 1.0e-3) {
            const double ef = jc_fracture_strain(M, dpl / dt, -P / seq, d.thermal ? T : 0.0);
            double D = d.damD[i] + dpl / ef;
            if (D > 1.0) D = 1.0;
            d.damD[i] = D;
          }
        }
      }
      if (pl > 1.e-6) d.flags->yielded = 1;
      StressRec rec;
      for (int k = 0; k < 6; ++k) rec.s[k] = sig[k];
      rec.rho = rho;
      rec.irho2 = 1.0 / (rho * rho);
      d.srec[i] = rec;
      if (MX) {
        StressRecF rf;
        for (int k = 0; k < 6; ++k) rf.t[k] = (float)(sig[k] * rec.irho2);
        rf.rho = (float)rho;
        rf.pad = 0.0f;
        d.srecf[i] = rf;
      }
      const double vol = d.vol[buf][i];
      // plastic work heat, Source/ThermalStuct.cpp:14-39 (delta_pl_strain is
      // never reset by the reference, so the last increment keeps heating)
      if (d.pl_heating) {
        double q = d.qpl[i];
        if (d.dpl[i] > 0.0) {
          double dp[6];
          for (int k = 0; k < 6; ++k) dp[k] = d.eplinc[k * np + i] * (1. / dt);
          q = sig[XX] * dp[XX] + 2.0 * sig[XY] * dp[XY] + 2.0 * sig[XZ] * dp[XZ] + sig[YY] * dp[YY] +
              2.0 * sig[YZ] * dp[YZ] + sig[ZZ] * dp[ZZ];
          d.qpl[i] = q;
        }
        d.psen[i] += q * dt;
        e_plw = vol * q;
      }
      if (d.thermal) {
        const double Tn = T + dt * d.dTdt[i];
        d.T[i] = Tn;
        tmax = Tn;
        tmin = Tn;
      } else {
        e_int = vol * (Sn[XX] * SR[XX] + 2.0 * Sn[XY] * SR[XY] + 2.0 * Sn[XZ] * SR[XZ] + Sn[YY] * SR[YY] +
                       2.0 * Sn[YZ] * SR[YZ] + Sn[ZZ] * SR[ZZ]);
      }
    }
  }
  const double r0 = block_reduce<0>(e_int, sh);
  const double r2 = block_reduce<0>(e_plw, sh);
  const double r3 = block_reduce<1>(tmax, sh);
  const double r4 = block_reduce<2>(tmin, sh);
  if (threadIdx.x == 0) {
    part(d, P_INT, blockIdx.x) = r0;
    part(d, P_PLW, blockIdx.x) = r2;
    part(d, P_MAXT, blockIdx.x) = r3;
    part(d, P_MINT, blockIdx.x) = r4;
  }
}

void launch_rates(const Dev &d, hipStream_t st, int buf) {
  if (d.mixed) {
    WF_LANES_GC_SWITCH(d.lanes_mx, d.gc_on, k_rates_mx, d, buf)
  } else {
    WF_LANES_SWITCH(d.lanes_f64, k_rates_f64, d, buf)
  }
}

void launch_stress(const Dev &d, hipStream_t st, int buf, double dt) {
  if (d.mixed)
    k_stress<true><<<d.nblocks, kBlock, 0, st>>>(d, buf, dt);
  else
    k_stress<false><<<d.nblocks, kBlock, 0, st>>>(d, buf, dt);
}

// ---------------------------------------------------------------------------
// conduction (CalcTempInc, Source/Thermal.cpp:53-219) + convection
// (CalcConvHeat, :339-389) + thermal expansion (Source/ThermalStuct.cpp:7-12)
// The reference credits temp[first] += m and temp[second] -= m with m built
// from the second particle's m/rho; "first" is the particle whose cell the
// pair was found from (Source/Neighbour.cpp:418-477): the higher index
// within one cell, else the particle whose half stencil holds the other.

__device__ inline bool pair_first(int ci, int cj, int oi, int oj, int ny, int nz) {
  if (ci == cj) return oi > oj;
  const int zi = ci % nz, yi = (ci / nz) % ny, xi = ci / (nz * ny);
  const int zj = cj % nz, yj = (cj / nz) % ny, xj = cj / (nz * ny);
  const int dx = xj - xi, dy = yj - yi, dz = zj - zi;
  return dz == 1 || (dz == 0 && dy == 1) || (dz == 0 && dy == 0 && dx == 1);
}

WF_KERNEL void k_thermal(Dev d, int buf, double dt) {
  if (dt < 0) dt = d.flags->dt;
  __shared__ double sh[kBlock / 64];
  const int i = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  const long long np = d.npad;
  double e_int = 0.0, e_frw = 0.0, e_chc = 0.0;
  if (i < d.n && d.group[i] >= 0) {
    const DevMat &M = d.mats[d.mat[i]];
    const double4 xi = d.xh[buf][i];
    const double rho = d.rho[buf][i];
    const double m = d.vm[i].w;
    const double Ti = d.T[i];
    const double voli = d.vol[buf][i];
    if (d.thbc[i] == 1) {
      // (m/rho)^(2/3), (m/rho)^(1/3) in plane strain (Source/Thermal.cpp:354-366)
      const double dS2 = pow(m / rho, d.dim == 2 ? 1.0 / 3.0 : 2.0 / 3.0);
      d.qconv[i] = rho * d.hconv[i] * dS2 * (d.tinf[i] - Ti) / m;
    }
    const double kk = 4. * (M.k_T * M.k_T) / (M.k_T + M.k_T);
    const int ci = d.cell[i], oi = d.perm[i];
    double temp = 0.0;
    const KernH ki = kern_h(xi.w, d.dim);
    const int cnt = d.nnb[i];
    for (int k = 0; k < cnt; ++k) {
      const int j = d.nbr[nb_slot(np, k, i)];
      const double4 xj = d.xh[buf][j];
      const double dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
      const double r2 = dx * dx + dy * dy + dz * dz;
      const double GK = grad_kernel_r2(r2, pair_h(ki, xi.w, xj.w));
      const bool first = pair_first(ci, d.cell[j], oi, d.perm[j], d.cellno[1], d.cellno[2]);
      const double vsec = first ? d.vol[buf][j] : voli;
      // dot(x_ij, GK x_ij) / |x_ij|^2 = GK
      temp += vsec * kk * (Ti - d.T[j]) * GK;
    }
    const double f = 1. / (rho * M.cp_T);
    const double qc = d.heat_cond ? d.qcont[i] : 0.0;
    double dTdt = f * (temp + d.qconv[i] + d.qsrc[i] + d.qpl[i] * d.pl_frac + qc);
    if (d.heat_fric) dTdt += f * d.qfric[i];
    d.dTdt[i] = dTdt;
    if (d.qfric) {
      // the contact heat totals (Source/Thermal.cpp:194-210)
      e_frw = d.qfric[i] * m / rho;
      e_chc = qc * m / rho;
    }
    double SR[6];
    for (int k = 0; k < 6; ++k) SR[k] = d.sr[k * np + i];
    const double ex = M.th_exp * dTdt;
    for (int k = 0; k < 3; ++k) {
      SR[k] += ex;
      d.sr[k * np + i] = SR[k];
    }
    double S[6];
    for (int k = 0; k < 6; ++k) S[k] = d.S[k * np + i];
    e_int = voli * (S[XX] * SR[XX] + 2.0 * S[XY] * SR[XY] + 2.0 * S[XZ] * SR[XZ] + S[YY] * SR[YY] +
                    2.0 * S[YZ] * SR[YZ] + S[ZZ] * SR[ZZ]);
  }
  const double r0 = block_reduce<0>(e_int, sh);
  if (threadIdx.x == 0) part(d, P_INT, blockIdx.x) = r0;
  if (d.qfric) {
    const double r1 = block_reduce<0>(e_frw, sh);
    const double r2 = block_reduce<0>(e_chc, sh);
    if (threadIdx.x == 0) {
      part(d, P_FRW, blockIdx.x) = r1;
      part(d, P_CHC, blockIdx.x) = r2;
    }
  }
}

void launch_thermal(const Dev &d, hipStream_t st, int buf, double dt) {
  k_thermal<<<d.nblocks, kBlock, 0, st>>>(d, buf, dt);
}

// ---------------------------------------------------------------------------
// end of step: fixed-order reduction of the block partials, scalar updates,
// reaction history, and the end-of-step half of the neighbour schedule.
// One block per active slot sums that slot's partials (each thread its
// blocks in order, four loads in flight, then the wavefront butterfly and
// the wavefronts in order), so the slots reduce in parallel across CUs; the
// block that finishes last (a ticket counter, reset by that block for the
// next step) then applies them with one thread.

struct FinList {
  int n;
  int slot[P_COUNT];
  int op[P_COUNT];  // 0 sum, 1 max, 2 min
};

__device__ inline double fin_op(int op, double v, double x) {
  return op == 0 ? v + x : (op == 1 ? fmax(v, x) : fmin(v, x));
}

__device__ void fin_apply(const Dev &d, double dt, const FinList &L) {
  if (dt < 0) dt = d.flags->dt;
  DevFlags *f = d.flags;
  const long long step = f->step;
  const long long slot = step % d.hist_cap;
  double e_kin = 0.0, e_int = 0.0, plw = 0.0, tmax = 0.0, tmin = 0.0, cfsum = 0.0, creact = 0.0, extw = 0.0;
  for (int k = 0; k < L.n; ++k) {
    const int s = L.slot[k];
    // written by the other blocks of this launch: read past the L1 cache
    const double r = __hip_atomic_load(&d.fin[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (s == P_KIN) e_kin = r;
    if (s == P_INT) e_int = r;
    if (s == P_PLW) plw = r;
    if (s == P_MAXT) tmax = r;
    if (s == P_MINT) tmin = r;
    if (s >= P_REACT && s < P_CFORCE) d.react_hist[slot * kMaxZones * 3 + (s - P_REACT)] = r;
    if (s >= P_CFORCE && s < P_CFSUM) d.cf_hist[slot * kMaxMesh * 3 + (s - P_CFORCE)] = r;
    if (s == P_CFSUM) cfsum = r;
    if (s == P_CREACT) creact = r;
    if (s == P_EXTW) extw = r;
    if (s == P_FRW) d.scalars[S_FRW] += r * dt;
    if (s == P_CHC) d.scalars[S_CHC] += r * dt;
  }
  if (d.contact) {
    d.scalars[S_CFSUM] = cfsum;
    d.scalars[S_CREACT] = creact;
    d.scalars[S_EXTW] += extw;
  }
  d.scalars[S_KIN] += e_kin * dt;
  d.scalars[S_INT] += e_int * dt;
  if (d.pl_heating) d.scalars[S_PLW] += plw * dt;
  if (d.thermal) {
    d.scalars[S_MAXT] = fmax(0.0, tmax);
    d.scalars[S_MINT] = fmin(1000.0, tmin);
  }
  if (!isfinite(e_int)) f->nan = 1;
  if (f->step_yield) {
    if (f->ts_i == f->ts_nb_inc - 1) f->cleared = 1;
    f->ts_i++;
    if (f->ts_i > f->ts_nb_inc - 1) f->ts_i = 0;
  }
  f->step = step + 1;
  f->time += dt;
}

constexpr int kFinThreads = 1024;
__global__ __launch_bounds__(kFinThreads) void k_finalize(Dev d, double dt, FinList L) {
  __shared__ double sh[kFinThreads / 64];
  __shared__ int last;
  const int s = L.slot[blockIdx.x], op = L.op[blockIdx.x];
  const int nb = d.nblocks;
  double v = op == 0 ? 0.0 : (op == 1 ? -INFINITY : INFINITY);
  for (int b = threadIdx.x; b < nb; b += 4 * blockDim.x) {
    double x[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int bu = b + u * (int)blockDim.x;
      x[u] = bu < nb ? part(d, s, bu) : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (b + u * (int)blockDim.x < nb) v = fin_op(op, v, x[u]);
  }
  double r;
  if (op == 0) r = block_reduce<0>(v, sh);
  else if (op == 1) r = block_reduce<1>(v, sh);
  else r = block_reduce<2>(v, sh);
  if (threadIdx.x == 0) {
    d.fin[s] = r;
    __threadfence();
    last = atomicAdd(&d.flags->fin_ticket, 1u) == (unsigned)(L.n - 1);
  }
  __syncthreads();
  if (!last || threadIdx.x != 0) return;
  __threadfence();
  fin_apply(d, dt, L);
  d.flags->fin_ticket = 0;
}

void launch_finalize(const Dev &d, hipStream_t st, double dt, int nz) {
  FinList L{};
  auto add = [&](int s, int op) {
    L.slot[L.n] = s;
    L.op[L.n] = op;
    ++L.n;
  };
  add(P_INT, 0);
  add(P_KIN, 0);
  if (d.pl_heating) add(P_PLW, 0);
  if (d.thermal) {
    add(P_MAXT, 1);
    add(P_MINT, 2);
  }
  for (int s = P_REACT; s < P_REACT + 3 * nz; ++s) add(s, 0);
  if (d.contact) {
    for (int s = P_CFORCE; s < P_CFORCE + 3 * d.nmesh; ++s) add(s, 0);
    add(P_CFSUM, 0);
    add(P_CREACT, 0);
    add(P_EXTW, 0);
    if (d.thermal && d.qfric) {
      add(P_FRW, 0);
      add(P_CHC, 0);
    }
  }
  k_finalize<<<L.n, kFinThreads, 0, st>>>(d, dt, L);
}

// ---------------------------------------------------------------------------
// rigid contact planes: free-surface detection (Domain::CalculateSurface,
// Source/Domain.cpp:1415-1491), contact pairs (AllocateNbPair's contact
// branch + ContactNbSearch, Source/Neighbour.cpp:375-396,
// Source/Contact.cpp:57-132), the Wang penalty force with Coulomb friction
// (Source/Contact_Wang.cpp:15-348) and the rigid-body motion of the planes
// (TriMesh::Update, Source/Mesh.cpp:273-296; UpdateContactParticles,
// Source/Contact.cpp:41-55).  FP64 in both precision modes.


// After a neighbour build.  A particle was on the free surface (ID 1) when
// the search ran; the new surface test then uses the new rows.  Contact
// partners are the rigid particles, in index order, of an adjacent cell
// closer than h_i + h_r, for particles on the surface before and after.
// setup = 1: the search before the first step only sets the surface.
// setup = 2: with damage, every step after the search (CalculateSurface(1),
// Source/SolverLeapfrog.cpp:204-206): the surface ids only.
WF_KERNEL void k_surface(Dev d, int cur, int setup) {
  if (setup != 2 && !d.flags->do_search) return;
  const long long np = d.npad;
  {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= d.n) return;
    const int old_surf = d.id[i] == 1;
    const double4 xi = d.xh[cur][i];
    double nx = 0.0, ny = 0.0, nz = 0.0;
    const int cnt = d.nnb[i];
    // a free-surface particle has at most 46 neighbours (12 in 2D,
    // Source/Domain.cpp:1472-1474); with more the test fails whatever the
    // normal, so the row is not walked.  Otherwise the row's groups of four
    // entries, their gathers issued together; the sum runs over the entries
    // in row order
    const int nb_max = d.dim == 2 ? 12 : 46;
    const int4 *row = reinterpret_cast<const int4 *>(d.nbr) + i;
    for (int g = 0; cnt <= nb_max && 4 * g < cnt; ++g) {
      const int4 e = row[(long long)g * np];
      const int jj[4] = {e.x, e.y, e.z, e.w};
      double4 xj[4];
      double mj[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        xj[u] = d.xh[cur][jj[u]];
        mj[u] = d.pmass[jj[u]];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (4 * g + u >= cnt) break;
        nx += mj[u] * (xi.x - xj[u].x);
        ny += mj[u] * (xi.y - xj[u].y);
        nz += mj[u] * (xi.z - xj[u].z);
      }
    }
    const double f = 1. / d.totmass;
    nx *= f;
    ny *= f;
    nz *= f;
    const int surf = cnt <= nb_max && sqrt(nx * nx + ny * ny + nz * nz) >= 0.25 * xi.w;
    d.id[i] = surf ? 1 : d.id_orig[i];
    if (setup == 2) return;
    int c = 0;
    if (!setup && old_surf && surf) {
      // ContactNbSearch: the rigid particles of the 27 cells around i's own
      // (binned by k_cell_hist / k_scatter at this search), |x_ij| < h_i +
      // h_r; partners are kept in rigid-particle index order
      const int ci = d.cell[i];
      const int gz = d.cellno[2], gy = d.cellno[1], gx = d.cellno[0];
      const int cz = ci % gz, cy = (ci / gz) % gy, cx = ci / (gz * gy);
      for (int ax = cx - 1; ax <= cx + 1; ++ax) {
        if (ax < 0 || ax >= gx) continue;
        for (int ay = cy - 1; ay <= cy + 1; ++ay) {
          if (ay < 0 || ay >= gy) continue;
          const int c0 = (ax * gy + ay) * gz;
          const int zlo = cz > 0 ? cz - 1 : 0, zhi = cz + 1 < gz ? cz + 1 : gz - 1;
          for (int p = d.rstart[c0 + zlo]; p < d.rstart[c0 + zhi + 1]; ++p) {
            const int r = d.rsorted[p];
            const double4 xr = d.rx[r];
            const double dx = xi.x - xr.x, dy = xi.y - xr.y, dz = xi.z - xr.z;
            if (sqrt(dx * dx + dy * dy + dz * dz) < xi.w + xr.w) {
              if (c < kMaxContact) {
                int q = c;
                while (q > 0 && d.cidx[(long long)(q - 1) * np + i] > r) {
                  d.cidx[(long long)q * np + i] = d.cidx[(long long)(q - 1) * np + i];
                  --q;
                }
                d.cidx[(long long)q * np + i] = r;
              }
              ++c;
            }
          }
        }
      }
      if (c > kMaxContact) d.flags->overflow = 1;
    }
    d.ccnt[i] = c < kMaxContact ? c : kMaxContact;
  }
}

// velocity of each plane and of its particles at the step's time
// (UserAcc's BC loop and the contact-mesh branch, Source/WeldForm.cpp:67-119)
WF_KERNEL void k_rigid_bc(Dev d, BcTable bc) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= d.nrig) return;
  const int m = d.rmesh[r];
  const int zone = d.mesh_zone[m];
  bool set = false, wset = false;
  double v[3] = {0, 0, 0}, w[3] = {0, 0, 0};
  for (int b = 0; b < bc.n; ++b) {
    if (bc.zone[b] != zone || bc.type[b] != 0) continue;
    if (bc.value_type[b] == 1) {
      if (bc.amp[b] < 0) continue;
      const double f = bc.amp_factor[b] * amp_value(d.amps[bc.amp[b]], d.flags->time);
      for (int k = 0; k < 3; ++k) v[k] = f * bc.v[b][k];
    } else {
      for (int k = 0; k < 3; ++k) v[k] = bc.v[b][k];
      // SetRotAxisVel: constant BCs only (Source/WeldForm.cpp:97-101)
      for (int k = 0; k < 3; ++k) w[k] = bc.w[b][k];
      wset = true;
    }
    set = true;
  }
  if (!set) return;
  d.rv[r] = make_double4(v[0], v[1], v[2], 0.0);
  if (r == d.mesh_elem0[m]) {
    d.mv[m] = make_double4(v[0], v[1], v[2], 0.0);
    if (wset) d.mw[m] = make_double4(w[0], w[1], w[2], 0.0);
  }
}

// Wang contact force of each particle over its partners, in order
WF_KERNEL void k_contact(Dev d, int cur, double dt) {
  if (dt < 0) dt = d.flags->dt;
  __shared__ double sh[kBlock / 64];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double cfm[kMaxMesh][3];
  for (int m = 0; m < kMaxMesh; ++m) cfm[m][0] = cfm[m][1] = cfm[m][2] = 0.0;
  double cfsum = 0.0, creact = 0.0, extw = 0.0, fts = INFINITY;
  if (i < d.n) {
    double cf[3] = {0, 0, 0};
    double qf = 0.0, fh = 0.0, cs = 0.0, qc = 0.0;
    double cfi[kMaxMesh][3];  // this particle's contact force per mesh
    for (int m = 0; m < kMaxMesh; ++m) cfi[m][0] = cfi[m][1] = cfi[m][2] = 0.0;
    const int cnt = d.ccnt[i];
    if (cnt > 0) {
      const double4 x4 = d.xh[cur][i];
      const double4 v4 = d.vm[i];
      double4 a4 = d.acc[i];
      const double x[3] = {x4.x, x4.y, x4.z}, v[3] = {v4.x, v4.y, v4.z};
      double a[3] = {a4.x, a4.y, a4.z};
      const double h = x4.w, m1 = v4.w;
      const DevMat &Mi = d.mats[d.mat[i]];
      double fr_sta = Mi.fr_sta >= 0.0 ? Mi.fr_sta : d.fr_sta;
      double fr_dyn = Mi.fr_sta >= 0.0 ? Mi.fr_dyn : d.fr_dyn;
      // friction_function Linear (Source/Contact_Wang.cpp:241-244)
      if (d.fric_linear) fr_sta = fr_dyn = d.fr_m * d.T[i] + d.fr_b;
      for (int q = 0; q < cnt; ++q) {
        const int r = d.cidx[(long long)q * d.npad + i];
        const int mi = d.rmesh[r];
        const double4 en = d.enorm[r];
        const double n2[3] = {en.x, en.y, en.z};
        const double4 v24 = d.rv[r];
        const double v2[3] = {v24.x, v24.y, v24.z};
        const double delta_ = -(n2[0] * (v[0] - v2[0]) + n2[1] * (v[1] - v2[1]) + n2[2] * (v[2] - v2[2]));
        double xp[3];
        for (int k = 0; k < 3; ++k) xp[k] = x[k] + v[k] * dt + a[k] * dt * dt / 2.0;
        const double dist = (n2[0] * xp[0] + n2[1] * xp[1] + n2[2] * xp[2]) - en.w;
        if (!(dist < h)) continue;
        double qj[3];
        for (int k = 0; k < 3; ++k) qj[k] = x[k] - dist * n2[k];
        const int4 el = d.elem[r];
        const int nd[3] = {el.x, el.y, el.z};
        bool inside = true;
        for (int ii = 0; ii < 3 && inside; ++ii) {
          const int jj = ii + 1 > 2 ? 0 : ii + 1;
          const double4 ni = d.node[nd[ii]], nj = d.node[nd[jj]];
          const double ea[3] = {nj.x - ni.x, nj.y - ni.y, nj.z - ni.z};
          const double eb[3] = {qj[0] - ni.x, qj[1] - ni.y, qj[2] - ni.z};
          double c[3];
          cross_d(ea, eb, c);
          if (c[0] * n2[0] + c[1] * n2[1] + c[2] * n2[2] < 0.0) inside = false;
        }
        if (!inside) continue;
        const double delta = h - dist;
        const double kij = 2.0 * m1 / (dt * dt);
        const double omega = sqrt(kij / m1);
        const double psi = 2. * m1 * omega * d.dfac;
        for (int k = 0; k < 3; ++k) cf[k] = (kij * delta - psi * delta_) * n2[k];
        {
          // dt_fext on the normal force (Source/Contact_Wang.cpp:226-233);
          // Domain::min_force_ts takes the least below 1000
          const double nv = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
          const double nc = sqrt(cf[0] * cf[0] + cf[1] * cf[1] + cf[2] * cf[2]);
          const double dt_fext = 1. * (m1 * 2. * nv / nc);
          if (dt_fext < 1000. && dt_fext < fts) fts = dt_fext;
        }
        for (int k = 0; k < 3; ++k) a[k] += cf[k] / m1;
        // Fraser 3-119: the contact area of the particle, formed whenever a
        // contact heat branch (the only readers, which divide by it) is on
        const double dens = d.rho[cur][i];
        const double dS2 = (d.heat_cond || d.heat_fric) ? pow(m1 / dens, 2.0 / 3.0) : 0.0;
        if (fr_sta > 0.) {
          double du[3], tg[3];
          for (int k = 0; k < 3; ++k) du[k] = xp[k] - x[k] - v2[k] * dt;
          const double dn = du[0] * n2[0] + du[1] * n2[1] + du[2] * n2[2];
          for (int k = 0; k < 3; ++k) tg[k] = kij * (du[k] - dn * n2[k]);
          const double ntg = sqrt(tg[0] * tg[0] + tg[1] * tg[1] + tg[2] * tg[2]);
          const double ncf = sqrt(cf[0] * cf[0] + cf[1] * cf[1] + cf[2] * cf[2]);
          if (ntg < fr_sta * ncf) {
            for (int k = 0; k < 3; ++k) {
              cf[k] -= tg[k] / m1;
              a[k] -= tg[k] / m1;
            }
          } else {
            double td[3];
            for (int k = 0; k < 3; ++k) {
              td[k] = fr_dyn * ncf * tg[k] / ntg;
              cf[k] -= td[k];
              a[k] -= td[k] / m1;
            }
            if (d.heat_fric) {
              // friction work as heat (Source/Contact_Wang.cpp:291-300)
              const double afv = fabs(td[0] * (v[0] - v2[0]) + td[1] * (v[1] - v2[1]) + td[2] * (v[2] - v2[2]));
              qf = afv * dens / m1;
              cs = sqrt(td[0] * td[0] + td[1] * td[1] + td[2] * td[2]) / dS2;
              fh = afv / dS2;
            }
          }
        }
        cfsum += sqrt(cf[0] * cf[0] + cf[1] * cf[1] + cf[2] * cf[2]);
        for (int m = 0; m < kMaxMesh; ++m)
          if (m == mi)
            for (int k = 0; k < 3; ++k) {
              cfm[m][k] += cf[k];
              cfi[m][k] += cf[k];
            }
        creact += (a[0] * n2[0] + a[1] * n2[1] + a[2] * n2[2]) * m1;
        extw += cf[0] * v2[0] + cf[1] * v2[1] + cf[2] * v2[2];
        // contact conductance, Fraser 3.121 (Source/Contact_Wang.cpp:320-324)
        if (d.heat_cond) qc = dens * d.hc * dS2 * (d.rT[r] - d.T[i]) / m1;
      }
      d.acc[i] = make_double4(a[0], a[1], a[2], a4.w);
    }
    d.cforce[i] = make_double4(cf[0], cf[1], cf[2], 0.0);
    if (d.gcf && cnt > 0)
      for (int m = 0; m < d.nmesh; ++m) d.gcf[(long long)m * d.npad + i] = make_double4(cfi[m][0], cfi[m][1], cfi[m][2], 0.0);
    if (d.qfric) {
      d.qfric[i] = qf;
      d.fhfl[i] = fh;
      d.cshear[i] = cs;
      d.qcont[i] = qc;
    }
  }
  for (int m = 0; m < d.nmesh; ++m)
    for (int k = 0; k < 3; ++k) {
      const double r = block_reduce<0>(cfm[m][k], sh);
      if (threadIdx.x == 0) part(d, P_CFORCE + 3 * m + k, blockIdx.x) = r;
    }
  const double r1 = block_reduce<0>(cfsum, sh);
  const double r2 = block_reduce<0>(creact, sh);
  const double r3 = block_reduce<0>(extw, sh);
  const double r4 = block_reduce<2>(fts, sh);
  if (threadIdx.x == 0) {
    part(d, P_FTS, blockIdx.x) = r4;
    part(d, P_CFSUM, blockIdx.x) = r1;
    part(d, P_CREACT, blockIdx.x) = r2;
    // the work takes the step size of the contact evaluation, before any
    // adaptive update of the step (Source/Contact_Wang.cpp:344)
    part(d, P_EXTW, blockIdx.x) = r3 * dt;
  }
}

// Fraser's (LAW 0, Source/Contact.cpp:223-554), Seo's (LAW 2,
// Source/Contact.cpp:562-770) and LS-Dyna's (LAW 4, Source/Contact_LS.cpp:
// 15-292) contact laws, per particle over its partners in index order.  Under
// Leapfrog none of them adds its force to a: contforce (assigned per pair,
// friction terms added or assigned after it) and the sums are outputs; only
// Fraser's Fr_Bound friction changes a, removing its tangential part.
//   Fraser: pairs approaching (delta_ = -n.v_r > 0) whose time to contact
//     t_c = (h + pplane - n.x) / (-delta_) is below dt, contact point
//     x + v t_c - h n; penetration (dt - t_c) delta_; stiffness
//     PFAC cont_stiff; a sliding pair (|v_t| > 1e-3) assigns
//     contforce = -friction_dyn |contforce| t / m.  The static branch reads
//     a force the reference never sets before it and adds nothing here.
//   Seo: dist = n.x - pplane < h, contact point x - dist n; stiffness
//     PFAC cont_stiff; static friction when |m a_t| < friction_sta
//     |contforce|: contforce -= PFAC cont_stiff (v_r dt)_t / m.
//   LS-Dyna: Wang's geometry (predicted position), stiffness
//     2 m / dt^2 PFAC; friction: the trial force -PFAC m v_r / dt returned
//     to the cone friction_sta |contforce| (the reference's old tangential
//     force is never set: zero) and added; contact_reaction_sum and the
//     external work as Wang's; the contact heat fields are reset.
// dt_fext = 2 m |v| / |contforce| per pair: Domain::min_force_ts takes the
// least positive one (Fraser, Seo) or the least (LS-Dyna) below 1000.
template <int LAW>
WF_KERNEL void k_contact_law(Dev d, int cur, double dt) {
  if (dt < 0) dt = d.flags->dt;
  __shared__ double sh[kBlock / 64];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double cfsum = 0.0, creact = 0.0, extw = 0.0, fts = INFINITY;
  if (i < d.n) {
    double cf[3] = {0, 0, 0};
    const int cnt = d.ccnt[i];
    if (cnt > 0) {
      const double4 x4 = d.xh[cur][i];
      const double4 v4 = d.vm[i];
      const double4 a4 = d.acc[i];
      const double x[3] = {x4.x, x4.y, x4.z}, v[3] = {v4.x, v4.y, v4.z};
      double a[3] = {a4.x, a4.y, a4.z};
      const double h = x4.w, m1 = v4.w;
      const DevMat &Mi = d.mats[d.mat[i]];
      double fr_sta = Mi.fr_sta >= 0.0 ? Mi.fr_sta : d.fr_sta;
      double fr_dyn = Mi.fr_sta >= 0.0 ? Mi.fr_dyn : d.fr_dyn;
      // friction_function Linear (Source/Contact_LS.cpp:217-220)
      if (LAW == 4 && d.fric_linear) fr_sta = fr_dyn = d.fr_m * d.T[i] + d.fr_b;
      const double ks = (LAW == 0 || LAW == 2) ? d.cstiff[i] : 0.0;
      const double nv = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      for (int q = 0; q < cnt; ++q) {
        const int r = d.cidx[(long long)q * d.npad + i];
        const double4 en = d.enorm[r];
        const double n2[3] = {en.x, en.y, en.z};
        const double4 v24 = d.rv[r];
        const double v2[3] = {v24.x, v24.y, v24.z};
        const double vr[3] = {v[0] - v2[0], v[1] - v2[1], v[2] - v2[2]};
        const double delta_ = -(n2[0] * vr[0] + n2[1] * vr[1] + n2[2] * vr[2]);
        double qj[3], delta;
        if (LAW == 0) {
          if (!(delta_ > 0)) continue;
          const double tc = (h + en.w - (n2[0] * x[0] + n2[1] * x[1] + n2[2] * x[2])) / (-delta_);
          if (!(tc < dt)) continue;
          for (int k = 0; k < 3; ++k) qj[k] = x[k] + (v[k] * tc) - (h * n2[k]);
          delta = (dt - tc) * delta_;
        } else {
          double xp[3];
          for (int k = 0; k < 3; ++k) xp[k] = LAW == 4 ? x[k] + v[k] * dt + a[k] * dt * dt / 2.0 : x[k];
          const double dist = (n2[0] * xp[0] + n2[1] * xp[1] + n2[2] * xp[2]) - en.w;
          if (!(dist < h)) continue;
          for (int k = 0; k < 3; ++k) qj[k] = x[k] - dist * n2[k];
          delta = h - dist;
        }
        const int4 el = d.elem[r];
        const int nd[3] = {el.x, el.y, el.z};
        bool inside = true;
        for (int ii = 0; ii < 3 && inside; ++ii) {
          const int jj = ii + 1 > 2 ? 0 : ii + 1;
          const double4 ni = d.node[nd[ii]], nj = d.node[nd[jj]];
          const double ea[3] = {nj.x - ni.x, nj.y - ni.y, nj.z - ni.z};
          const double eb[3] = {qj[0] - ni.x, qj[1] - ni.y, qj[2] - ni.z};
          double c[3];
          cross_d(ea, eb, c);
          if (c[0] * n2[0] + c[1] * n2[1] + c[2] * n2[2] < 0.0) inside = false;
        }
        if (!inside) continue;
        const double kij = LAW == 4 ? 2.0 * m1 / (dt * dt) * d.pfac : d.pfac * ks;
        const double omega = sqrt(kij / m1);
        const double psi = 2. * m1 * omega * d.dfac;
        const double ncf0 = kij * delta - psi * delta_;
        for (int k = 0; k < 3; ++k) cf[k] = ncf0 * n2[k];
        const double nc = sqrt(cf[0] * cf[0] + cf[1] * cf[1] + cf[2] * cf[2]);
        if (LAW != 4) cfsum += nc;
        const double dt_fext = 1. * (m1 * 2. * nv / nc);
        if (dt_fext < 1000. && dt_fext < fts && (LAW == 4 || dt_fext > 0)) fts = dt_fext;
        const double an = a[0] * n2[0] + a[1] * n2[1] + a[2] * n2[2];
        const double atg[3] = {a[0] - an * n2[0], a[1] - an * n2[1], a[2] - an * n2[2]};
        if (LAW == 0) {
          if (d.fric_bound) {
            for (int k = 0; k < 3; ++k) a[k] -= atg[k];
          } else if (fr_dyn > 0.) {
            double tg[3];
            for (int k = 0; k < 3; ++k) tg[k] = vr[k] + delta_ * n2[k];
            const double ntg = sqrt(tg[0] * tg[0] + tg[1] * tg[1] + tg[2] * tg[2]);
            if (ntg > 1.e-3)
              for (int k = 0; k < 3; ++k) cf[k] = -(fr_dyn * nc * (tg[k] / ntg)) / m1;
          }
        } else if (LAW == 2) {
          double du[3], tg[3], rt[3];
          for (int k = 0; k < 3; ++k) du[k] = vr[k] * dt;
          const double dn = du[0] * n2[0] + du[1] * n2[1] + du[2] * n2[2];
          for (int k = 0; k < 3; ++k) tg[k] = kij * (du[k] - dn * n2[k]);
          for (int k = 0; k < 3; ++k) rt[k] = atg[k] * m1;
          if (fr_sta > 0. && sqrt(rt[0] * rt[0] + rt[1] * rt[1] + rt[2] * rt[2]) < fr_sta * nc)
            for (int k = 0; k < 3; ++k) cf[k] -= tg[k] / m1;
        } else {
          if (fr_sta > 0.) {
            double fs[3];
            for (int k = 0; k < 3; ++k) fs[k] = 0.0 - d.pfac * m1 * vr[k] / dt;
            const double fy = fr_sta * nc;
            const double nfs = sqrt(fs[0] * fs[0] + fs[1] * fs[1] + fs[2] * fs[2]);
            for (int k = 0; k < 3; ++k) cf[k] += nfs > fy ? fy * fs[k] / nfs : fs[k];
          }
          cfsum += sqrt(cf[0] * cf[0] + cf[1] * cf[1] + cf[2] * cf[2]);
          creact += an * m1;
          extw += cf[0] * v2[0] + cf[1] * v2[1] + cf[2] * v2[2];
        }
      }
      if (LAW == 0 && d.fric_bound) d.acc[i] = make_double4(a[0], a[1], a[2], a4.w);
    }
    d.cforce[i] = make_double4(cf[0], cf[1], cf[2], 0.0);
    // the laws leave m_contact_force alone: no per-group contact sums
    if (d.gcf && cnt > 0)
      for (int m = 0; m < d.nmesh; ++m) d.gcf[(long long)m * d.npad + i] = make_double4(0.0, 0.0, 0.0, 0.0);
    if (d.qfric) {
      d.qfric[i] = 0.0;
      d.fhfl[i] = 0.0;
      d.cshear[i] = 0.0;
      d.qcont[i] = 0.0;
    }
  }
  const double r1 = block_reduce<0>(cfsum, sh);
  const double r2 = block_reduce<0>(creact, sh);
  const double r3 = block_reduce<0>(extw, sh);
  const double r4 = block_reduce<2>(fts, sh);
  if (threadIdx.x == 0) {
    for (int m = 0; m < d.nmesh; ++m)
      for (int k = 0; k < 3; ++k) part(d, P_CFORCE + 3 * m + k, blockIdx.x) = 0.0;
    part(d, P_FTS, blockIdx.x) = r4;
    part(d, P_CFSUM, blockIdx.x) = r1;
    part(d, P_CREACT, blockIdx.x) = r2;
    part(d, P_EXTW, blockIdx.x) = r3 * dt;
  }
}

// Domain::CalcForce2233 (Source/Interaction.cpp:25-333) as
// LastComputeAcceleration applies it to every pair (Source/Domain.cpp:
// 1700-1709), as one FP64 gather per particle over its row: the pair's
// terms for particle i are the reference's for P1 with (P1, P2) = (i, j)
// (each term is odd or even under the swap, so the second particle's terms
// are the same formula from its side).  Per neighbour: a += m_j g.(sigma_i/
// rho_i^2 + sigma_j/rho_j^2 + Pi), dDensity from (m_j/rho_j) v_ij.g (times
// rho_i), StrainRate / RotationRate with the float weight (float)(m_j/rho_j)
// (`float mj_dj`, Source/Interaction.cpp:262), ZWab += (float)(m_j/rho_j) W,
// VXSPH += XSPH m_j / ((rho_i + rho_j)/2) W (v_j - v_i) and, when shepard,
// SumDen += m_j W; g = GK x_ij or (GK L_i) x_ij under the gradient
// correction.  GradientType 0, free-free pairs, no tensile instability.
template <bool GC>
__global__ __launch_bounds__(kBlock) void k_force2233(Dev d, int cur, double xsph, int shepard) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= d.n) return;
  using S = Src<false>;
  const long long np = d.npad;
  double a[3] = {0, 0, 0}, s = 0, sr[6] = {0, 0, 0, 0, 0, 0}, w[3] = {0, 0, 0}, zw = 0, sd = 0;
  double vx[3] = {0, 0, 0};
  const double rhoi = d.rho[cur][i];
  if (d.group[i] >= 0) {
    double Li[6];
    own_corr<GC>(d, i, Li);
    const DevMat &M = d.mats[d.mat[i]];
    const double alpha = M.alpha, beta = M.beta, cs = M.Cs;
    const bool av = (alpha != 0.0 || beta != 0.0);
    const double4 xi = S::x(d, cur, i), vi = S::v(d, i);
    double ti[6], rho_i;
    S::srec(d, i, ti, rho_i);
    const KH<double> ki = kern_h_t<double>(xi.w, d.dim);
    const HbmJ<false> g{d, cur, i};
    for_each_nb(g, d.nnb[i], 0, 1, [&](int j) {
      const double4 xj = g.x(j), vj = g.v(j);
      double tj[6], rho_j;
      g.srec(j, tj, rho_j);
      const KH<double> kh = S::kh(ki, xi, xj);
      const double dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
      const double r2 = dx * dx + dy * dy + dz * dz;
      const double GK = gk_t<double>(r2, kh);
      const double q = r2 > 0.0 ? sqrt(r2) * kh.ih : 0.0;
      // W's constant C (Source/Functions.cpp:32): 1/(pi h^3), 10/(7 pi h^2) in 2D
      const double K = w_shape<double>(q) * (d.dim == 2 ? kh.ih * kh.ih * (10.0 / (7.0 * M_PI))
                                                        : kh.ih * kh.ih * kh.ih * (1.0 / M_PI));
      const double vdx = vi.x - vj.x, vdy = vi.y - vj.y, vdz = vi.z - vj.z;
      const double vdot = vdx * dx + vdy * dy + vdz * dz;
      double pi = 0.0;
      if (av && vdot < 0.0) {
        const double mu = kh.h * vdot / (r2 + 0.01 * kh.h * kh.h);
        pi = (alpha * cs * mu + beta * mu * mu) / (0.5 * (rho_i + rho_j));
      }
      double ex, ey, ez;
      corr_dir<GC>(Li, dx, dy, dz, ex, ey, ez);
      const double gx = GK * ex, gy = GK * ey, gz = GK * ez;
      const double Mxx = ti[XX] + tj[XX] + pi, Myy = ti[YY] + tj[YY] + pi, Mzz = ti[ZZ] + tj[ZZ] + pi;
      const double Mxy = ti[XY] + tj[XY], Mxz = ti[XZ] + tj[XZ], Myz = ti[YZ] + tj[YZ];
      const double mj = vj.w;
      a[0] += mj * (gx * Mxx + gy * Mxy + gz * Mxz);
      a[1] += mj * (gx * Mxy + gy * Myy + gz * Myz);
      a[2] += mj * (gx * Mxz + gy * Myz + gz * Mzz);
      const double volj = S::vol(d, cur, j, xj);
      s += volj * (vdx * gx + vdy * gy + vdz * gz);
      const double fw = (double)(float)volj;
      const double c = fw * (-0.5 * GK);
      sr[XX] += c * (2.0 * vdx * ex);
      sr[YY] += c * (2.0 * vdy * ey);
      sr[ZZ] += c * (2.0 * vdz * ez);
      sr[XY] += c * (vdx * ey + vdy * ex);
      sr[XZ] += c * (vdx * ez + vdz * ex);
      sr[YZ] += c * (vdy * ez + vdz * ey);
      w[0] += c * (vdx * ey - vdy * ex);
      w[1] += c * (vdx * ez - vdz * ex);
      w[2] += c * (vdy * ez - vdz * ey);
      // rows are padded with the particle itself: its pair terms above are
      // exact zeros, the kernel sums must skip it
      const bool real = j != i;
      zw += real ? fw * K : 0.0;
      const double xf = xsph * mj / (0.5 * (rhoi + d.rho[cur][j])) * K;
      vx[0] += xf * -vdx;
      vx[1] += xf * -vdy;
      vx[2] += xf * -vdz;
      if (shepard && real) sd += mj * K;
    });
  }
  if (d.dim == 2 && !GC) a[2] = 0.0;  // plane strain (Source/Interaction.cpp:237)
  const double4 a4 = d.acc[i];
  d.acc[i] = make_double4(a[0], a[1], a[2], a4.w);
  d.drho[i] = s * rhoi;
  for (int k = 0; k < 6; ++k) d.sr[k * np + i] = sr[k];
  for (int k = 0; k < 3; ++k) d.rr[k * np + i] = w[k];
  d.zwab[i] = zw;
  d.vxsph[i] = make_double4(vx[0], vx[1], vx[2], 0.0);
  d.sumden[i] = sd;
}

void launch_force2233(const Dev &d, hipStream_t st, int cur, double xsph, int shepard) {
  if (d.gc_on)
    k_force2233<true><<<d.nblocks, kBlock, 0, st>>>(d, cur, xsph, shepard);
  else
    k_force2233<false><<<d.nblocks, kBlock, 0, st>>>(d, cur, xsph, shepard);
}

// Per-group sums (wf_set_group_sums): one block per group walks the group's
// particles in internal order (thread t takes every kGroupThreads-th) and
// reduces in a fixed order, so the sums are deterministic.  what 0: the
// reaction of zone gzone, sum of m a over the group's particles carrying
// that id, read where the zone reactions are (after the acceleration pass,
// before contact and the BCs); what 1: the contact force each rigid mesh
// exerts on the group (after contact).
constexpr int kGroupThreads = 256;
__global__ __launch_bounds__(kGroupThreads) void k_group_sums(Dev d, int what) {
  __shared__ double sh[kGroupThreads / 64];
  const int g = blockIdx.x;
  const long long slot = (d.flags->step % d.ghist_cap) * d.ngroups + g;
  const int b = d.goff[g], e = d.goff[g + 1];
  if (what == 0) {
    double s[3] = {0, 0, 0};
    for (int p = b + threadIdx.x; p < e; p += kGroupThreads) {
      const int i = d.gidx[p];
      if (d.id[i] != d.gzone) continue;
      const double4 a = d.acc[i];
      const double m = d.vm[i].w;
      s[0] += m * a.x, s[1] += m * a.y, s[2] += m * a.z;
    }
    for (int k = 0; k < 3; ++k) {
      const double r = block_reduce<0>(s[k], sh);
      if (threadIdx.x == 0) d.gr_hist[slot * 3 + k] = r;
    }
  } else {
    for (int m = 0; m < d.nmesh; ++m) {
      double s[3] = {0, 0, 0};
      for (int p = b + threadIdx.x; p < e; p += kGroupThreads) {
        const int i = d.gidx[p];
        if (d.ccnt[i] <= 0) continue;
        const double4 f = d.gcf[(long long)m * d.npad + i];
        s[0] += f.x, s[1] += f.y, s[2] += f.z;
      }
      for (int k = 0; k < 3; ++k) {
        const double r = block_reduce<0>(s[k], sh);
        if (threadIdx.x == 0) d.gc_hist[(slot * kMaxMesh + m) * 3 + k] = r;
      }
    }
  }
}

void launch_group_sums(const Dev &d, hipStream_t st, int what) {
  if (d.ngroups > 0) k_group_sums<<<d.ngroups, kGroupThreads, 0, st>>>(d, what);
}

// the drift of the rigid particles (the reference's loop covers them)
WF_KERNEL void k_rigid_drift(Dev d, double dt) {
  if (dt < 0) dt = d.flags->dt;
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= d.nrig) return;
  double4 x = d.rx[r];
  const double4 v = d.rv[r];
  x.x += v.x * dt;
  x.y += v.y * dt;
  x.z += v.z * dt;
  d.rx[r] = x;
}

// one block per mesh (TriMesh::Update, Source/Mesh.cpp:273-296): node
// velocity m_v + m_w x node at the node's position before it moves, nodes
// move with it, element normals and plane coefficients are recomputed from
// the winding, particles take the mean node velocity
__global__ __launch_bounds__(1024) void k_mesh_update(Dev d, double dt) {
  if (dt < 0) dt = d.flags->dt;
  const int m = blockIdx.x;
  const double4 mv = d.mv[m], mw = d.mw[m];
  const int n0 = d.mesh_node0[m], nn = d.mesh_nnode[m];
  for (int k = threadIdx.x; k < nn; k += blockDim.x) {
    double4 p = d.node[n0 + k];
    const double wp[3] = {mw.y * p.z - mw.z * p.y, mw.z * p.x - mw.x * p.z, mw.x * p.y - mw.y * p.x};
    const double4 nv = make_double4(mv.x + wp[0], mv.y + wp[1], mv.z + wp[2], 0.0);
    p.x += nv.x * dt;
    p.y += nv.y * dt;
    p.z += nv.z * dt;
    d.node[n0 + k] = p;
    d.nodev[n0 + k] = nv;
  }
  __syncthreads();
  const int e0 = d.mesh_elem0[m], ne = d.mesh_nelem[m];
  for (int e = threadIdx.x; e < ne; e += blockDim.x) {
    const int4 el = d.elem[e0 + e];
    const double4 a = d.node[el.x], b = d.node[el.y], c = d.node[el.z];
    const double u[3] = {b.x - a.x, b.y - a.y, b.z - a.z}, v[3] = {c.x - a.x, c.y - a.y, c.z - a.z};
    double w[3];
    cross_d(u, v, w);
    const double nw = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    const double n[3] = {w[0] / nw, w[1] / nw, w[2] / nw};
    const int fa = el.w == 0 ? el.x : (el.w == 1 ? el.y : el.z);
    const double4 pf = d.node[fa];
    const double pp = pf.x * n[0] + pf.y * n[1] + pf.z * n[2];
    d.enorm[e0 + e] = make_double4(n[0], n[1], n[2], pp);
    const double4 va = d.nodev[el.x], vb = d.nodev[el.y], vc = d.nodev[el.z];
    const double vm = ((0.0 + va.x) + vb.x) + vc.x;
    const double vn = ((0.0 + va.y) + vb.y) + vc.y;
    const double vo = ((0.0 + va.z) + vb.z) + vc.z;
    d.rv[e0 + e] = make_double4(vm / 3, vn / 3, vo / 3, 0.0);
  }
}

void launch_surface(const Dev &d, hipStream_t st, int cur, int setup) {
  k_surface<<<nblk(d.n), kBlock, 0, st>>>(d, cur, setup);
}
void launch_contact(const Dev &d, hipStream_t st, const BcTable &bc, int cur, double dt, int law) {
  k_rigid_bc<<<(d.nrig + kBlock - 1) / kBlock, kBlock, 0, st>>>(d, bc);
  switch (law) {
    case 0: k_contact_law<0><<<d.nblocks, kBlock, 0, st>>>(d, cur, dt); break;
    case 2: k_contact_law<2><<<d.nblocks, kBlock, 0, st>>>(d, cur, dt); break;
    case 4: k_contact_law<4><<<d.nblocks, kBlock, 0, st>>>(d, cur, dt); break;
    default: k_contact<<<d.nblocks, kBlock, 0, st>>>(d, cur, dt);
  }
}
void launch_rigid_drift(const Dev &d, hipStream_t st, double dt) {
  k_rigid_drift<<<(d.nrig + kBlock - 1) / kBlock, kBlock, 0, st>>>(d, dt);
}
void launch_mesh_update(const Dev &d, hipStream_t st, double dt) { k_mesh_update<<<d.nmesh, 1024, 0, st>>>(d, dt); }

// ---------------------------------------------------------------------------
// utilities

WF_KERNEL void k_max_pl(Dev d) {
  __shared__ double sh[kBlock / 64];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const double v = i < d.n ? d.pl[i] : 0.0;
  const double r = block_reduce<1>(v, sh);
  if (threadIdx.x == 0) part(d, P_MAXT, blockIdx.x) = r;
}

__global__ __launch_bounds__(256) void k_max_pl_final(Dev d) {
  __shared__ double sh[16];
  double v = 0.0;
  for (int b = threadIdx.x; b < d.nblocks; b += blockDim.x) v = fmax(v, part(d, P_MAXT, b));
  const double r = block_reduce<1>(v, sh);
  if (threadIdx.x == 0) d.scalars[S_MAXPL] = r;
}

// rebuild every FP32 shadow from the FP64 state (set-up and host writes)
WF_KERNEL void k_shadows(Dev d, int cur) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= d.n) return;
  const double *o = d.mats[d.mat[i]].xc;
  for (int b = 0; b < 2; ++b) {
    const double4 x = d.xh[b][i];
    d.xhf[b][i] = make_float4((float)(x.x - o[0]), (float)(x.y - o[1]), (float)(x.z - o[2]), (float)d.vol[b][i]);
  }
  const double4 v = d.vm[i];
  d.vmf[i] = make_float4((float)v.x, (float)v.y, (float)v.z, (float)v.w);
  const StressRec r = d.srec[i];
  StressRecF f;
  for (int k = 0; k < 6; ++k) f.t[k] = (float)(r.s[k] * r.irho2);
  f.rho = (float)r.rho;
  f.pad = 0.0f;
  d.srecf[i] = f;
  const DevMat &M = d.mats[d.mat[i]];
  const double rho = d.rho[cur][i];
  d.prf[i] = (float)((M.P0 + M.Cs * M.Cs * (rho - M.ref_density)) / (rho * rho));
}

void launch_shadows(const Dev &d, hipStream_t st, int cur) {
  if (d.mixed) k_shadows<<<nblk(d.n), kBlock, 0, st>>>(d, cur);
}

void launch_max_pl(const Dev &d, hipStream_t st) {
  k_max_pl<<<d.nblocks, kBlock, 0, st>>>(d);
  k_max_pl_final<<<1, 256, 0, st>>>(d);
}

}  // namespace wf
