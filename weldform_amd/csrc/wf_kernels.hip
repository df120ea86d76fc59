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
#include <math.h>

#include "wf_internal.h"

namespace wf {

#define WF_KERNEL __global__ __launch_bounds__(kBlock)

static inline int nblk(long long n) { return (int)((n + kBlock - 1) / kBlock); }

// ---------------------------------------------------------------------------
// reductions

__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ inline double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ inline double wave_min(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  return v;
}

// block reduction in a fixed order; result valid in thread 0
template <int OP>
__device__ inline double block_reduce(double v, double *sh) {
  v = OP == 0 ? wave_sum(v) : (OP == 1 ? wave_max(v) : wave_min(v));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    r = sh[0];
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k)
      r = OP == 0 ? r + sh[k] : (OP == 1 ? fmax(r, sh[k]) : fmin(r, sh[k]));
  }
  __syncthreads();
  return r;
}

// ---------------------------------------------------------------------------
// kernel functions (Source/Functions.cpp:74-87) and the neighbour predicate

// cubic spline gradient factor: GK * x_ij = grad_i W_ij (3D)
__device__ inline double grad_kernel(double q, double h) {
  const double C = 1.0 / (h * h * h * h * M_PI);
  if (q == 0.0) return C / h * (-3.0);
  if (q < 1.0) return C / (q * h) * (-3.0 * q + 2.25 * q * q);
  if (q < 2.0) return C / (q * h) * (-0.75 * (2.0 - q) * (2.0 - q));
  return 0.0;
}

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

WF_KERNEL void k_schedule(Dev d, int isfirst, int force) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  DevFlags *f = d.flags;
  const int mx = f->yielded;
  int search = 0;
  if (force) {
    search = 1;
  } else {
    if (mx && !f->isyielding) {
      search = 1;
      f->isyielding = 1;
      f->cleared = 0;
    } else if ((mx || isfirst) && f->ts_i == 0 && f->cleared) {
      search = 1;
      f->cleared = 0;
    }
  }
  f->step_yield = mx;
  f->do_search = search;
  if (search) {
    f->nb_searches += 1;
    f->max_count = 0;
    f->directed_pairs = 0;
  }
}

// ---------------------------------------------------------------------------
// neighbour build: counting sort of particles by reference cell, then one
// lane per particle scans the 27 surrounding cells with the exact predicate.

WF_KERNEL void k_zero_int(Dev d, int *p, int n) {
  if (!d.flags->do_search) return;
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) p[i] = 0;
}

WF_KERNEL void k_cell_hist(Dev d, int cur) {
  if (!d.flags->do_search) return;
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= d.n) return;
  const double4 x = d.xh[cur][i];
  const int cx = cell_coord(x.x, d.blpf[0], d.cellsize[0], d.cellno[0]);
  const int cy = cell_coord(x.y, d.blpf[1], d.cellsize[1], d.cellno[1]);
  const int cz = cell_coord(x.z, d.blpf[2], d.cellsize[2], d.cellno[2]);
  const int c = (cx * d.cellno[1] + cy) * d.cellno[2] + cz;
  d.cell[i] = c;
  atomicAdd(&d.cell_count[c], 1);
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

// block-wide exclusive scan of 1024 values; returns the block total in *tot
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

__global__ __launch_bounds__(1024) void k_scan1(Dev d) {
  if (!d.flags->do_search) return;
  __shared__ int sh[33];
  const int i = blockIdx.x * 1024 + threadIdx.x;
  const int v = i < d.ncells ? d.cell_count[i] : 0;
  int tot;
  const int e = block_scan_excl(v, sh, &tot);
  if (i < d.ncells) d.cell_start[i] = e;
  if (threadIdx.x == 0) d.scan_tmp[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void k_scan2(Dev d, int nb) {
  if (!d.flags->do_search) return;
  __shared__ int sh[33];
  int carry = 0;
  for (int base = 0; base < nb; base += 1024) {
    const int i = base + threadIdx.x;
    const int v = i < nb ? d.scan_tmp[i] : 0;
    int tot;
    const int e = block_scan_excl(v, sh, &tot);
    if (i < nb) d.scan_tmp[i] = e + carry;
    carry += tot;
  }
  if (threadIdx.x == 0) d.cell_start[d.ncells] = carry;
}

__global__ __launch_bounds__(1024) void k_scan3(Dev d) {
  if (!d.flags->do_search) return;
  const int i = blockIdx.x * 1024 + threadIdx.x;
  if (i < d.ncells) d.cell_start[i] += d.scan_tmp[blockIdx.x];
}

WF_KERNEL void k_scatter(Dev d) {
  if (!d.flags->do_search) return;
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= d.n) return;
  const int c = d.cell[i];
  const int pos = d.cell_start[c] + atomicAdd(&d.cell_cursor[c], 1);
  d.sorted[pos] = i;
}

// cells hold ~10-30 particles: an insertion sort makes the within-cell
// order (and with it every neighbour row) independent of atomic order
WF_KERNEL void k_cell_sort(Dev d) {
  if (!d.flags->do_search) return;
  int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= d.ncells) return;
  const int s = d.cell_start[c], e = d.cell_start[c + 1];
  for (int a = s + 1; a < e; ++a) {
    const int v = d.sorted[a];
    int b = a - 1;
    while (b >= s && d.sorted[b] > v) {
      d.sorted[b + 1] = d.sorted[b];
      --b;
    }
    d.sorted[b + 1] = v;
  }
}

WF_KERNEL void k_build_nbr(Dev d, int cur) {
  if (!d.flags->do_search) return;
  __shared__ double sh[kBlock / 64];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  int cnt = 0;
  if (i < d.n) {
    const int g = d.group[i];
    if (g >= 0) {
      const double4 xi = d.xh[cur][i];
      const int c = d.cell[i];
      const int nz = d.cellno[2], ny = d.cellno[1], nx = d.cellno[0];
      const int cz = c % nz, cy = (c / nz) % ny, cx = c / (nz * ny);
      for (int ax = cx - 1; ax <= cx + 1; ++ax) {
        if (ax < 0 || ax >= nx) continue;
        for (int ay = cy - 1; ay <= cy + 1; ++ay) {
          if (ay < 0 || ay >= ny) continue;
          for (int az = cz - 1; az <= cz + 1; ++az) {
            if (az < 0 || az >= nz) continue;
            const int cc = (ax * ny + ay) * nz + az;
            const int s = d.cell_start[cc], e = d.cell_start[cc + 1];
            for (int p = s; p < e; ++p) {
              const int j = d.sorted[p];
              if (j == i || d.group[j] != g) continue;
              if (!in_radius(xi, d.xh[cur][j], d.cellfac)) continue;
              if (cnt < d.max_nb) d.nbr[(long long)cnt * d.npad + i] = j;
              ++cnt;
            }
          }
        }
      }
    }
    d.nnb[i] = cnt < d.max_nb ? cnt : d.max_nb;
    if (cnt > d.max_nb) d.flags->overflow = 1;
  }
  // deterministic totals via block reduction + integer atomics
  const double tot = block_reduce<0>((double)cnt, sh);
  const double mx = block_reduce<1>((double)cnt, sh);
  if (threadIdx.x == 0) {
    atomicAdd((unsigned long long *)&d.flags->directed_pairs, (unsigned long long)tot);
    atomicMax(&d.flags->max_count, (int)mx);
  }
}

void launch_schedule(const Dev &d, hipStream_t st, int isfirst, int force) {
  k_schedule<<<1, 64, 0, st>>>(d, isfirst, force);
}

void launch_neighbours(const Dev &d, hipStream_t st, int cur) {
  const int nbc = nblk(d.ncells);
  k_zero_int<<<nbc, kBlock, 0, st>>>(d, d.cell_count, d.ncells);
  k_zero_int<<<nbc, kBlock, 0, st>>>(d, d.cell_cursor, d.ncells);
  k_cell_hist<<<nblk(d.n), kBlock, 0, st>>>(d, cur);
  const int ns = (d.ncells + 1023) / 1024;
  k_scan1<<<ns, 1024, 0, st>>>(d);
  k_scan2<<<1, 1024, 0, st>>>(d, ns);
  k_scan3<<<ns, 1024, 0, st>>>(d);
  k_scatter<<<nblk(d.n), kBlock, 0, st>>>(d);
  k_cell_sort<<<nbc, kBlock, 0, st>>>(d);
  k_build_nbr<<<nblk(d.n), kBlock, 0, st>>>(d, cur);
}

// ---------------------------------------------------------------------------
// pass 1: acceleration (CalcAccel + AccelReduction, GradientType 0, TI = 0)

WF_KERNEL void k_accel(Dev d, int cur) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= d.n) return;
  double ax = 0.0, ay = 0.0, az = 0.0;
  if (d.group[i] >= 0) {
    const DevMat &M = d.mats[d.mat[i]];
    const double alpha = M.alpha, beta = M.beta, cs = M.Cs;
    const bool av = (alpha != 0.0 || beta != 0.0);
    const double4 xi = d.xh[cur][i];
    const double4 vi = d.vm[i];
    const StressRec ri = d.srec[i];
    const double sxx = ri.s[XX] * ri.irho2, syy = ri.s[YY] * ri.irho2, szz = ri.s[ZZ] * ri.irho2;
    const double sxy = ri.s[XY] * ri.irho2, sxz = ri.s[XZ] * ri.irho2, syz = ri.s[YZ] * ri.irho2;
    const int cnt = d.nnb[i];
    const int *nb = d.nbr + i;
    for (int k = 0; k < cnt; ++k) {
      const int j = nb[(long long)k * d.npad];
      const double4 xj = d.xh[cur][j];
      const double4 vj = d.vm[j];
      const StressRec rj = d.srec[j];
      const double h = (xi.w + xj.w) / 2;
      const double dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      const double GK = grad_kernel(r / h, h);
      double pi = 0.0;
      if (av) {
        const double vdot = (vi.x - vj.x) * dx + (vi.y - vj.y) * dy + (vi.z - vj.z) * dz;
        if (vdot < 0.0) {
          const double mu = h * vdot / (r * r + 0.01 * h * h);
          pi = (alpha * cs * mu + beta * mu * mu) / (0.5 * (ri.rho + rj.rho));
        }
      }
      const double Mxx = sxx + rj.s[XX] * rj.irho2 + pi;
      const double Myy = syy + rj.s[YY] * rj.irho2 + pi;
      const double Mzz = szz + rj.s[ZZ] * rj.irho2 + pi;
      const double Mxy = sxy + rj.s[XY] * rj.irho2;
      const double Mxz = sxz + rj.s[XZ] * rj.irho2;
      const double Myz = syz + rj.s[YZ] * rj.irho2;
      const double gx = GK * dx, gy = GK * dy, gz = GK * dz;
      const double mj = vj.w;
      ax += mj * (gx * Mxx + gy * Mxy + gz * Mxz);
      ay += mj * (gx * Mxy + gy * Myy + gz * Myz);
      az += mj * (gx * Mxz + gy * Myz + gz * Mzz);
    }
  }
  d.acc[i] = make_double4(ax, ay, az, 0.0);
}

void launch_accel(const Dev &d, hipStream_t st, int cur) {
  k_accel<<<nblk(d.n), kBlock, 0, st>>>(d, cur);
}

// ---------------------------------------------------------------------------
// velocity kicks with the UserAcc BC semantics (Source/WeldForm.cpp:51-92)
// and the reaction observer at the first GeneralAfter of the step.
//   Leapfrog  (Source/SolverLeapfrog.cpp:228-273):
//       [first step] v -= a*0.5*dt ;  v += a*0.5*(dt_prev + dt)
//   KickDrift (Source/SolverKickDrift.cpp:208-231): v += a*0.5*dt_prev

WF_KERNEL void k_kick(Dev d, BcTable bc, int scheme, int isfirst, double dt_prev, double dt) {
  __shared__ double sh[kBlock / 64];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double4 a = make_double4(0, 0, 0, 0);
  double m = 0.0;
  int id = -0x7fffffff;
  if (i < d.n) {
    a = d.acc[i];
    const double4 vm = d.vm[i];
    m = vm.w;
    id = d.id[i];
    double v[3] = {vm.x, vm.y, vm.z};
    double av[3] = {a.x, a.y, a.z};
    if (!isfinite(a.x) || !isfinite(a.y) || !isfinite(a.z)) d.flags->nan = 1;
    bool fixed = false;
    for (int b = 0; b < bc.n; ++b) {
      if (bc.zone[b] != id) continue;
      if (bc.type[b] == 0) {
        fixed = true;
        for (int k = 0; k < 3; ++k) {
          av[k] = 0.0;
          v[k] = bc.v[b][k];
        }
      } else if (bc.type[b] == 2) {
        d.T[i] = bc.T[b];
      }
    }
    if (!fixed) {
      if (scheme == WF_LEAPFROG) {
        if (isfirst)
          for (int k = 0; k < 3; ++k) v[k] -= av[k] * 0.5 * dt;
        for (int k = 0; k < 3; ++k) v[k] += av[k] * 0.5 * (dt_prev + dt);
      } else {
        for (int k = 0; k < 3; ++k) v[k] += av[k] * 0.5 * dt_prev;
      }
    }
    d.vm[i] = make_double4(v[0], v[1], v[2], vm.w);
    d.acc[i] = make_double4(av[0], av[1], av[2], 0.0);
  }
  for (int z = 0; z < bc.nz; ++z) {
    const bool in = (id == bc.rzone[z]);
    const double fx = block_reduce<0>(in ? m * a.x : 0.0, sh);
    const double fy = block_reduce<0>(in ? m * a.y : 0.0, sh);
    const double fz = block_reduce<0>(in ? m * a.z : 0.0, sh);
    if (threadIdx.x == 0) {
      double *p = d.partials + (long long)blockIdx.x * P_COUNT + P_REACT + 3 * z;
      p[0] = fx;
      p[1] = fy;
      p[2] = fz;
    }
  }
}

// KickDrift's second half kick, v += a*dt/2 (Source/SolverKickDrift.cpp:261-265)
WF_KERNEL void k_kick2(Dev d, double dt) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= d.n) return;
  const double4 a = d.acc[i];
  double4 v = d.vm[i];
  v.x += a.x * dt / 2.;
  v.y += a.y * dt / 2.;
  v.z += a.z * dt / 2.;
  d.vm[i] = v;
}

void launch_kick(const Dev &d, hipStream_t st, const BcTable &bc, int scheme, int isfirst, double dt_prev,
                 double dt) {
  k_kick<<<d.nblocks, kBlock, 0, st>>>(d, bc, scheme, isfirst, dt_prev, dt);
}

void launch_kick2(const Dev &d, hipStream_t st, double dt) { k_kick2<<<nblk(d.n), kBlock, 0, st>>>(d, dt); }

// ---------------------------------------------------------------------------
// pass 2: density rate + density update + drift
// (CalcDensInc + DensReduction, Source/InteractionAlt.cpp:483-640;
//  Source/SolverLeapfrog.cpp:280-301).  Reads x_n, v_{n+1/2}, m/rho_n and
// writes rho_{n+1}, m/rho_{n+1}, x_{n+1} into the other buffer.

WF_KERNEL void k_dens_drift(Dev d, int cur, double dt, int do_update) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= d.n) return;
  const double4 xi = d.xh[cur][i];
  const double4 vi = d.vm[i];
  double s = 0.0;
  if (d.group[i] >= 0) {
    const int cnt = d.nnb[i];
    const int *nb = d.nbr + i;
    for (int k = 0; k < cnt; ++k) {
      const int j = nb[(long long)k * d.npad];
      const double4 xj = d.xh[cur][j];
      const double4 vj = d.vm[j];
      const double volj = d.vol[cur][j];
      const double h = (xi.w + xj.w) / 2;
      const double dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      const double GK = grad_kernel(r / h, h);
      const double t = (vi.x - vj.x) * (GK * dx) + (vi.y - vj.y) * (GK * dy) + (vi.z - vj.z) * (GK * dz);
      s += volj * t;
    }
  }
  const double rhoi = d.rho[cur][i];
  const double drho = s * rhoi;
  d.drho[i] = drho;
  if (!do_update) return;
  const int nxt = cur ^ 1;
  const double rn = rhoi + dt * drho;
  d.rho[nxt][i] = rn;
  d.vol[nxt][i] = vi.w / rn;
  const double ux = vi.x * dt, uy = vi.y * dt, uz = vi.z * dt;
  double4 u = d.disp[i];
  u.x += ux;
  u.y += uy;
  u.z += uz;
  d.disp[i] = u;
  d.xh[nxt][i] = make_double4(xi.x + ux, xi.y + uy, xi.z + uz, xi.w);
}

void launch_dens_drift(const Dev &d, hipStream_t st, int cur, double dt, int do_update) {
  k_dens_drift<<<nblk(d.n), kBlock, 0, st>>>(d, cur, dt, do_update);
}

// ---------------------------------------------------------------------------
// pass 3: strain-rate / rotation-rate gather fused with the per-particle
// stress update (CalcRateTensors + RateTensorsReduction,
// Source/InteractionAlt.cpp:263-479; Particle::CalcStressStrain,
// Source/Particle.cpp:960-1067), plastic work heat (Source/Particle.cpp:922-939),
// the temperature update of ThermalCalcs (Source/Thermal.cpp:726-734) and
// the energy terms (Source/Domain.cpp:1752-1829).

__device__ inline double jc_yield(const DevMat &M, double strain, double sr, double T) {
  const double T_h = (T - M.jc_Tt) / (M.jc_Tm - M.jc_Tt);
  double f = 1.0;
  if (sr > M.jc_eps0) f = 1.0 + M.jc_C * log(sr / M.jc_eps0);
  return (M.jc_A + M.jc_B * pow(strain, M.jc_n)) * f * (1.0 - pow(T_h, M.jc_m));
}

__device__ inline double hol_yield(const DevMat &M, double strain) {
  if (strain + M.hol_eps0 > M.hol_eps1) return M.hol_K * pow(strain + M.hol_eps0, M.hol_m);
  return M.hol_sy0;
}

WF_KERNEL void k_rates_stress(Dev d, int buf, double dt, int do_rates, int do_stress) {
  __shared__ double sh[kBlock / 64];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const long long np = d.npad;
  double e_int = 0.0, e_kin = 0.0, e_plw = 0.0, tmax = -INFINITY, tmin = INFINITY;
  if (i < d.n) {
    const int g = d.group[i];
    double SR[6], W[3];  // W = (w01, w02, w12)
    if (do_rates) {
      for (int k = 0; k < 6; ++k) SR[k] = 0.0;
      W[0] = W[1] = W[2] = 0.0;
      if (g >= 0) {
        const DevMat &M = d.mats[d.mat[i]];
        const double4 xi = d.xh[buf][i];
        const double4 vi = d.vm[i];
        const double rhoi = d.rho[buf][i];
        const double Pi = M.P0 + (M.Cs * M.Cs) * (rhoi - M.ref_density);
        const double pri = Pi / (rhoi * rhoi);
        const double mi = vi.w;
        const int cnt = d.nnb[i];
        const int *nb = d.nbr + i;
        for (int k = 0; k < cnt; ++k) {
          const int j = nb[(long long)k * np];
          const double4 xj = d.xh[buf][j];
          const double4 vj = d.vm[j];
          const double volj = d.vol[buf][j];
          const double rhoj = d.rho[buf][j];
          const double h = (xi.w + xj.w) / 2;
          const double dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
          const double r = sqrt(dx * dx + dy * dy + dz * dz);
          const double GK = grad_kernel(r / h, h);
          const double vx = vi.x - vj.x, vy = vi.y - vj.y, vz = vi.z - vj.z;
          const double c = volj * (-0.5 * GK);
          SR[XX] += c * (2.0 * vx * dx);
          SR[YY] += c * (2.0 * vy * dy);
          SR[ZZ] += c * (2.0 * vz * dz);
          SR[XY] += c * (vx * dy + vy * dx);
          SR[XZ] += c * (vx * dz + vz * dx);
          SR[YZ] += c * (vy * dz + vz * dy);
          W[0] += c * (vx * dy - vy * dx);
          W[1] += c * (vx * dz - vz * dx);
          W[2] += c * (vy * dz - vz * dy);
          // kinetic-energy pair term (Source/Domain.cpp:1776-1792): the pair
          // is oriented lower-index first; its sum over particles vanishes
          const double Pj = M.P0 + (M.Cs * M.Cs) * (rhoj - M.ref_density);
          const double msec = i < j ? vj.w : mi;
          const double term = 0.5 * msec * (pri + Pj / (rhoj * rhoj)) * (vx * (GK * dx) + vy * (GK * dy) + vz * (GK * dz));
          e_kin += i < j ? term : -term;
        }
      }
      for (int k = 0; k < 6; ++k) d.sr[k * np + i] = SR[k];
      for (int k = 0; k < 3; ++k) d.rr[k * np + i] = W[k];
    } else {
      for (int k = 0; k < 6; ++k) SR[k] = d.sr[k * np + i];
      for (int k = 0; k < 3; ++k) W[k] = d.rr[k * np + i];
    }
    if (do_stress && g >= 0) {
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
      if (pl > 1.e-6) d.flags->yielded = 1;
      StressRec rec;
      for (int k = 0; k < 6; ++k) rec.s[k] = sig[k];
      rec.rho = rho;
      rec.irho2 = 1.0 / (rho * rho);
      d.srec[i] = rec;
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
  const double r1 = block_reduce<0>(e_kin, sh);
  const double r2 = block_reduce<0>(e_plw, sh);
  const double r3 = block_reduce<1>(tmax, sh);
  const double r4 = block_reduce<2>(tmin, sh);
  if (threadIdx.x == 0) {
    double *p = d.partials + (long long)blockIdx.x * P_COUNT;
    p[P_INT] = r0;
    p[P_KIN] = r1;
    p[P_PLW] = r2;
    p[P_MAXT] = r3;
    p[P_MINT] = r4;
  }
}

void launch_rates_stress(const Dev &d, hipStream_t st, int buf, double dt, int do_rates, int do_stress) {
  k_rates_stress<<<d.nblocks, kBlock, 0, st>>>(d, buf, dt, do_rates, do_stress);
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
  __shared__ double sh[kBlock / 64];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const long long np = d.npad;
  double e_int = 0.0;
  if (i < d.n && d.group[i] >= 0) {
    const DevMat &M = d.mats[d.mat[i]];
    const double4 xi = d.xh[buf][i];
    const double rho = d.rho[buf][i];
    const double m = d.vm[i].w;
    const double Ti = d.T[i];
    const double voli = d.vol[buf][i];
    if (d.thbc[i] == 1) {
      const double dS2 = pow(m / rho, 2.0 / 3.0);
      d.qconv[i] = rho * d.hconv[i] * dS2 * (d.tinf[i] - Ti) / m;
    }
    const double kk = 4. * (M.k_T * M.k_T) / (M.k_T + M.k_T);
    const int ci = d.cell[i], oi = d.perm[i];
    double temp = 0.0;
    const int cnt = d.nnb[i];
    const int *nb = d.nbr + i;
    for (int k = 0; k < cnt; ++k) {
      const int j = nb[(long long)k * np];
      const double4 xj = d.xh[buf][j];
      const double h = (xi.w + xj.w) / 2.0;
      const double dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      const double GK = grad_kernel(r / h, h);
      const bool first = pair_first(ci, d.cell[j], oi, d.perm[j], d.cellno[1], d.cellno[2]);
      const double vsec = first ? d.vol[buf][j] : voli;
      temp += vsec * kk * (Ti - d.T[j]) * (GK * (dx * dx + dy * dy + dz * dz)) / (r * r);
    }
    const double f = 1. / (rho * M.cp_T);
    const double dTdt = f * (temp + d.qconv[i] + d.qsrc[i] + d.qpl[i] * d.pl_frac + 0.0);
    d.dTdt[i] = dTdt;
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
  if (threadIdx.x == 0) d.partials[(long long)blockIdx.x * P_COUNT + P_INT] = r0;
}

void launch_thermal(const Dev &d, hipStream_t st, int buf, double dt) {
  k_thermal<<<d.nblocks, kBlock, 0, st>>>(d, buf, dt);
}

// ---------------------------------------------------------------------------
// end of step: fixed-order reduction of the block partials, scalar updates,
// reaction history, and the end-of-step half of the neighbour schedule.

__global__ __launch_bounds__(256) void k_finalize(Dev d, double dt, long long step, int nz, int full_step) {
  __shared__ double sh[4];
  const int nb = d.nblocks;
  double res[P_COUNT];
  for (int s = 0; s < P_COUNT; ++s) {
    if (s >= P_REACT + 3 * nz) {
      res[s] = 0.0;
      continue;
    }
    const int op = (s == P_MAXT) ? 1 : (s == P_MINT ? 2 : 0);
    double v = op == 0 ? 0.0 : (op == 1 ? -INFINITY : INFINITY);
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
      const double x = d.partials[(long long)b * P_COUNT + s];
      v = op == 0 ? v + x : (op == 1 ? fmax(v, x) : fmin(v, x));
    }
    res[s] = op == 0 ? block_reduce<0>(v, sh) : (op == 1 ? block_reduce<1>(v, sh) : block_reduce<2>(v, sh));
  }
  if (threadIdx.x != 0 || !full_step) return;
  d.scalars[S_KIN] += res[P_KIN] * dt;
  d.scalars[S_INT] += res[P_INT] * dt;
  if (d.pl_heating) d.scalars[S_PLW] += res[P_PLW] * dt;
  if (d.thermal) {
    d.scalars[S_MAXT] = fmax(0.0, res[P_MAXT]);
    d.scalars[S_MINT] = fmin(1000.0, res[P_MINT]);
  }
  if (!isfinite(res[P_INT])) d.flags->nan = 1;
  const long long slot = step % d.hist_cap;
  for (int z = 0; z < nz; ++z)
    for (int k = 0; k < 3; ++k) d.react_hist[(slot * kMaxZones + z) * 3 + k] = res[P_REACT + 3 * z + k];
  DevFlags *f = d.flags;
  if (f->step_yield) {
    if (f->ts_i == f->ts_nb_inc - 1) f->cleared = 1;
    f->ts_i++;
    if (f->ts_i > f->ts_nb_inc - 1) f->ts_i = 0;
  }
}

void launch_finalize(const Dev &d, hipStream_t st, double dt, long long step, int nz, int full_step) {
  k_finalize<<<1, 256, 0, st>>>(d, dt, step, nz, full_step);
}

// ---------------------------------------------------------------------------
// utilities

WF_KERNEL void k_max_pl(Dev d) {
  __shared__ double sh[kBlock / 64];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const double v = i < d.n ? d.pl[i] : 0.0;
  const double r = block_reduce<1>(v, sh);
  if (threadIdx.x == 0) d.partials[(long long)blockIdx.x * P_COUNT + P_MAXT] = r;
}

__global__ __launch_bounds__(256) void k_max_pl_final(Dev d) {
  __shared__ double sh[4];
  double v = 0.0;
  for (int b = threadIdx.x; b < d.nblocks; b += blockDim.x) v = fmax(v, d.partials[(long long)b * P_COUNT + P_MAXT]);
  const double r = block_reduce<1>(v, sh);
  if (threadIdx.x == 0) d.scalars[S_MAXPL] = r;
}

void launch_max_pl(const Dev &d, hipStream_t st) {
  k_max_pl<<<d.nblocks, kBlock, 0, st>>>(d);
  k_max_pl_final<<<1, 256, 0, st>>>(d);
}

}  // namespace wf
