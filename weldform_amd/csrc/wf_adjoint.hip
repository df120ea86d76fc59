// wf_adjoint.hip -- reverse-mode adjoint of one FP64 Leapfrog step.
//
// The forward step (wf_kernels.hip, Source/SolverLeapfrog.cpp:124-362) maps
// the state s_n = (x, v, rho, S, Sy, pl) to s_{n+1} through
//   accel  a0 = sum_j m_j GK_ij (sigma_i/rho_i^2 + sigma_j/rho_j^2 + Pi_ij) x_ij
//   contact a1 = Wang penalty + friction on a0    (per particle, in order)
//   kick   v' = v + c a1  (c = dt, dt/2 on the first step; BC particles fixed)
//   density rho' = rho (1 + dt D),  D_i = sum_j (m_j/rho_j) v'_ij . GK x_ij
//   drift  x' = x + v' dt
//   rates  SR, W = sum_j (m_j/rho'_j) (-GK/2) sym/skew(v'_ij x_ij)  at x', v'
//   stress (S', Sy', pl') = Jaumann + radial return (S, Sy, pl, SR, W, T)
// with sigma = S - P(rho) I.  The loss seeds are the plane contact forces or
// the zone reactions of the step.  These kernels run that map backwards,
// one kernel per forward phase in reverse order, each re-evaluating its
// forward phase from the tape so every branch (yield, contact activity,
// friction regime, viscosity switch) is the forward's.
//
// Pair passes are gathers over each particle's own neighbour row, as in the
// forward (kAdjLanes threads per particle taking alternate entries, partial
// sums added in lane order): the pair term of (i, j) depends on i's and j's
// inputs, and its
// adjoint with respect to i's inputs needs only the pair's combined seed
// (m_j abar_i - m_i abar_j for the acceleration, vol_j sbar_i + vol_i sbar_j
// for the density and rate sums), so every particle accumulates its own
// adjoint in registers: no atomics, deterministic.
#include "wf_device.h"

namespace wf {

constexpr int kAdjLanes = 4;  // threads per particle in the reverse gathers

// d GK / d r2 of gk_t (the pair term depends on x_ij through r2 = x_ij.x_ij)
__device__ inline double gk_dr2(double r2, const KH<double> &k) {
  if (r2 == 0.0) return 0.0;
  const double ri = rsqrt_t(r2);
  const double q = r2 * ri * k.ih;
  if (q < 1.0) return 1.125 * k.c5 * k.ih * ri;
  if (q < 2.0) {
    const double t = 2.0 - q;
    return 0.75 * k.c5 * k.h * ri * ri * (t * k.ih + 0.5 * t * t * ri);
  }
  return 0.0;
}

// ---------------------------------------------------------------------------
// stress update (Particle::CalcStressStrain, Source/Particle.cpp:960-1067)

// partial derivatives of the Johnson-Cook yield stress
// (Source/Material.cpp:6-25): Sy = (A + B pl^n) f (1 - Th^m),
// f = 1 + C ln(sr / eps0) above eps0.  At pl = 0 the pl- and n-derivatives
// are taken as 0: pl = 0 only where the particle has never yielded, where
// pl does not depend on the parameters.
struct JcGrad {
  double dpl, dsr, dth[5];
};
__device__ inline JcGrad jc_grad(const DevMat &M, double pl, double sr, double T) {
  JcGrad g;
  const double Th = (T - M.jc_Tt) / (M.jc_Tm - M.jc_Tt);
  const double thm = pow(Th, M.jc_m);
  const double thf = 1.0 - thm;
  const bool rate = sr > M.jc_eps0;
  const double lr = rate ? log(sr / M.jc_eps0) : 0.0;
  const double f = rate ? 1.0 + M.jc_C * lr : 1.0;
  const double pn = pow(pl, M.jc_n);
  const double base = M.jc_A + M.jc_B * pn;
  g.dth[0] = f * thf;
  g.dth[1] = pn * f * thf;
  g.dth[2] = pl > 0.0 ? M.jc_B * pn * log(pl) * f * thf : 0.0;
  g.dth[3] = rate ? base * lr * thf : 0.0;
  g.dth[4] = Th > 0.0 ? -base * f * thm * log(Th) : 0.0;
  g.dpl = pl > 0.0 ? M.jc_B * M.jc_n * pow(pl, M.jc_n - 1.0) * f * thf : 0.0;
  g.dsr = rate ? base * (M.jc_C / sr) * thf : 0.0;
  return g;
}

WF_KERNEL void k_adj_stress(Dev d, AdjSlot s0, AdjBufs b, double dt, int adj_mat) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= d.n) return;
  const long long np = d.npad;
  double SRb[6] = {0, 0, 0, 0, 0, 0}, Wb[3] = {0, 0, 0};
  if (d.group[i] >= 0) {
    const DevMat &M = d.mats[d.mat[i]];
    double S[6], SR[6], W[3], Sbar_out[6];
    for (int k = 0; k < 6; ++k) {
      S[k] = s0.S[k * np + i];
      SR[k] = s0.sr[k * np + i];
      Sbar_out[k] = b.Sb[k * np + i];
    }
    for (int k = 0; k < 3; ++k) W[k] = s0.rr[k * np + i];
    const double Sy = s0.sigy[i], pl = s0.pl[i], T = d.T[i];
    // forward, as k_stress
    const double Sf[3][3] = {{S[XX], S[XY], S[XZ]}, {S[XY], S[YY], S[YZ]}, {S[XZ], S[YZ], S[ZZ]}};
    const double Wf[3][3] = {{0.0, W[0], W[1]}, {-W[0], 0.0, W[2]}, {-W[1], -W[2], 0.0}};
    const int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
    const double c3 = 1.0 / 3.0 * (SR[XX] + SR[YY] + SR[ZZ]);
    const double G2 = 2.0 * M.G;
    double Sn[6];
    for (int k = 0; k < 6; ++k) {
      const int a = ia[k], bb = ib[k];
      double jt = 0.0;
      for (int q = 0; q < 3; ++q) jt += Sf[a][q] * Wf[bb][q] + Wf[a][q] * Sf[q][bb];
      const double dev = k < 3 ? SR[k] - c3 : SR[k];
      Sn[k] = dt * (G2 * dev + jt) + S[k];
    }
    const double E = 0.5 * ((SR[XX] - SR[YY]) * (SR[XX] - SR[YY]) + (SR[YY] - SR[ZZ]) * (SR[YY] - SR[ZZ]) +
                            (SR[ZZ] - SR[XX]) * (SR[ZZ] - SR[XX])) +
                     3.0 * (SR[XY] * SR[XY] + SR[YZ] * SR[YZ] + SR[XZ] * SR[XZ]);
    const double effsr = sqrt(E);
    const double J2 = 0.5 * (Sn[XX] * Sn[XX] + 2.0 * Sn[XY] * Sn[XY] + 2.0 * Sn[XZ] * Sn[XZ] + Sn[YY] * Sn[YY] +
                             2.0 * Sn[YZ] * Sn[YZ] + Sn[ZZ] * Sn[ZZ]);
    const double st = sqrt(3.0 * J2);
    const double ratio = Sy / st;
    const bool capped = !(1.0 < ratio);
    const double fac = capped ? ratio : 1.0;
    double Sy2 = Sy;
    if (M.model == WF_HOLLOMON) Sy2 = hol_yield(M, pl);
    else if (M.model == WF_JOHNSON_COOK) Sy2 = jc_yield(M, pl, effsr, T);
    const bool yield = st > Sy2;
    const double H = 3. * M.G + M.Ep;
    // reverse
    const double plb = b.plb[i], syb = b.syb[i];
    double plbar = plb, sybar_old = 0.0, stbar = 0.0, effb = 0.0;
    const double depb = plb + ((yield && M.model == WF_BILINEAR) ? syb * M.Ep : 0.0);
    double sy2b = syb;
    if (yield) {
      stbar += depb / H;
      sy2b -= depb / H;
    }
    if (M.model == WF_BILINEAR) {
      sybar_old += sy2b;
    } else if (M.model == WF_HOLLOMON) {
      if (pl + M.hol_eps0 > M.hol_eps1)
        plbar += sy2b * M.hol_K * M.hol_m * pow(pl + M.hol_eps0, M.hol_m - 1.0);
    } else {
      const JcGrad g = jc_grad(M, pl, effsr, T);
      plbar += sy2b * g.dpl;
      effb += sy2b * g.dsr;
      if (d.mat[i] == adj_mat)
        for (int p = 0; p < 5; ++p) b.thb[p * np + i] += sy2b * g.dth[p];
    }
    double Snb[6];
    double facb = 0.0;
    for (int k = 0; k < 6; ++k) {
      Snb[k] = Sbar_out[k] * fac;
      facb += Sbar_out[k] * Sn[k];
    }
    if (capped) {
      sybar_old += facb / st;
      stbar -= facb * Sy / (st * st);
    }
    const double J2b = st > 0.0 ? stbar * 1.5 / st : 0.0;
    for (int k = 0; k < 6; ++k) Snb[k] += J2b * Sn[k] * (k < 3 ? 1.0 : 2.0);
    if (effsr > 0.0) {
      const double Eb = effb / (2.0 * effsr);
      SRb[XX] += Eb * (2.0 * SR[XX] - SR[YY] - SR[ZZ]);
      SRb[YY] += Eb * (2.0 * SR[YY] - SR[XX] - SR[ZZ]);
      SRb[ZZ] += Eb * (2.0 * SR[ZZ] - SR[XX] - SR[YY]);
      SRb[XY] += Eb * 6.0 * SR[XY];
      SRb[XZ] += Eb * 6.0 * SR[XZ];
      SRb[YZ] += Eb * 6.0 * SR[YZ];
    }
    double Sb_new[6];
    double Sfb[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, Wfb[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    double c3b = 0.0;
    for (int k = 0; k < 6; ++k) {
      Sb_new[k] = Snb[k];
      const double devb = Snb[k] * dt * G2, jtb = Snb[k] * dt;
      SRb[k] += devb;
      if (k < 3) c3b -= devb;
      const int a = ia[k], bb = ib[k];
      for (int q = 0; q < 3; ++q) {
        Sfb[a][q] += jtb * Wf[bb][q];
        Wfb[bb][q] += jtb * Sf[a][q];
        Wfb[a][q] += jtb * Sf[q][bb];
        Sfb[q][bb] += jtb * Wf[a][q];
      }
    }
    for (int k = 0; k < 3; ++k) SRb[k] += c3b * (1.0 / 3.0);
    Sb_new[XX] += Sfb[0][0];
    Sb_new[YY] += Sfb[1][1];
    Sb_new[ZZ] += Sfb[2][2];
    Sb_new[XY] += Sfb[0][1] + Sfb[1][0];
    Sb_new[XZ] += Sfb[0][2] + Sfb[2][0];
    Sb_new[YZ] += Sfb[1][2] + Sfb[2][1];
    Wb[0] = Wfb[0][1] - Wfb[1][0];
    Wb[1] = Wfb[0][2] - Wfb[2][0];
    Wb[2] = Wfb[1][2] - Wfb[2][1];
    for (int k = 0; k < 6; ++k) b.Sb[k * np + i] = Sb_new[k];
    b.syb[i] = sybar_old;
    b.plb[i] = plbar;
  }
  b.rsb[3 * i] = make_double4(SRb[XX], SRb[YY], SRb[ZZ], SRb[XY]);
  b.rsb[3 * i + 1] = make_double4(SRb[XZ], SRb[YZ], Wb[0], Wb[1]);
  b.rsb[3 * i + 2] = make_double4(Wb[2], 0.0, 0.0, 0.0);
}

// the packed rate-adjoint record of particle j: strain rate [6], rotation [3]
__device__ inline void rate_seeds(const AdjBufs &b, int j, double sb[6], double wb[3]) {
  const double4 a = b.rsb[3 * j], c = b.rsb[3 * j + 1], e = b.rsb[3 * j + 2];
  sb[XX] = a.x, sb[YY] = a.y, sb[ZZ] = a.z, sb[XY] = a.w, sb[XZ] = c.x, sb[YZ] = c.y;
  wb[0] = c.z, wb[1] = c.w, wb[2] = e.x;
}

// ---------------------------------------------------------------------------
// rate-tensor gather (CalcRateTensors, Source/InteractionAlt.cpp:263-479) at
// x', v', rho', then the drift x' = x + v' dt and the density update
// rho' = rho (1 + dt D) backwards

// LANES threads per particle take alternate row entries; their partial
// sums are added in lane order
template <int LANES>
__global__ __launch_bounds__(kBlock *LANES) void k_adj_rates(Dev d, AdjSlot s0, AdjSlot s1, AdjBufs b, double dt) {
  __shared__ double red[(LANES > 1 ? LANES - 1 : 1) * 7 * kBlock];
  const int sub = threadIdx.x / kBlock;
  const int i = blockIdx.x * kBlock + (int)(threadIdx.x % kBlock);
  const long long np = d.npad;
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};  // xb[3], vb[3], volb
  const bool on = i < d.n;
  const double4 xi = on ? s1.xh[i] : make_double4(0, 0, 0, 1), vi = on ? s1.vm[i] : make_double4(0, 0, 0, 0);
  const double rhoi1 = on ? s1.rho[i] : 1.0;
  if (on && d.group[i] >= 0) {
    const double voli = vi.w / rhoi1;
    double sbi[6], wbi[3];
    rate_seeds(b, i, sbi, wbi);
    const KH<double> ki = kern_h_t<double>(xi.w);
    const int cnt = s0.nnb[i];
    for (int q = sub; q < cnt; q += LANES) {
      const int j = s0.nbr[nb_slot(np, q, i)];
      const double4 xj = s1.xh[j], vj = s1.vm[j];
      const double volj = vj.w / s1.rho[j];
      double sbj[6], wbj[3];
      rate_seeds(b, j, sbj, wbj);
      const KH<double> kh = pair_h_t<double>(ki, xi.w, xj.w);
      const double dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
      const double r2 = dx * dx + dy * dy + dz * dz;
      const double GK = gk_t<double>(r2, kh);
      const double vx = vi.x - vj.x, vy = vi.y - vj.y, vz = vi.z - vj.z;
      const double cp = -0.5 * GK;
      const double e[6] = {2 * vx * dx, 2 * vy * dy, 2 * vz * dz, vx * dy + vy * dx, vx * dz + vz * dx, vy * dz + vz * dy};
      const double w[3] = {vx * dy - vy * dx, vx * dz - vz * dx, vy * dz - vz * dy};
      double eb[6], wbv[3], cpb = 0.0;
      for (int k = 0; k < 6; ++k) {
        eb[k] = volj * sbi[k] + voli * sbj[k];
        cpb += eb[k] * e[k];
        acc[6] += sbj[k] * cp * e[k];
      }
      for (int k = 0; k < 3; ++k) {
        wbv[k] = volj * wbi[k] + voli * wbj[k];
        cpb += wbv[k] * w[k];
        acc[6] += wbj[k] * cp * w[k];
      }
      for (int k = 0; k < 6; ++k) eb[k] *= cp;
      for (int k = 0; k < 3; ++k) wbv[k] *= cp;
      double bvx = 2 * eb[0] * dx + eb[3] * dy + eb[4] * dz + wbv[0] * dy + wbv[1] * dz;
      double bvy = 2 * eb[1] * dy + eb[3] * dx + eb[5] * dz - wbv[0] * dx + wbv[2] * dz;
      double bvz = 2 * eb[2] * dz + eb[4] * dx + eb[5] * dy - wbv[1] * dx - wbv[2] * dy;
      double bdx = 2 * eb[0] * vx + eb[3] * vy + eb[4] * vz - wbv[0] * vy - wbv[1] * vz;
      double bdy = 2 * eb[1] * vy + eb[3] * vx + eb[5] * vz + wbv[0] * vx - wbv[2] * vz;
      double bdz = 2 * eb[2] * vz + eb[4] * vx + eb[5] * vy + wbv[1] * vx + wbv[2] * vy;
      const double r2b = -0.5 * cpb * gk_dr2(r2, kh);
      bdx += 2 * r2b * dx;
      bdy += 2 * r2b * dy;
      bdz += 2 * r2b * dz;
      acc[3] += bvx, acc[4] += bvy, acc[5] += bvz;
      acc[0] += bdx, acc[1] += bdy, acc[2] += bdz;
    }
  }
  combine_lanes<LANES>(acc, red);
  if (!on || sub != 0) return;
  const double xb[3] = {acc[0], acc[1], acc[2]}, vb[3] = {acc[3], acc[4], acc[5]}, volb = acc[6];
  // x' and v' adjoints: incoming state adjoint plus the gather
  const double4 xin = b.xb[i], vin = b.vb[i];
  const double xpb[3] = {xin.x + xb[0], xin.y + xb[1], xin.z + xb[2]};
  const double rho0 = s0.rho[i];
  const double rhopb = b.rhob[i] + volb * (-vi.w / (rhoi1 * rhoi1));
  // drift: x' = x + v' dt
  b.xb[i] = make_double4(xpb[0], xpb[1], xpb[2], 0.0);
  b.vpb[i] = make_double4(vin.x + vb[0] + dt * xpb[0], vin.y + vb[1] + dt * xpb[1], vin.z + vb[2] + dt * xpb[2], 0.0);
  // density update: rho' = rho + dt (D rho), D = drho / rho
  const double D = s0.drho[i] / rho0;
  b.rhob[i] = rhopb * (1.0 + dt * D);
  b.db[i] = rhopb * dt * rho0;
}

// ---------------------------------------------------------------------------
// density-rate gather (CalcDensInc, Source/InteractionAlt.cpp:483-640) at
// x, v', rho backwards, then the kick v' = v + c a (Source/SolverLeapfrog.
// cpp:241-273) with the velocity BCs (Source/WeldForm.cpp:51-92) and the
// zone-reaction seeds of the observer

template <int LANES>
__global__ __launch_bounds__(kBlock *LANES) void k_adj_dens(Dev d, AdjSlot s0, AdjSlot s1, AdjBufs b, BcTable bc,
                                                            AdjSeeds seeds, double kick) {
  __shared__ double red[(LANES > 1 ? LANES - 1 : 1) * 7 * kBlock];
  const int sub = threadIdx.x / kBlock;
  const int i = blockIdx.x * kBlock + (int)(threadIdx.x % kBlock);
  const long long np = d.npad;
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};  // xb[3], vb[3], volb
  const bool on = i < d.n;
  const double4 xi = on ? s0.xh[i] : make_double4(0, 0, 0, 1), vi = on ? s1.vm[i] : make_double4(0, 0, 0, 0);
  const double rhoi = on ? s0.rho[i] : 1.0;
  if (on && d.group[i] >= 0) {
    const double voli = vi.w / rhoi;
    const double dbi = b.db[i];
    const KH<double> ki = kern_h_t<double>(xi.w);
    const int cnt = s0.nnb[i];
    for (int q = sub; q < cnt; q += LANES) {
      const int j = s0.nbr[nb_slot(np, q, i)];
      const double4 xj = s0.xh[j], vj = s1.vm[j];
      const double volj = vj.w / s0.rho[j];
      const KH<double> kh = pair_h_t<double>(ki, xi.w, xj.w);
      const double dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
      const double r2 = dx * dx + dy * dy + dz * dz;
      const double GK = gk_t<double>(r2, kh);
      const double vx = vi.x - vj.x, vy = vi.y - vj.y, vz = vi.z - vj.z;
      const double vd = vx * dx + vy * dy + vz * dz;
      const double dbj = b.db[j];
      const double wb = volj * dbi + voli * dbj;
      acc[6] += dbj * GK * vd;
      acc[3] += wb * GK * dx, acc[4] += wb * GK * dy, acc[5] += wb * GK * dz;
      const double r2b = wb * vd * gk_dr2(r2, kh);
      acc[0] += wb * GK * vx + 2 * r2b * dx;
      acc[1] += wb * GK * vy + 2 * r2b * dy;
      acc[2] += wb * GK * vz + 2 * r2b * dz;
    }
  }
  combine_lanes<LANES>(acc, red);
  if (!on || sub != 0) return;
  const double xb[3] = {acc[0], acc[1], acc[2]}, vb[3] = {acc[3], acc[4], acc[5]}, volb = acc[6];
  double4 x0 = b.xb[i];
  b.xb[i] = make_double4(x0.x + xb[0], x0.y + xb[1], x0.z + xb[2], 0.0);
  b.rhob[i] += volb * (-vi.w / (rhoi * rhoi));
  const double4 vp = b.vpb[i];
  const double vpb[3] = {vp.x + vb[0], vp.y + vb[1], vp.z + vb[2]};
  // the kick: BC particles take the prescribed velocity and a = 0
  const int id = s0.id[i];
  bool fixed = false;
  for (int k = 0; k < bc.n; ++k) {
    if (bc.zone[k] != id || bc.type[k] != 0) continue;
    if (bc.value_type[k] == 1 && bc.amp[k] < 0) continue;
    fixed = true;
  }
  double ab[3] = {0, 0, 0};
  if (fixed) {
    b.vb[i] = make_double4(0, 0, 0, 0);
  } else {
    b.vb[i] = make_double4(vpb[0], vpb[1], vpb[2], 0.0);
    for (int k = 0; k < 3; ++k) ab[k] = kick * vpb[k];
  }
  // reaction observer: R_z = sum over the zone of m a (before the BC)
  for (int z = 0; z < bc.nz; ++z)
    if (id == bc.rzone[z])
      for (int k = 0; k < 3; ++k) ab[k] += seeds.react[z][k] * vi.w;
  b.ab[i] = make_double4(ab[0], ab[1], ab[2], 0.0);
}

// ---------------------------------------------------------------------------
// Wang contact (Source/Contact_Wang.cpp:15-348) backwards.  Each partner
// updates the acceleration the next partner's predicted position uses, so
// the partners are reversed in turn; the acceleration entering partner q is
// recomputed by running the forward over partners 0..q-1.

struct CPair {
  bool active, fric, stat;
  double n[3], v2[3], xp[3], cf0[3], tg[3], ncf, ntg, kij, psi;
  int mesh;
};

// forward of one partner: updates a and returns the pair's quantities
__device__ inline CPair contact_pair(const Dev &d, const AdjSlot &s0, int r, const double x[3], const double v[3],
                                     double a[3], double h, double m1, double dt) {
  CPair p;
  p.active = false;
  const double4 en = s0.enorm[r];
  const double n2[3] = {en.x, en.y, en.z};
  const double4 v24 = s0.rv[r];
  const double v2[3] = {v24.x, v24.y, v24.z};
  for (int k = 0; k < 3; ++k) {
    p.n[k] = n2[k];
    p.v2[k] = v2[k];
  }
  p.mesh = d.rmesh[r];
  const double delta_ = -(n2[0] * (v[0] - v2[0]) + n2[1] * (v[1] - v2[1]) + n2[2] * (v[2] - v2[2]));
  for (int k = 0; k < 3; ++k) p.xp[k] = x[k] + v[k] * dt + a[k] * dt * dt / 2.0;
  const double dist = (n2[0] * p.xp[0] + n2[1] * p.xp[1] + n2[2] * p.xp[2]) - en.w;
  if (!(dist < h)) return p;
  double qj[3];
  for (int k = 0; k < 3; ++k) qj[k] = x[k] - dist * n2[k];
  const int4 el = d.elem[r];
  const int nd[3] = {el.x, el.y, el.z};
  for (int ii = 0; ii < 3; ++ii) {
    const int jj = ii + 1 > 2 ? 0 : ii + 1;
    const double4 ni = s0.node[nd[ii]], nj = s0.node[nd[jj]];
    const double ea[3] = {nj.x - ni.x, nj.y - ni.y, nj.z - ni.z};
    const double eb[3] = {qj[0] - ni.x, qj[1] - ni.y, qj[2] - ni.z};
    double c[3];
    cross_d(ea, eb, c);
    if (c[0] * n2[0] + c[1] * n2[1] + c[2] * n2[2] < 0.0) return p;
  }
  p.active = true;
  const double delta = h - dist;
  p.kij = 2.0 * m1 / (dt * dt);
  const double omega = sqrt(p.kij / m1);
  p.psi = 2. * m1 * omega * d.dfac;
  for (int k = 0; k < 3; ++k) p.cf0[k] = (p.kij * delta - p.psi * delta_) * n2[k];
  for (int k = 0; k < 3; ++k) a[k] += p.cf0[k] / m1;
  p.fric = d.fr_sta > 0.;
  p.stat = false;
  if (p.fric) {
    double du[3];
    for (int k = 0; k < 3; ++k) du[k] = p.xp[k] - x[k] - v2[k] * dt;
    const double dn = du[0] * n2[0] + du[1] * n2[1] + du[2] * n2[2];
    for (int k = 0; k < 3; ++k) p.tg[k] = p.kij * (du[k] - dn * n2[k]);
    p.ntg = sqrt(p.tg[0] * p.tg[0] + p.tg[1] * p.tg[1] + p.tg[2] * p.tg[2]);
    p.ncf = sqrt(p.cf0[0] * p.cf0[0] + p.cf0[1] * p.cf0[1] + p.cf0[2] * p.cf0[2]);
    p.stat = p.ntg < d.fr_sta * p.ncf;
    if (p.stat) {
      for (int k = 0; k < 3; ++k) a[k] -= p.tg[k] / m1;
    } else {
      for (int k = 0; k < 3; ++k) a[k] -= d.fr_dyn * p.ncf * p.tg[k] / p.ntg / m1;
    }
  }
  return p;
}

WF_KERNEL void k_adj_contact(Dev d, AdjSlot s0, AdjBufs b, AdjSeeds seeds, double dt) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= d.n) return;
  const int cnt = s0.ccnt[i];
  if (cnt == 0) return;
  const long long np = d.npad;
  const double4 x4 = s0.xh[i], v4 = s0.vm[i], a4 = s0.apre[i];
  const double x[3] = {x4.x, x4.y, x4.z}, v[3] = {v4.x, v4.y, v4.z};
  const double h = x4.w, m1 = v4.w;
  const double4 ab4 = b.ab[i];
  double ab[3] = {ab4.x, ab4.y, ab4.z};  // adjoint of a after partner q
  double xb[3] = {0, 0, 0}, vb[3] = {0, 0, 0}, musb = 0.0, mudb = 0.0;
  for (int q = cnt - 1; q >= 0; --q) {
    double a[3] = {a4.x, a4.y, a4.z};
    for (int p = 0; p < q; ++p) (void)contact_pair(d, s0, s0.cidx[(long long)p * np + i], x, v, a, h, m1, dt);
    const CPair P = contact_pair(d, s0, s0.cidx[(long long)q * np + i], x, v, a, h, m1, dt);
    if (!P.active) continue;
    // seed of this partner's force in its plane's sum
    double cfb[3] = {0, 0, 0};
    for (int m = 0; m < kMaxMesh; ++m)
      if (m == P.mesh)
        for (int k = 0; k < 3; ++k) cfb[k] = seeds.cf[m][k];
    double cf0b[3] = {cfb[0], cfb[1], cfb[2]};
    double xpb[3] = {0, 0, 0};
    if (P.fric) {
      double tgb[3];
      if (P.stat) {
        for (int k = 0; k < 3; ++k) tgb[k] = -cfb[k] / m1 - ab[k] / m1;
      } else {
        // td = mu_d ncf tg / ntg; cf -= td; a -= td / m
        double tdb[3], tdtg = 0.0;
        for (int k = 0; k < 3; ++k) {
          tdb[k] = -cfb[k] - ab[k] / m1;
          tdtg += tdb[k] * P.tg[k];
        }
        mudb += P.ncf * tdtg / P.ntg;
        const double ncfb = d.fr_dyn * tdtg / P.ntg;
        const double s = d.fr_dyn * P.ncf / P.ntg;
        for (int k = 0; k < 3; ++k) tgb[k] = s * (tdb[k] - tdtg * P.tg[k] / (P.ntg * P.ntg));
        for (int k = 0; k < 3; ++k) cf0b[k] += ncfb * P.cf0[k] / P.ncf;
      }
      // tg = kij (du - (du.n) n), du = xp - x - v2 dt
      const double tn = tgb[0] * P.n[0] + tgb[1] * P.n[1] + tgb[2] * P.n[2];
      for (int k = 0; k < 3; ++k) {
        const double dub = P.kij * (tgb[k] - tn * P.n[k]);
        xpb[k] += dub;
        xb[k] -= dub;
      }
    }
    // a += cf0 / m
    for (int k = 0; k < 3; ++k) cf0b[k] += ab[k] / m1;
    // cf0 = (kij (h - dist) - psi delta_) n, dist = n.xp - p, delta_ = -n.(v - v2)
    const double cn = cf0b[0] * P.n[0] + cf0b[1] * P.n[1] + cf0b[2] * P.n[2];
    const double distb = -P.kij * cn, deltab_ = -P.psi * cn;
    for (int k = 0; k < 3; ++k) {
      xpb[k] += distb * P.n[k];
      vb[k] += -deltab_ * P.n[k];
    }
    // xp = x + v dt + a dt^2 / 2 (a entering this partner)
    for (int k = 0; k < 3; ++k) {
      xb[k] += xpb[k];
      vb[k] += xpb[k] * dt;
      ab[k] += xpb[k] * dt * dt / 2.0;
    }
  }
  b.ab[i] = make_double4(ab[0], ab[1], ab[2], 0.0);
  const double4 xo = b.xb[i], vo = b.vb[i];
  b.xb[i] = make_double4(xo.x + xb[0], xo.y + xb[1], xo.z + xb[2], 0.0);
  b.vb[i] = make_double4(vo.x + vb[0], vo.y + vb[1], vo.z + vb[2], 0.0);
  b.thb[5 * np + i] += musb;
  b.thb[6 * np + i] += mudb;
}

// ---------------------------------------------------------------------------
// acceleration gather (CalcAccel, Source/InteractionAlt.cpp:12-260) at x, v,
// rho, sigma = S - P(rho) I backwards

// the step's sigma / rho^2 (sigma = S - P(rho) I) packed per particle for
// the acceleration gather (two records, see AdjBufs::tsig)
WF_KERNEL void k_adj_tsig(Dev d, AdjSlot s0, AdjBufs b) {
  const int j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= d.n) return;
  const DevMat &M = d.mats[d.mat[j]];
  const double rho = s0.rho[j];
  const double P = M.P0 + (M.Cs * M.Cs) * (rho - M.ref_density);
  const double irho2 = 1.0 / (rho * rho);
  const long long np = d.npad;
  double t[6];
  for (int k = 0; k < 6; ++k) {
    const double sg = k < 3 ? -P + s0.S[k * np + j] : s0.S[k * np + j];
    t[k] = sg * irho2;
  }
  b.tsig[2 * j] = make_double4(t[XX], t[YY], t[ZZ], t[XY]);
  b.tsig[2 * j + 1] = make_double4(t[XZ], t[YZ], rho, irho2);
}
__device__ inline void sigma_t(const AdjBufs &b, int j, double t[6], double &rho, double &irho2) {
  const double4 a = b.tsig[2 * j], c = b.tsig[2 * j + 1];
  t[XX] = a.x, t[YY] = a.y, t[ZZ] = a.z, t[XY] = a.w, t[XZ] = c.x, t[YZ] = c.y;
  rho = c.z;
  irho2 = c.w;
}

template <int LANES>
__global__ __launch_bounds__(kBlock *LANES) void k_adj_accel(Dev d, AdjSlot s0, AdjBufs b) {
  __shared__ double red[(LANES > 1 ? LANES - 1 : 1) * 13 * kBlock];
  const int sub = threadIdx.x / kBlock;
  const int i = blockIdx.x * kBlock + (int)(threadIdx.x % kBlock);
  const bool on = i < d.n && d.group[i] >= 0;
  const long long np = d.npad;
  const DevMat &M = d.mats[on ? d.mat[i] : 0];
  const double beta = M.beta, acs = M.alpha * M.Cs;
  const bool av = (M.alpha != 0.0 || M.beta != 0.0);
  double acc[13] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};  // xb[3], vb[3], tb[6], rhob
  double ti[6] = {0, 0, 0, 0, 0, 0}, rhoi = 1.0, irho2i = 1.0;
  if (on) {
    const double4 xi = s0.xh[i], vi = s0.vm[i], abi = b.ab[i];
    sigma_t(b, i, ti, rhoi, irho2i);
    const KH<double> ki = kern_h_t<double>(xi.w);
    const int cnt = s0.nnb[i];
    for (int q = sub; q < cnt; q += LANES) {
      const int j = s0.nbr[nb_slot(np, q, i)];
      const double4 xj = s0.xh[j], vj = s0.vm[j], abj = b.ab[j];
      double tj[6], rhoj, irho2j;
      sigma_t(b, j, tj, rhoj, irho2j);
      const KH<double> kh = pair_h_t<double>(ki, xi.w, xj.w);
      const double dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
      const double r2 = dx * dx + dy * dy + dz * dz;
      const double GK = gk_t<double>(r2, kh);
      const double vx = vi.x - vj.x, vy = vi.y - vj.y, vz = vi.z - vj.z;
      const double vdot = vx * dx + vy * dy + vz * dz;
      const double D1 = r2 + 0.01 * kh.h * kh.h, D2 = 0.5 * (rhoi + rhoj);
      const double hv = kh.h * vdot / (D1 * D2);
      const double pv = hv * (acs + beta * (hv * D2));
      const bool act = av && vdot < 0.0;
      const double pi = act ? pv : 0.0;
      const double Mxx = ti[XX] + tj[XX] + pi, Myy = ti[YY] + tj[YY] + pi, Mzz = ti[ZZ] + tj[ZZ] + pi;
      const double Mxy = ti[XY] + tj[XY], Mxz = ti[XZ] + tj[XZ], Myz = ti[YZ] + tj[YZ];
      // combined seed of the pair term P_ij = GK M x_ij (a_i += m_j P, a_j -= m_i P)
      const double cx = vj.w * abi.x - vi.w * abj.x, cy = vj.w * abi.y - vi.w * abj.y,
                   cz = vj.w * abi.z - vi.w * abj.z;
      const double mdx = Mxx * dx + Mxy * dy + Mxz * dz, mdy = Mxy * dx + Myy * dy + Myz * dz,
                   mdz = Mxz * dx + Myz * dy + Mzz * dz;
      const double GKb = cx * mdx + cy * mdy + cz * mdz;
      const double mb[6] = {GK * cx * dx, GK * cy * dy, GK * cz * dz, GK * (cx * dy + cy * dx), GK * (cx * dz + cz * dx),
                            GK * (cy * dz + cz * dy)};
      for (int k = 0; k < 6; ++k) acc[6 + k] += mb[k];
      double bdx = GK * (Mxx * cx + Mxy * cy + Mxz * cz);
      double bdy = GK * (Mxy * cx + Myy * cy + Myz * cz);
      double bdz = GK * (Mxz * cx + Myz * cy + Mzz * cz);
      double r2b = GKb * gk_dr2(r2, kh);
      if (act) {
        const double pib = mb[0] + mb[1] + mb[2];
        const double hvb = pib * (acs + 2.0 * beta * hv * D2);
        double D2b = pib * beta * hv * hv;
        const double vdb = hvb * kh.h / (D1 * D2);
        const double D1b = -hvb * hv / D1;
        D2b += -hvb * hv / D2;
        r2b += D1b;
        acc[12] += 0.5 * D2b;
        acc[3] += vdb * dx, acc[4] += vdb * dy, acc[5] += vdb * dz;
        bdx += vdb * vx, bdy += vdb * vy, bdz += vdb * vz;
      }
      acc[0] += bdx + 2 * r2b * dx;
      acc[1] += bdy + 2 * r2b * dy;
      acc[2] += bdz + 2 * r2b * dz;
    }
  }
  combine_lanes<LANES>(acc, red);
  if (!on || sub != 0) return;
  const double *tb = acc + 6;
  double rhob = acc[12];
  // t = sigma / rho^2, sigma = S - P(rho) I
  double sgb[6], tsig = 0.0;
  for (int k = 0; k < 6; ++k) {
    sgb[k] = tb[k] * irho2i;
    tsig += tb[k] * ti[k];  // t:tbar = (sigma:tbar) / rho^2
  }
  rhob += -2.0 * tsig / rhoi;
  rhob += -(M.Cs * M.Cs) * (sgb[XX] + sgb[YY] + sgb[ZZ]);
  for (int k = 0; k < 6; ++k) b.Sb[k * np + i] += sgb[k];
  b.rhob[i] += rhob;
  const double4 xo = b.xb[i], vo = b.vb[i];
  b.xb[i] = make_double4(xo.x + acc[0], xo.y + acc[1], xo.z + acc[2], 0.0);
  b.vb[i] = make_double4(vo.x + acc[3], vo.y + acc[4], vo.z + acc[5], 0.0);
}

void launch_adj_step(const Dev &d, hipStream_t st, const AdjSlot &s0, const AdjSlot &s1, const AdjBufs &b,
                     const BcTable &bc, const AdjSeeds &seeds, double dt, double kick, int adj_mat) {
  const int nb = nblk(d.n);
  constexpr int L = kAdjLanes;
  k_adj_tsig<<<nb, kBlock, 0, st>>>(d, s0, b);
  k_adj_stress<<<nb, kBlock, 0, st>>>(d, s0, b, dt, adj_mat);
  k_adj_rates<L><<<nb, kBlock * L, 0, st>>>(d, s0, s1, b, dt);
  k_adj_dens<L><<<nb, kBlock * L, 0, st>>>(d, s0, s1, b, bc, seeds, kick);
  if (d.contact) k_adj_contact<<<nb, kBlock, 0, st>>>(d, s0, b, seeds, dt);
  k_adj_accel<L><<<nb, kBlock * L, 0, st>>>(d, s0, b);
}

// parameter adjoints: fixed-order block partials, then one block
WF_KERNEL void k_adj_part(Dev d, AdjBufs b) {
  __shared__ double sh[kBlock / 64];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  for (int p = 0; p < kAdjParams; ++p) {
    const double v = i < d.n ? b.thb[(long long)p * d.npad + i] : 0.0;
    const double r = block_reduce<0>(v, sh);
    if (threadIdx.x == 0) b.part[(long long)p * d.nblocks + blockIdx.x] = r;
  }
}
__global__ __launch_bounds__(1024) void k_adj_final(Dev d, AdjBufs b, double *out) {
  __shared__ double sh[16];
  for (int p = 0; p < kAdjParams; ++p) {
    double v = 0.0;
    for (int k = threadIdx.x; k < d.nblocks; k += blockDim.x) v += b.part[(long long)p * d.nblocks + k];
    const double r = block_reduce<0>(v, sh);
    if (threadIdx.x == 0) out[p] = r;
  }
}
void launch_adj_reduce(const Dev &d, hipStream_t st, const AdjBufs &b, double *out) {
  k_adj_part<<<d.nblocks, kBlock, 0, st>>>(d, b);
  k_adj_final<<<1, 1024, 0, st>>>(d, b, out);
}

// tape: copy the neighbour and contact lists into a slot when this step
// searched (or when forced, for a segment's first step); ver[slot] = 1 marks
// a stored version
WF_KERNEL void k_tape_lists(Dev d, AdjSlot s, int force, int *ver, int slot) {
  if (!(force || d.flags->do_search)) return;
  const long long np = d.npad;
  const long long nrow = (long long)nb_rows(d.max_nb) * np;
  const long long stride = (long long)gridDim.x * kBlock;
  const long long t0 = (long long)blockIdx.x * kBlock + threadIdx.x;
  for (long long k = t0; k < nrow; k += stride) s.nbr[k] = d.nbr[k];
  for (long long k = t0; k < np; k += stride) {
    s.nnb[k] = d.nnb[k];
    if (d.contact) {
      s.ccnt[k] = d.ccnt[k];
      s.id[k] = d.id[k];
    } else {
      s.id[k] = d.id[k];
    }
  }
  if (d.contact)
    for (long long k = t0; k < (long long)kMaxContact * np; k += stride) s.cidx[k] = d.cidx[k];
  if (t0 == 0) ver[slot] = 1;
}
void launch_tape_lists(const Dev &d, hipStream_t st, const AdjSlot &s, int force, int *ver, int slot) {
  k_tape_lists<<<nblk_small(d.npad), kBlock, 0, st>>>(d, s, force, ver, slot);
}

}  // namespace wf
