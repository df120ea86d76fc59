// wf_output.hip -- the particle table of the reference's output writers,
// packed on the device in the caller's particle order so that one copy
// brings it to the host:
//   * Domain::WriteXDMF (Source/Output.cpp:75-330): the float32 / int32
//     HDF5 datasets, dataset after dataset, each n x width values;
//   * Domain::WriteCSV (Source/Output.cpp:521-554): the 17 columns of one
//     row per particle, in FP64 (the CSV prints doubles).
// A thread packs one particle: it reads each own-state array once
// (coalesced, internal order) and scatters to perm[k], the particle's
// position in the caller's order.
#include <mutex>

#include "wf_device.h"

namespace wf {

// Particle::CalculateEquivalentStress (Source/Particle.cpp:623-633) on the
// deviatoric stress S [6][np], in the reference's order of operations and
// without contraction into fused multiply-adds, so the value is the one the
// reference's host code forms from the same FP64 ShearStress.
__device__ double sigma_eq(const double *S, long long np, long long k) {
#pragma clang fp contract(off)
  const double xx = S[XX * np + k], yy = S[YY * np + k], zz = S[ZZ * np + k];
  const double xy = S[XY * np + k], xz = S[XZ * np + k], yz = S[YZ * np + k];
  const double j2 = 0.5 * (xx * xx + 2.0 * xy * xy + 2.0 * xz * xz + yy * yy + 2.0 * yz * yz + zz * zz);
  return sqrt(3.0 * j2);
}

// a symmetric tensor in the reference's dataset order (00, 11, 22, 01, 12, 02)
__device__ inline void put_t6(float *dst, const double *t, long long np, long long k) {
  dst[0] = (float)t[XX * np + k];
  dst[1] = (float)t[YY * np + k];
  dst[2] = (float)t[ZZ * np + k];
  dst[3] = (float)t[XY * np + k];
  dst[4] = (float)t[YZ * np + k];
  dst[5] = (float)t[XZ * np + k];
}

WF_KERNEL void k_pack_xdmf(Dev d, int cur, float *f, int *iv) {
  const long long k = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (k >= d.n) return;
  const long long n = d.n, np = d.npad;
  const long long o = d.perm[k];
  auto col = [&](int off, int w) { return f + (long long)off * n + o * w; };
  const double4 x = d.xh[cur][k], v = d.vm[k], a = d.acc[k], u = d.disp[k];
  float *p = col(WF_XDMF_POSITION, 3);
  p[0] = (float)x.x, p[1] = (float)x.y, p[2] = (float)x.z;
  p = col(WF_XDMF_VELOCITY, 3);
  p[0] = (float)v.x, p[1] = (float)v.y, p[2] = (float)v.z;
  p = col(WF_XDMF_ACCELERATION, 3);
  p[0] = (float)a.x, p[1] = (float)a.y, p[2] = (float)a.z;
  *col(WF_XDMF_PRESSURE, 1) = (float)d.pres[k];
  *col(WF_XDMF_DENSITY, 1) = (float)d.rho[cur][k];
  // UserOutput's default (Source/Domain.cpp:53-58) sets all three to 0
  *col(WF_XDMF_PROPERTY1, 1) = 0.f;
  *col(WF_XDMF_PROPERTY2, 1) = 0.f;
  *col(WF_XDMF_PROPERTY3, 1) = 0.f;
  *col(WF_XDMF_MASS, 1) = (float)v.w;
  *col(WF_XDMF_H, 1) = (float)x.w;
  {
    const StressRec &r = d.srec[k];
    float *s = col(WF_XDMF_SIGMA, 6);
    s[0] = (float)r.s[XX], s[1] = (float)r.s[YY], s[2] = (float)r.s[ZZ];
    s[3] = (float)r.s[XY], s[4] = (float)r.s[YZ], s[5] = (float)r.s[XZ];
    // the reference writes its Sigma buffer under the name ShearS
    // (Source/Output.cpp:250-251)
    float *q = col(WF_XDMF_SHEARS, 6);
    for (int c = 0; c < 6; ++c) q[c] = s[c];
  }
  put_t6(col(WF_XDMF_STRAIN, 6), d.strain, np, k);
  put_t6(col(WF_XDMF_STRAIN_RATE, 6), d.sr, np, k);
  put_t6(col(WF_XDMF_STRAIN_PL, 6), d.epl, np, k);
  if (d.gc_on) {
    put_t6(col(WF_XDMF_GRADCORRMAT, 6), d.gcm, np, k);
  } else {
    // gradCorrM is only formed by the kernel-gradient correction; the
    // reference leaves it unset without it
    float *g = col(WF_XDMF_GRADCORRMAT, 6);
    for (int c = 0; c < 6; ++c) g[c] = 0.f;
  }
  *col(WF_XDMF_SIGMA_EQ, 1) = (float)sigma_eq(d.S, np, k);
  *col(WF_XDMF_TEMPERATURE, 1) = (float)d.T[k];
  *col(WF_XDMF_PL_STRAIN, 1) = (float)d.pl[k];
  // friction heat flux (friction_hfl) and contact shear stress are set by
  // the contact's friction-heating branch (Source/Contact_Wang.cpp:289-298)
  // and reset to 0 by every contact evaluation (:34-39)
  *col(WF_XDMF_Q_FRICTION, 1) = d.fhfl ? (float)d.fhfl[k] : 0.f;
  *col(WF_XDMF_C_SHEARABS, 1) = d.cshear ? (float)d.cshear[k] : 0.f;
  *col(WF_XDMF_PS_EN, 1) = (float)d.psen[k];
  p = col(WF_XDMF_DISPLACEMENT, 3);
  p[0] = (float)u.x, p[1] = (float)u.y, p[2] = (float)u.z;
  p = col(WF_XDMF_CONTACT_FORCE, 3);
  if (d.contact) {
    const double4 cf = d.cforce[k];
    p[0] = (float)cf.x, p[1] = (float)cf.y, p[2] = (float)cf.z;
  } else {
    p[0] = p[1] = p[2] = 0.f;
  }
  iv[WF_XDMF_TAG * n + o] = d.id[k];
  iv[WF_XDMF_NEIGHBORS * n + o] = d.nnb[k];
  iv[WF_XDMF_CONTNEIB * n + o] = d.contact ? d.ccnt[k] : 0;
}

WF_KERNEL void k_pack_csv(Dev d, int cur, double *rows) {
  const long long k = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (k >= d.n) return;
  const long long np = d.npad;
  double *r = rows + (long long)d.perm[k] * WF_CSV_NCOL;
  const double4 x = d.xh[cur][k], v = d.vm[k], a = d.acc[k];
  r[0] = x.x, r[1] = x.y, r[2] = x.z;
  r[3] = (double)d.id[k];
  r[4] = sigma_eq(d.S, np, k);
  r[5] = d.pl[k];
  r[6] = v.x, r[7] = v.y, r[8] = v.z;
  r[9] = a.x, r[10] = a.y, r[11] = a.z;
  if (d.contact) {
    const double4 cf = d.cforce[k];
    r[12] = cf.x, r[13] = cf.y, r[14] = cf.z;
  } else {
    r[12] = r[13] = r[14] = 0.0;
  }
  r[15] = d.pres[k];
  r[16] = d.T[k];
}

void launch_pack_xdmf(const Dev &d, hipStream_t st, int cur, float *f, int *iv) {
  k_pack_xdmf<<<nblk(d.n), kBlock, 0, st>>>(d, cur, f, iv);
}

// ---------------------------------------------------------------------------
// Static device scratch of the allocation-free output entry point
// (wf_pack_xdmf_noalloc): kOutScratchBytes per GPU, held from the library's
// first use of a device until the process ends.  Host code never passes the
// array's name to a kernel or a copy: on the host it is only a stand-in for
// the symbol.  out_scratch() looks up the array's device address once per
// device (with that device current) and returns it.
static __device__ __align__(256) unsigned char scratch[kOutScratchBytes];

namespace {
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t out_scratch(int device, void **ptr) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int prev = 0;
    hipError_t e = hipGetDevice(&prev);
    if (e != hipSuccess) return e;
    if (prev != device && (e = hipSetDevice(device)) != hipSuccess) return e;
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (prev != device) (void)hipSetDevice(prev);
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = p;
  }
  *ptr = g_scratch_ptr[device];
  return hipSuccess;
}

void launch_pack_csv(const Dev &d, hipStream_t st, int cur, double *rows) {
  k_pack_csv<<<nblk(d.n), kBlock, 0, st>>>(d, cur, rows);
}

}  // namespace wf
