// Device-side data layout and per-element GFN2 math for the MI355X executor.
//
// Everything the kernels touch lives in one device arena described by Topo
// (structure of arrays, ragged batch-major).  Integer offsets are int32 except
// the dense-matrix offsets (int64).  Dense n x n matrices are row-major; the
// symmetric ones are therefore also column-major.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace xtb {

// Systems of at most kMaxAO AOs keep their per-system work vectors (the
// eigensolver's tridiagonal and eigenvalue arrays, the reduction's partial
// sums, the conditioning estimate's vectors, per-atom force accumulators) in
// LDS.  Larger systems, up to kMaxSystemAO, run instantiations of the same
// kernels that keep these vectors in a global slice of kBigPerAO doubles per
// AO plus kBigPerSys per system, at big_offset(); the host launches them over
// the prefix of the size order above kMaxAO.
constexpr int kMaxAO = 1024;
constexpr int kMaxSystemAO = 8192;
constexpr int kBigPerAO = 12;
constexpr int kBigPerSys = 1024;
constexpr int kD4MaxRef = 7;
constexpr double kPi = 3.141592653589793238462643383279502884;

struct Topo {
  int B;
  // per system
  const int* atom_off;      // [B+1]
  const int* shell_off;     // [B+1]
  const int* ao_off;        // [B+1]
  const int* vec_off;       // [B+1]  mixer vector (ns + 9N)
  const int* pc_off;        // [B+1]  point charges
  const int64_t* mat_off;   // [B+1]  dense n*n
  const int64_t* gam_off;   // [B+1]  shell-pair ns*ns
  const int64_t* apair_off; // [B+1]  ordered atom pairs N*N
  const double* nalpha;     // [B]
  const double* nbeta;      // [B]
  // per atom
  const int* Z;
  const double* xyz;        // [3N]
  const int* atom_sys;      // [N]
  const int* atom_sh0;      // [N+1] global first shell of the atom
  const double* at_cnrad;   // GFN2 CN radius 4/3*r_cov (bohr)
  const double* at_h0rad;   // H0 polynomial radius (bohr)
  const double* at_en;      // GFN2 electronegativity
  const double* at_dk;      // AES2 dipole kernel
  const double* at_qk;      // AES2 quadrupole kernel
  const double* at_mprad;   // AES2 multipole radius
  const double* at_mpvcn;   // AES2 valence CN
  const double* at_sqrt_arep;
  const double* at_zeff;
  const double* at_d4rcov;
  const double* at_d4en;
  const double* at_d4zeff;
  const double* at_d4hard;
  const double* at_d4r4r2;
  const int* at_d4ref;      // first global reference
  const int* at_d4nref;
  const int* perm;          // [N] atoms of each system grouped by D4 element (stable)
  // per shell (global index)
  const int* sh_atom;       // global atom
  const int* sh_l;
  const int* sh_ao;         // global first AO
  const int* sh_prim;       // first primitive
  const int* sh_nprim;
  const double* sh_level;   // Hartree
  const double* sh_kcn;     // Hartree
  const double* sh_kpoly;
  const double* sh_zeta;
  const double* sh_refocc;
  const double* sh_eta;     // ES2 hardness
  const double* sh_gam3;    // ES3 Gamma3
  // per AO (global)
  const int* ao_sh;
  const int* ao_atom;
  // primitives
  const double* alpha;
  const double* coef;
  // point charges
  const double* pc_xyz;
  const double* pc_q;
  const double* pc_gam;
  // uniform electric field per system [3B] (Hartree / (e bohr)), or null when
  // no system of the call carries one
  const double* field;
  // external charge operator: b per atom [N] and A per system at apair_off
  // (row-major n_atoms^2), each null when the call has none
  const double* pshift;
  const double* presp;
  // unrestricted systems: each runs as two systems, its alpha channel (the
  // batch member) and a beta copy appended after every member; partner[s] is
  // the other one (-1 for a restricted system).  The alpha system's SCC
  // state holds the charge block [qsh, dip, quad] and then the magnetization
  // block.  sh_wspin[3 sh + k] couples shell sh to the k-th shell of its atom
  // (spin constants W_ll').  Both null when the call has no such system.
  const int* partner;
  const double* sh_wspin;
  // global tables
  const double* d4_c6;      // [262*262]
  const double* d4_refcn;
  const double* d4_refq;
  const int* d4_refng;
  const double* kpair;      // [9]
};

// First double of system s's slice of the global work vectors (systems above
// kMaxAO AOs only; the slice holds kBigPerAO n + kBigPerSys doubles)
__device__ __forceinline__ int64_t big_offset(const Topo& t, int s) {
  return (int64_t)kBigPerAO * t.ao_off[s] + (int64_t)kBigPerSys * s;
}

// Per-system scalar state of the SCC driver (device resident).
struct SysState {
  int iterations;
  int status;        // xtbloom status code
  int converged;
  int kocc;          // orbitals with nonzero occupation
  double f_old;      // free energy of the previous iteration
  double free_energy;
  double entropy;
  double e_core, e_es2, e_es3, e_aes2, e_d4, e_pc;
  double residual_rms;
};

// Systems are visited in a size-sorted order (largest first) so that the
// long-running blocks start first; every per-system kernel maps
// blockIdx -> order[blockIdx].
__device__ __forceinline__ double logistic(double x) {
  if (x >= 0.0) {
    double e = exp(-x);
    return 1.0 / (1.0 + e);
  }
  double e = exp(x);
  return e / (1.0 + e);
}

// GFN2 double-exponential coordination count and its radial derivative
// (reference gfn2/coordination.cpp:170-188).
__device__ __forceinline__ void dexp_count(double r, double rc, double& val, double& der) {
  double f1 = logistic(10.0 * (rc / r - 1.0));
  double f2 = logistic(20.0 * ((rc + 2.0) / r - 1.0));
  val = f1 * f2;
  der = -(1.0 / (r * r)) *
        (10.0 * rc * f1 * (1.0 - f1) * f2 + 20.0 * (rc + 2.0) * f2 * (1.0 - f2) * f1);
}

// AES2 CN-dependent multipole radius (reference gfn2/aes2.cpp:366-383).
__device__ __forceinline__ double mp_radius(const Topo& t, int a, double cn, double* dr) {
  const double rad = t.at_mprad[a];
  const double f = logistic(4.0 * (cn - t.at_mpvcn[a] - 1.2));
  if (dr) *dr = (5.0 - rad) * 4.0 * f * (1.0 - f);
  return rad + (5.0 - rad) * f;
}

// AES2 damped inverse powers (reference gfn2/aes2.cpp:385-398).
__device__ __forceinline__ void aes2_kernels(double r, double rho, double& k3, double& k5) {
  double ir = 1.0 / r, ir3 = ir * ir * ir, ir5 = ir3 * ir * ir, sc = rho * ir;
  double sc3 = sc * sc * sc;
  k3 = ir3 / (1.0 + 6.0 * sc3);
  k5 = ir5 / (1.0 + 6.0 * sc3 * sc);
}

// D4 charge scaling zeta(q) and d zeta/dq (reference gfn2/d4.cpp:283-296).
__device__ __forceinline__ double zeta_scale(double c, double qref, double qmod) {
  if (qmod < 0.0) return exp(3.0);
  return exp(3.0 * (1.0 - exp(c * (1.0 - qref / qmod))));
}
__device__ __forceinline__ double zeta_scale_dq(double c, double qref, double qmod) {
  if (qmod < 0.0) return 0.0;
  double inner = exp(c * (1.0 - qref / qmod));
  return -3.0 * c * inner * zeta_scale(c, qref, qmod) * qref / (qmod * qmod);
}

// D4 reference weights of one atom (reference gfn2/d4.cpp:298-366).
__device__ inline void d4_weights(const Topo& t, int a, double cn, double q, double* w,
                                  double* dwdcn, double* dwdq) {
  const int off = t.at_d4ref[a], nref = t.at_d4nref[a];
  double norm = 0.0, dnorm = 0.0, maxcn = -1.0e300;
  for (int r = 0; r < nref; ++r) {
    const double rc = t.d4_refcn[off + r];
    maxcn = fmax(maxcn, rc);
    for (int k = 1; k <= t.d4_refng[off + r]; ++k) {
      double f = 6.0 * k, dl = cn - rc, v = exp(-f * dl * dl);
      norm += v;
      dnorm += 2.0 * f * (rc - cn) * v;
    }
  }
  const double inorm = fabs(norm) > 1.4916681462400413e-154 ? 1.0 / norm : 0.0;
  const double qmod = q + t.at_d4zeff[a], c = t.at_d4hard[a] * 2.0;
  for (int r = 0; r < kD4MaxRef; ++r) {
    w[r] = 0.0;
    if (dwdcn) dwdcn[r] = 0.0;
    if (dwdq) dwdq[r] = 0.0;
  }
  for (int r = 0; r < nref; ++r) {
    const double rc = t.d4_refcn[off + r];
    double num = 0.0, dnum = 0.0;
    for (int k = 1; k <= t.d4_refng[off + r]; ++k) {
      double f = 6.0 * k, dl = cn - rc, v = exp(-f * dl * dl);
      num += v;
      dnum += 2.0 * f * (rc - cn) * v;
    }
    double cw = num * inorm;
    if (!isfinite(cw) || inorm == 0.0) cw = fabs(maxcn - rc) < 1e-12 ? 1.0 : 0.0;
    double dcw = inorm * (dnum - num * dnorm * inorm);
    if (!isfinite(dcw) || inorm == 0.0) dcw = 0.0;
    const double qref = t.d4_refq[off + r] + t.at_d4zeff[a];
    const double z = zeta_scale(c, qref, qmod);
    w[r] = cw * z;
    if (dwdcn) dwdcn[r] = dcw * z;
    if (dwdq) dwdq[r] = cw * zeta_scale_dq(c, qref, qmod);
  }
}

// sum_rs wa_r wb_s C6ref[r][s]
__device__ __forceinline__ double d4_contract(const Topo& t, int a, int b, const double* wa,
                                              const double* wb) {
  const int oa = t.at_d4ref[a], na = t.at_d4nref[a];
  const int ob = t.at_d4ref[b], nb = t.at_d4nref[b];
  double c = 0.0;
  for (int r = 0; r < na; ++r) {
    const double* row = t.d4_c6 + (size_t)(oa + r) * 262 + ob;
    double s = 0.0;
    for (int k = 0; k < nb; ++k) s += wb[k] * row[k];
    c += wa[r] * s;
  }
  return c;
}

// D4 rational damping of one pair (reference gfn2/d4.cpp:774-786): value and dD/d(r^2)*2.
__device__ __forceinline__ void d4_damping(const Topo& t, int a, int b, double r2, double& damp,
                                           double& ddamp) {
  damp = 0.0;
  ddamp = 0.0;
  if (r2 > 2500.0) return;
  const double rrij = 3.0 * t.at_d4r4r2[a] * t.at_d4r4r2[b];
  const double r0 = 0.52 * sqrt(rrij) + 5.0;
  const double r4 = r2 * r2, r6 = r4 * r2, r02 = r0 * r0, r04 = r02 * r02, r06 = r04 * r02;
  const double t6 = 1.0 / (r6 + r06), t8 = 1.0 / (r4 * r4 + r04 * r04);
  damp = t6 + 2.7 * rrij * t8;
  ddamp = -6.0 * r4 * t6 * t6 + 2.7 * rrij * (-8.0 * r6 * t8 * t8);
}

__device__ __forceinline__ double d4_r0(const Topo& t, int a, int b) {
  return 0.52 * sqrt(3.0 * t.at_d4r4r2[a] * t.at_d4r4r2[b]) + 5.0;
}

// D4 electronegativity-weighted erf coordination count (reference gfn2/d4.cpp:758-767).
__device__ __forceinline__ void d4_cn_count(const Topo& t, int a, int b, double r, double& val,
                                            double& der) {
  const double rc = t.at_d4rcov[a] + t.at_d4rcov[b];
  const double den = fabs(t.at_d4en[a] - t.at_d4en[b]);
  const double x5 = den + 19.08857;
  const double enf = 4.10451 * exp(-(x5 * x5) / (2.0 * 11.28174 * 11.28174));
  const double ex = 7.5 * (r - rc) / rc;
  val = 0.5 * enf * (1.0 + erf(-ex));
  der = -enf * 7.5 * exp(-ex * ex) * 0.5641895835477562869480794515607726 / rc;
}

// ---------------------------------------------------------------------------
// Wavefront / block reductions (64-lane wavefronts) built on DPP moves: two
// quad permutations and two row rotations leave every lane with its 16-lane row
// sum, row broadcasts 15 and 31 accumulate the rows into lane 63, which is then
// read back as a wave-uniform value.  update_dpp with old = 0 makes lanes
// without a source (or outside the row mask) contribute zero.
// ---------------------------------------------------------------------------
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_move(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROW_MASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_move<0xb1, 0xf>(v);   // quad_perm [1,0,3,2]
  v += dpp_move<0x4e, 0xf>(v);   // quad_perm [2,3,0,1]
  v += dpp_move<0x124, 0xf>(v);  // row_ror:4
  v += dpp_move<0x128, 0xf>(v);  // row_ror:8
  v += dpp_move<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_move<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, 63);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Block-wide sum for blockDim.x <= 1024; red must hold blockDim.x/64 doubles.
__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < nw; ++i) s += red[i];
  __syncthreads();
  return s;
}

// K sums at once; red holds at least K * (blockDim/64) doubles.  Three barriers
// for all K values instead of three per value.
template <int K>
__device__ __forceinline__ void block_sum_n(double (&v)[K], double* red) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) red[k * nw + wid] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double s = 0.0;
    for (int i = 0; i < nw; ++i) s += red[k * nw + i];
    v[k] = s;
  }
  __syncthreads();
}

}  // namespace xtb
