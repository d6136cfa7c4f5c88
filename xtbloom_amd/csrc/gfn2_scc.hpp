// Geometry, integral, SCC-iteration kernels of the MI355X GFN2 executor.
#pragma once

#include "gfn2_device.hpp"
#include "gfn2_integrals.hpp"

namespace xtb {



// Mixer arrays, one slice per system at vec_off (vectors) / mem*vec_off (history).
struct MixerArrays {
  double* x;         // current SCC state used to build potentials ([qsh, dip, quad])
  double* raw;       // Mulliken output of the latest iteration
  double* cur;       // Broyden current input
  double* prev;
  double* prev_res;
  double* df;        // [mem * dim]
  double* u;         // [mem * dim]
  double* omega;     // [B * mem]
  double* ovl;       // [B * mem * mem]  <df_r|df_c> by slot pair
  int mem;
  double damping;
  double qtol, etol, kt;
  int max_iter;
};

// ---------------------------------------------------------------------------
// geometry: CN (GFN2 dexp), D4 CN, AES2 radii, ES2 Gamma, point-charge shell
// potential, FRESH SAD state and mixer reset.  One block per system.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_geometry(Topo t, const int* order, double* cn, double* cnd4,
                                                  double* mrad, double* dmrad, double* gam,
                                                  double* vpc, MixerArrays mx, SysState* st,
                                                  const double* charges) {
  const int s = order[blockIdx.x];
  const int a0 = t.atom_off[s], a1 = t.atom_off[s + 1], N = a1 - a0;
  const int s0 = t.shell_off[s], s1 = t.shell_off[s + 1], ns = s1 - s0;
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int a = a0 + tid; a < a1; a += nt) {
    double c = 0.0, c4 = 0.0;
    const double* ra = t.xyz + 3 * a;
    for (int b = a0; b < a1; ++b) {
      if (b == a) continue;
      const double* rb = t.xyz + 3 * b;
      const double dx = ra[0] - rb[0], dy = ra[1] - rb[1], dz = ra[2] - rb[2];
      const double r2 = dx * dx + dy * dy + dz * dz;
      const double r = sqrt(r2);
      if (r2 <= 625.0) {
        double val, der;
        dexp_count(r, t.at_cnrad[a] + t.at_cnrad[b], val, der);
        c += val;
      }
      if (r2 <= 900.0) {
        double val, der;
        d4_cn_count(t, a, b, r, val, der);
        c4 += val;
      }
    }
    cn[a] = c;
    cnd4[a] = c4;
    double dr;
    mrad[a] = mp_radius(t, a, c, &dr);
    dmrad[a] = dr;
  }
  // ES2 Gamma (reference gfn2/es2.cpp:211-300): arithmetic hardness average
  double* G = gam + t.gam_off[s];
  for (int idx = tid; idx < ns * ns; idx += nt) {
    const int i = idx / ns, j = idx % ns;
    const int si = s0 + i, sj = s0 + j;
    const int A = t.sh_atom[si], C = t.sh_atom[sj];
    const double eta = 0.5 * (t.sh_eta[si] + t.sh_eta[sj]);
    if (A == C) {
      G[idx] = eta;
    } else {
      const double dx = t.xyz[3 * A] - t.xyz[3 * C], dy = t.xyz[3 * A + 1] - t.xyz[3 * C + 1],
                   dz = t.xyz[3 * A + 2] - t.xyz[3 * C + 2];
      G[idx] = 1.0 / sqrt(dx * dx + dy * dy + dz * dz + 1.0 / (eta * eta));
    }
  }
  // point-charge shell potential (reference external_point_charges.hpp:59)
  // and the scalar potential -E.R_A of a uniform field on the charge of the
  // shell's atom (reference gfn2_cpu_execution.cpp:1294-1302)
  const int p0 = t.pc_off[s], p1 = t.pc_off[s + 1];
  for (int sh = s0 + tid; sh < s1; sh += nt) {
    const int A = t.sh_atom[sh];
    double v = 0.0;
    for (int p = p0; p < p1; ++p) {
      const double dx = t.xyz[3 * A] - t.pc_xyz[3 * p], dy = t.xyz[3 * A + 1] - t.pc_xyz[3 * p + 1],
                   dz = t.xyz[3 * A + 2] - t.pc_xyz[3 * p + 2];
      const double ia = 2.0 / (t.sh_eta[sh] + t.pc_gam[p]);
      v += t.pc_q[p] / sqrt(dx * dx + dy * dy + dz * dz + ia * ia);
    }
    if (t.field) {
      const double* E = t.field + 3 * s;
      v -= E[0] * t.xyz[3 * A] + E[1] * t.xyz[3 * A + 1] + E[2] * t.xyz[3 * A + 2];
    }
    vpc[sh] = v;
  }
  // FRESH state (reference wavefunction.cpp:733-776)
  const int vo = t.vec_off[s], dim = t.vec_off[s + 1] - vo;
  for (int i = tid; i < dim; i += nt) {
    double val = 0.0;
    if (i < ns) {
      const int sh = s0 + i, A = t.sh_atom[sh];
      double ref = 0.0;
      for (int q = t.atom_sh0[A]; q < t.atom_sh0[A + 1]; ++q) ref += t.sh_refocc[q];
      val = (t.sh_refocc[sh] / ref) * (charges[s] / N);
    }
    mx.x[vo + i] = val;
    mx.cur[vo + i] = val;
    mx.prev[vo + i] = 0.0;
    mx.prev_res[vo + i] = 0.0;
  }
  if (tid == 0) {
    SysState z = {};
    z.status = st[s].status;  // keep setup failures
    st[s] = z;
  }
}

// ---------------------------------------------------------------------------
// S, D, Q and H0 of every ordered shell pair, one thread per pair.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double shell_level(const Topo& t, int sh, const double* cn) {
  return t.sh_level[sh] - t.sh_kcn[sh] * cn[t.sh_atom[sh]];
}

__device__ __forceinline__ double h0_pair_scale(const Topo& t, int si, int sj) {
  const double zs = t.sh_zeta[si], zt = t.sh_zeta[sj];
  const double zeta = sqrt(2.0 * sqrt(zs * zt) / (zs + zt));  // wexp = 0.5
  const double den = t.at_en[t.sh_atom[si]] - t.at_en[t.sh_atom[sj]];
  return zeta * (1.0 + 0.02 * den * den) * t.kpair[t.sh_l[si] * 3 + t.sh_l[sj]];
}

// spatial scale and its derivative d/dR (reference h0.cpp:407-440, 537-560)
__device__ __forceinline__ void h0_spatial(const Topo& t, int si, int sj, double r, double& sp,
                                           double& dsp) {
  const int A = t.sh_atom[si], C = t.sh_atom[sj];
  const double rr = sqrt(r / (t.at_h0rad[A] + t.at_h0rad[C]));
  const double ks = t.sh_kpoly[si], kt = t.sh_kpoly[sj];
  const double p1 = 1.0 + ks * rr, p2 = 1.0 + kt * rr;
  const double kp = h0_pair_scale(t, si, sj);
  sp = kp * p1 * p2;
  dsp = kp * (ks * p2 + kt * p1) * rr / (2.0 * r);
}

// D and traceless Q of element (n, m) from those of element (m, n), m on atom
// ar and n on atom ac ([xx, xy, yy, xz, yz, zz])
__device__ __forceinline__ void transposed_moments(const Topo& t, int ar, int ac, double sv,
                                                   const double (&dm)[3], const double (&qm)[6],
                                                   double (&dt)[3], double (&qt)[6]) {
  const double dl[3] = {t.xyz[3 * ac] - t.xyz[3 * ar], t.xyz[3 * ac + 1] - t.xyz[3 * ar + 1],
                        t.xyz[3 * ac + 2] - t.xyz[3 * ar + 2]};
#pragma unroll
  for (int x = 0; x < 3; ++x) dt[x] = dm[x] + dl[x] * sv;
  const double dd = dl[0] * dm[0] + dl[1] * dm[1] + dl[2] * dm[2];
  const double l2 = dl[0] * dl[0] + dl[1] * dl[1] + dl[2] * dl[2];
  const double trc = 0.5 * (2.0 * dd + l2 * sv);
  qt[0] = qm[0] + 1.5 * (2.0 * dl[0] * dm[0] + dl[0] * dl[0] * sv) - trc;
  qt[1] = qm[1] + 1.5 * (dl[0] * dm[1] + dl[1] * dm[0] + dl[0] * dl[1] * sv);
  qt[2] = qm[2] + 1.5 * (2.0 * dl[1] * dm[1] + dl[1] * dl[1] * sv) - trc;
  qt[3] = qm[3] + 1.5 * (dl[0] * dm[2] + dl[2] * dm[0] + dl[0] * dl[2] * sv);
  qt[4] = qm[4] + 1.5 * (dl[1] * dm[2] + dl[2] * dm[1] + dl[1] * dl[2] * sv);
  qt[5] = qm[5] + 1.5 * (2.0 * dl[2] * dm[2] + dl[2] * dl[2] * sv) - trc;
}

// Tile edge of the Fock and Mulliken kernels (kFockTile below).
constexpr int kIntMirrorTile = 32;

template <int LB, int LK>
__device__ void integral_block(const Topo& t, int s, int si, int sj, const double* cn, double* S,
                               double* D, double* Q, double* H0) {
  const int A = t.sh_atom[si], C = t.sh_atom[sj];
  const double v[3] = {t.xyz[3 * C] - t.xyz[3 * A], t.xyz[3 * C + 1] - t.xyz[3 * A + 1],
                       t.xyz[3 * C + 2] - t.xyz[3 * A + 2]};
  PairBlock<LB, LK> blk;
  pair_integrals<LB, LK, true>(t, si, sj, v, blk);
  double f = 0.5 * (shell_level(t, si, cn) + shell_level(t, sj, cn));
  if (A != C) {
    const double r = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    double sp, dsp;
    h0_spatial(t, si, sj, r, sp, dsp);
    f *= sp;
  }
  const int n = t.ao_off[s + 1] - t.ao_off[s];
  const int64_t nn = (int64_t)n * n;
  const int64_t mo = t.mat_off[s];
  const int r0 = t.sh_ao[si] - t.ao_off[s], c0 = t.sh_ao[sj] - t.ao_off[s];
#pragma unroll
  for (int u = 0; u < 2 * LB + 1; ++u)
#pragma unroll
    for (int w = 0; w < 2 * LK + 1; ++w) {
      const int64_t idx = mo + (int64_t)(r0 + u) * n + (c0 + w);
      // S is mirrored from the bra <= ket triangle so it is exactly symmetric
      const double sval = (si <= sj) ? blk.s[u][w] : 0.0;
      if (si <= sj) {
        S[idx] = sval;
        S[mo + (int64_t)(c0 + w) * n + (r0 + u)] = sval;
        H0[idx] = sval * f;
        H0[mo + (int64_t)(c0 + w) * n + (r0 + u)] = sval * f;
      }
      const int64_t loc = (int64_t)(r0 + u) * n + (c0 + w);
#pragma unroll
      for (int c = 0; c < 3; ++c) D[3 * mo + c * nn + loc] = blk.m[c][u][w];
#pragma unroll
      for (int c = 0; c < 6; ++c) Q[6 * mo + c * nn + loc] = blk.m[3 + c][u][w];
      // the (sj, si) block from the transposition identities (moments about
      // the other atom): D_nm = D_mn + d S_mn and the traceless Q likewise.
      // Only elements inside a diagonal Fock tile are ever read (the Fock and
      // Mulliken kernels take the transposed moments of off-diagonal tiles from
      // the same identities, the force kernel reads forward blocks only), so
      // the other mirrored elements are not stored.
      if (si < sj && ((r0 + u) / kIntMirrorTile) == ((c0 + w) / kIntMirrorTile)) {
        double dm[3], qm[6], dt[3], qt[6];
#pragma unroll
        for (int c = 0; c < 3; ++c) dm[c] = blk.m[c][u][w];
#pragma unroll
        for (int c = 0; c < 6; ++c) qm[c] = blk.m[3 + c][u][w];
        transposed_moments(t, A, C, sval, dm, qm, dt, qt);
        const int64_t tl = (int64_t)(c0 + w) * n + (r0 + u);
#pragma unroll
        for (int c = 0; c < 3; ++c) D[3 * mo + c * nn + tl] = dt[c];
#pragma unroll
        for (int c = 0; c < 6; ++c) Q[6 * mo + c * nn + tl] = qt[c];
      }
    }
}

// A subset of the batch's systems with a prefix of per-system work items;
// kernels find their system by binary search over the prefix.
struct SysList {
  const int* sys;      // [count] system indices
  const int64_t* off;  // [count + 1] work-item prefix
  int count;
};

__device__ __forceinline__ int syslist_find(const SysList& l, int64_t g, int64_t& loc) {
  int lo = 0, hi = l.count;  // off[lo] <= g < off[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (l.off[mid] <= g) lo = mid; else hi = mid;
  }
  loc = g - l.off[lo];
  return l.sys[lo];
}

// One thread per unordered shell pair i <= j of a system (ns (ns + 1) / 2
// items, p = j (j + 1) / 2 + i); the block writes both (i, j) and (j, i).
// HAS_D selects the instantiation: systems whose shells are all s/p use the
// smaller one, so they do not pay the register budget of d-shell blocks.
template <bool HAS_D>
__global__ __launch_bounds__(128) void k_integrals(Topo t, SysList list, int64_t total,
                                                   const double* cn, double* S, double* D,
                                                   double* Q, double* H0) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  int64_t loc;
  const int s = syslist_find(list, g, loc);
  int jj = (int)((sqrt(8.0 * (double)loc + 1.0) - 1.0) * 0.5);
  while ((int64_t)jj * (jj + 1) / 2 > loc) --jj;
  while ((int64_t)(jj + 1) * (jj + 2) / 2 <= loc) ++jj;
  const int ii = (int)(loc - (int64_t)jj * (jj + 1) / 2);
  const int si = t.shell_off[s] + ii, sj = t.shell_off[s] + jj;
  const int key = t.sh_l[si] * 3 + t.sh_l[sj];
  if constexpr (!HAS_D) {
    switch (key) {
      case 0: integral_block<0, 0>(t, s, si, sj, cn, S, D, Q, H0); break;
      case 1: integral_block<0, 1>(t, s, si, sj, cn, S, D, Q, H0); break;
      case 3: integral_block<1, 0>(t, s, si, sj, cn, S, D, Q, H0); break;
      default: integral_block<1, 1>(t, s, si, sj, cn, S, D, Q, H0); break;
    }
  } else {
    switch (key) {
      case 0: integral_block<0, 0>(t, s, si, sj, cn, S, D, Q, H0); break;
      case 1: integral_block<0, 1>(t, s, si, sj, cn, S, D, Q, H0); break;
      case 2: integral_block<0, 2>(t, s, si, sj, cn, S, D, Q, H0); break;
      case 3: integral_block<1, 0>(t, s, si, sj, cn, S, D, Q, H0); break;
      case 4: integral_block<1, 1>(t, s, si, sj, cn, S, D, Q, H0); break;
      case 5: integral_block<1, 2>(t, s, si, sj, cn, S, D, Q, H0); break;
      case 6: integral_block<2, 0>(t, s, si, sj, cn, S, D, Q, H0); break;
      case 7: integral_block<2, 1>(t, s, si, sj, cn, S, D, Q, H0); break;
      default: integral_block<2, 2>(t, s, si, sj, cn, S, D, Q, H0); break;
    }
  }
}

// ---------------------------------------------------------------------------
// Atom-pair sums shared by the potentials and the SCC energy: the AES2
// multipole potentials of every atom from all other atoms (reference
// aes2.cpp:451-545) and the D4 two-body contraction (reference
// d4.cpp:896-1032).
//
// One block per system.  The atoms' data are staged as a structure of arrays
// in the system's element-grouped order (Topo::perm): in LDS for systems of at
// most kPairStageAtoms atoms, in a global scratch slice otherwise.  G lanes
// share one atom (G = the largest power of two with G*N <= blockDim, at most
// 64); each takes a contiguous run of partners and the G partial sums are
// combined with xor shuffles inside the wavefront.
//
// D4: sum_b damp_ab w_a^T C(a,b) w_b = w_a^T sum_j C(e_a, j) y_j with
// y_j = sum over partners b of element j of damp_ab w_b.  Partners are visited
// grouped by element, so the 7x7 reference block of C6 is applied once per
// element run instead of once per pair.
// ---------------------------------------------------------------------------
constexpr int kPairStageAtoms = 256;
constexpr int kPairStageDoubles = 22;  // doubles per staged atom
constexpr int kPairStageInts = 2;      // ints per staged atom

struct PairStage {
  double* x;    // [3][N] positions
  double* mr;   // [N] AES2 multipole radius
  double* q;    // [N] atomic charge
  double* dp;   // [3][N] dipole
  double* qp;   // [6][N] traceless quadrupole (xx xy yy xz yz zz)
  double* w;    // [7][N] D4 reference weights
  double* r42;  // [N] D4 <r4>/<r2>
  int* ref;     // [N] first D4 reference of the element
  int* nref;    // [N] number of D4 references
};

__device__ __forceinline__ PairStage pair_stage(double* buf, int* ibuf, int N) {
  PairStage p;
  p.x = buf;
  p.mr = buf + 3 * N;
  p.q = buf + 4 * N;
  p.dp = buf + 5 * N;
  p.qp = buf + 8 * N;
  p.w = buf + 14 * N;
  p.r42 = buf + 21 * N;
  p.ref = ibuf;
  p.nref = ibuf + N;
  return p;
}

// Stage atom p (element-grouped position) of the system: geometry, radius,
// multipoles from the state vector, D4 weights at charge q (written to w4 and
// optionally dw4 = dw/dq as well), and the element's D4 data.
// dip / quad: the system's dipole and quadrupole blocks of the state vector;
// la: the atom's index inside the system.
__device__ __forceinline__ void pair_stage_atom(const Topo& t, const PairStage& S, int N, int p,
                                                int a, int la, const double* dip,
                                                const double* quad, double q, const double* cnd4,
                                                const double* mrad, double* w4, double* dw4) {
  for (int c = 0; c < 3; ++c) {
    S.x[c * N + p] = t.xyz[3 * a + c];
    S.dp[c * N + p] = dip[3 * la + c];
  }
  for (int c = 0; c < 6; ++c) S.qp[c * N + p] = quad[6 * la + c];
  S.mr[p] = mrad[a];
  S.q[p] = q;
  S.r42[p] = t.at_d4r4r2[a];
  S.ref[p] = t.at_d4ref[a];
  S.nref[p] = t.at_d4nref[a];
  d4_weights(t, a, cnd4[a], q, w4 + kD4MaxRef * a, nullptr,
             dw4 ? dw4 + kD4MaxRef * a : nullptr);
  for (int r = 0; r < kD4MaxRef; ++r) S.w[r * N + p] = w4[kD4MaxRef * a + r];
}

struct PairSums {
  double pq;      // potential on the atom's charge
  double pd[3];   // potential on its dipole
  double pQ[6];   // potential on its quadrupole (off-diagonals doubled)
  double z[7];    // sum_j C(e_a, j) y_j
};

// Partners pb0 <= pb < pb1 (element-grouped positions) of atom pa.
__device__ __forceinline__ void pair_sums(const Topo& t, const PairStage& S, int N, int pa,
                                          int pb0, int pb1, PairSums& o) {
  const double xa = S.x[pa], ya = S.x[N + pa], za = S.x[2 * N + pa];
  const double mra = S.mr[pa], r42a = S.r42[pa];
  const int oa = S.ref[pa], na = S.nref[pa];
  o.pq = 0.0;
  for (int c = 0; c < 3; ++c) o.pd[c] = 0.0;
  for (int c = 0; c < 6; ++c) o.pQ[c] = 0.0;
  for (int r = 0; r < 7; ++r) o.z[r] = 0.0;
  double y[7];
  for (int k = 0; k < 7; ++k) y[k] = 0.0;
  int cur = -1, curn = 0;
  auto flush = [&]() __attribute__((always_inline)) {
    if (cur < 0) return;
#pragma unroll
    for (int r = 0; r < 7; ++r) {
      if (r < na) {
        const double* row = t.d4_c6 + (size_t)(oa + r) * 262 + cur;
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < 7; ++k)
          if (k < curn) acc = fma(row[k], y[k], acc);
        o.z[r] += acc;
      }
    }
  };
  for (int pb = pb0; pb < pb1; ++pb) {
    const int ob = S.ref[pb];
    if (ob != cur) {
      flush();
      cur = ob;
      curn = S.nref[pb];
#pragma unroll
      for (int k = 0; k < 7; ++k) y[k] = 0.0;
    }
    const double r[3] = {xa - S.x[pb], ya - S.x[N + pb], za - S.x[2 * N + pb]};
    const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    const bool self = pb == pa;
    // AES2: the self "pair" gets zero kernels (r = 0), which keeps the loop
    // body branch-free.
    double k3, k5;
    aes2_kernels(self ? 1.0 : sqrt(r2), 0.5 * (mra + S.mr[pb]), k3, k5);
    k3 = self ? 0.0 : k3;
    k5 = self ? 0.0 : k5;
    const double db[3] = {S.dp[pb], S.dp[N + pb], S.dp[2 * N + pb]};
    const double sq[6] = {r[0] * r[0] * k5, 2.0 * r[0] * r[1] * k5, r[1] * r[1] * k5,
                          2.0 * r[0] * r[2] * k5, 2.0 * r[1] * r[2] * k5, r[2] * r[2] * k5};
    const double rdb = r[0] * db[0] + r[1] * db[1] + r[2] * db[2];
    double sqQ = 0.0;
#pragma unroll
    for (int c = 0; c < 6; ++c) sqQ = fma(sq[c], S.qp[c * N + pb], sqQ);
    const double qb = S.q[pb];
    o.pq += k3 * rdb + sqQ;
#pragma unroll
    for (int c = 0; c < 3; ++c) o.pd[c] += -qb * k3 * r[c] + k5 * (r2 * db[c] - 3.0 * r[c] * rdb);
#pragma unroll
    for (int c = 0; c < 6; ++c) o.pQ[c] += qb * sq[c];
    // D4 rational damping (reference d4.cpp:774-786)
    double damp = 0.0;
    if (!self && r2 <= 2500.0) {
      const double rrij = 3.0 * r42a * S.r42[pb];
      const double r0 = 0.52 * sqrt(rrij) + 5.0;
      const double r4 = r2 * r2, r6 = r4 * r2, r02 = r0 * r0, r04 = r02 * r02, r06 = r04 * r02;
      damp = 1.0 / (r6 + r06) + 2.7 * rrij / (r4 * r4 + r04 * r04);
    }
#pragma unroll
    for (int k = 0; k < 7; ++k) y[k] = fma(damp, S.w[k * N + pb], y[k]);
  }
  flush();
}

// Sum the PairSums of the G consecutive lanes that share one atom.
__device__ __forceinline__ void pair_sums_reduce(PairSums& o, int G) {
  for (int off = G >> 1; off > 0; off >>= 1) {
    o.pq += __shfl_xor(o.pq, off);
#pragma unroll
    for (int c = 0; c < 3; ++c) o.pd[c] += __shfl_xor(o.pd[c], off);
#pragma unroll
    for (int c = 0; c < 6; ++c) o.pQ[c] += __shfl_xor(o.pQ[c], off);
#pragma unroll
    for (int r = 0; r < 7; ++r) o.z[r] += __shfl_xor(o.z[r], off);
  }
}

__device__ __forceinline__ int pair_lanes(int N, int nt) {
  int G = 1;
  while (G < 64 && 2 * G * N <= nt) G *= 2;
  return G;
}

// Potentials from the state x: v_sh (with the atom potential folded in), v_D, v_Q
// (reference scc_driver.cpp:2363-2560, aes2.cpp:451-545, d4.cpp:896-1032).
// STAGED = the instantiation for systems of at most kPairStageAtoms atoms; the
// other one serves the rest from the global staging slice.
// Bounded to 128 VGPRs (four waves per SIMD): two blocks share a CU, so the
// whole batch runs in one wave of blocks (a few spilled values cost less).
template <bool STAGED>
__global__ __launch_bounds__(512, 4) void k_potentials(Topo t, const int* order,
                                                    const unsigned char* active, const double* xs,
                                                    const double* cnd4, const double* mrad,
                                                    const double* gam, const double* vpc,
                                                    double* qat, double* w4, double* dw4,
                                                    double* vat, double* vsh, double* vd,
                                                    double* vqq, double* gstage, int* gstagei) {
  const int s = order[blockIdx.x];
  if (active && !active[s]) return;
  const int a0 = t.atom_off[s], a1 = t.atom_off[s + 1], N = a1 - a0;
  if (STAGED != (N <= kPairStageAtoms)) return;
  const int s0 = t.shell_off[s], s1 = t.shell_off[s + 1], ns = s1 - s0;
  const double* x = xs + t.vec_off[s];
  const double* qsh = x;
  const int tid = threadIdx.x, nt = blockDim.x;
  __shared__ double lbuf[STAGED ? kPairStageDoubles * kPairStageAtoms : 1];
  __shared__ int libuf[STAGED ? kPairStageInts * kPairStageAtoms : 1];
  double* buf = STAGED ? lbuf : gstage + (size_t)kPairStageDoubles * a0;
  int* ibuf = STAGED ? libuf : gstagei + (size_t)kPairStageInts * a0;
  const PairStage S = pair_stage(buf, ibuf, N);
  for (int p = tid; p < N; p += nt) {
    const int a = t.perm[a0 + p];
    double q = 0.0;
    for (int sh = t.atom_sh0[a]; sh < t.atom_sh0[a + 1]; ++sh) q += qsh[sh - s0];
    qat[a] = q;
    pair_stage_atom(t, S, N, p, a, a - a0, x + ns, x + ns + 3 * N, q, cnd4, mrad, w4, dw4);
  }
  __syncthreads();
  // ES2 + ES3 + point charges; Gamma is symmetric, so column i is read
  // (coalesced across the block) in place of row i.
  const double* Gm = gam + t.gam_off[s];
  for (int i = tid; i < ns; i += nt) {
    double v = 0.0;
    for (int j = 0; j < ns; ++j) v += Gm[j * ns + i] * qsh[j];
    const int sh = s0 + i;
    v += t.sh_gam3[sh] * qsh[i] * qsh[i];
    v += vpc[sh];
    vsh[sh] = v;
  }
  const double qscale[6] = {1.0, 2.0, 1.0, 2.0, 2.0, 1.0};
  const int G = pair_lanes(N, nt), g = tid & (G - 1), per = nt / G;
  const int L = (N + G - 1) / G;
  for (int base = 0; base < N; base += per) {
    const int pa = base + tid / G;
    if (pa < N) {
      PairSums o;
      pair_sums(t, S, N, pa, min(N, g * L), min(N, (g + 1) * L), o);
      pair_sums_reduce(o, G);
      if (g == 0) {
        const int a = t.perm[a0 + pa];
        double pd4 = 0.0;
        for (int r = 0; r < kD4MaxRef; ++r) pd4 -= dw4[kD4MaxRef * a + r] * o.z[r];
        // external charge operator b_A + sum_B A_AB q_B, before D4 as in
        // the reference (scc_driver.cpp:2515-2531); qat of the whole system
        // was written before the barrier above
        double vop = 0.0;
        if (t.presp) {
          const double* row = t.presp + t.apair_off[s] + (int64_t)(a - a0) * N;
          for (int c = 0; c < N; ++c) vop = fma(row[c], qat[a0 + c], vop);
        }
        if (t.pshift) vop += t.pshift[a];
        vat[a] = (o.pq + vop) + pd4;
        // a uniform field adds -E to every atomic dipole potential
        const double* E = t.field ? t.field + 3 * s : nullptr;
        for (int c = 0; c < 3; ++c)
          vd[3 * a + c] = 2.0 * t.at_dk[a] * S.dp[c * N + pa] + o.pd[c] - (E ? E[c] : 0.0);
        for (int c = 0; c < 6; ++c)
          vqq[6 * a + c] = 2.0 * t.at_qk[a] * qscale[c] * S.qp[c * N + pa] + o.pQ[c];
      }
    }
  }
  __syncthreads();
  for (int sh = s0 + tid; sh < s1; sh += nt) vsh[sh] += vat[t.sh_atom[sh]];
}

// H = H0 - S (u_m+u_n)/2 - D_mn.v(A_n)/2 - D_nm.v(A_m)/2 - (Q likewise)   (mulliken.hpp:222-237)
//
// The transposed blocks follow from the element's own S, D, Q (moments about
// the ket atom): with d = R(A_n) - R(A_m),
//   D_nm,a = D_mn,a + d_a S_mn
//   M_nm,ab = M_mn,ab + d_a D_mn,b + d_b D_mn,a + d_a d_b S_mn      (raw second moments)
// and the traceless Q = 3/2 M - 1/2 tr(M) 1.  Element (m, n) therefore needs
// only S, D, Q, H0 at (m, n), and H is symmetric: one block reads a 32 x 32
// tile on or above the diagonal and writes it and, through LDS, its mirror
// below the diagonal, so S, D, Q and H0 are read about once per pair.
struct FockTile {
  int sys;
  int r0, c0;  // tile origin, r0 <= c0
};
constexpr int kFockTile = 32;
static_assert(kFockTile == kIntMirrorTile, "the integral kernel stores the mirrored D/Q of diagonal Fock tiles");

__device__ __forceinline__ double fock_element(const Topo& t, int o, int r, int c, double sv,
                                               const double (&dm)[3], const double (&qm)[6],
                                               double h0, const double* __restrict__ vsh,
                                               const double* __restrict__ vd,
                                               const double* __restrict__ vqq) {
  const int ar = t.ao_atom[o + r], ac = t.ao_atom[o + c];
  double dt[3], qt[6];
  transposed_moments(t, ar, ac, sv, dm, qm, dt, qt);
  double h = -0.5 * sv * (vsh[t.ao_sh[o + r]] + vsh[t.ao_sh[o + c]]);
#pragma unroll
  for (int x = 0; x < 3; ++x) h -= 0.5 * (dm[x] * vd[3 * ac + x] + dt[x] * vd[3 * ar + x]);
#pragma unroll
  for (int x = 0; x < 6; ++x) h -= 0.5 * (qm[x] * vqq[6 * ac + x] + qt[x] * vqq[6 * ar + x]);
  return h0 + h;
}

// One block = one tile (host-built list of the tiles with r0 <= c0).  Thread
// (rg, cl) = (tid / 32, tid % 32) handles rows rg, rg + 8, rg + 16, rg + 24
// of column cl: row segments of 32 doubles are read and written whole.
__global__ __launch_bounds__(256) void k_fock(Topo t, const FockTile* tiles,
                                              const unsigned char* active,
                                              const double* __restrict__ S,
                                              const double* __restrict__ D,
                                              const double* __restrict__ Q,
                                              const double* __restrict__ H0,
                                              const double* __restrict__ vsh,
                                              const double* __restrict__ vd,
                                              const double* __restrict__ vqq,
                                              double* __restrict__ H) {
  const FockTile tl = tiles[blockIdx.x];
  const int s = tl.sys;
  if (active && !active[s]) return;
  const int n = t.ao_off[s + 1] - t.ao_off[s];
  const int o = t.ao_off[s];
  const int64_t mo = t.mat_off[s], nn = (int64_t)n * n;
  const double* Ds = D + 3 * mo;
  const double* Qs = Q + 6 * mo;
  __shared__ double mirror[kFockTile][kFockTile + 1];
  const int cl = threadIdx.x & 31, rg = threadIdx.x >> 5;
  const int c = tl.c0 + cl;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int rl = rg + 8 * i, r = tl.r0 + rl;
    if (r < n && c < n) {
      const int64_t e = (int64_t)r * n + c;
      double dm[3], qm[6];
#pragma unroll
      for (int x = 0; x < 3; ++x) dm[x] = Ds[x * nn + e];
#pragma unroll
      for (int x = 0; x < 6; ++x) qm[x] = Qs[x * nn + e];
      const double h = fock_element(t, o, r, c, S[mo + e], dm, qm, H0[mo + e], vsh, vd, vqq);
      H[mo + e] = h;
      mirror[cl][rl] = h;
    }
  }
  if (tl.r0 == tl.c0) return;  // a diagonal tile holds both halves itself (uniform per block)
  __syncthreads();
  // H(c, r) = H(r, c): row c0 + rg + 8 i of H, columns r0 .. r0 + 31
  const int rc = tl.r0 + cl;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int ml = rg + 8 * i, m = tl.c0 + ml;
    if (m < n && rc < n) H[mo + (int64_t)m * n + rc] = mirror[ml][cl];
  }
}

// ---------------------------------------------------------------------------
// Mulliken raw multipoles and Tr(P H0) (reference mulliken_kernels_impl.hpp:40-115),
// in two deterministic phases:
//   k_mulliken_rows    one block per Fock tile (32 x 32, on or above the
//                      diagonal).  For every column c the partial sums -P.S,
//                      -P.D(3), -P.Q(6) and P.H0 over the 32 rows of one tile
//                      row go to part[off(chunk) + 11 c + k].  Tile (R, C)
//                      gives the partials of its columns over rows R (moments
//                      as stored) and, P being symmetric, those of its rows
//                      over columns C from the transposed moments, so P, S,
//                      D, Q and H0 are read once per pair of elements.
//   k_mulliken_finish  one block per system: chunk partials added in chunk
//                      order, then shell charges, atomic dipoles/quadrupoles
//                      (by ket AO) and the core energy.
// ---------------------------------------------------------------------------
constexpr int kMullRows = 32;
static_assert(kMullRows == kFockTile, "Mulliken chunks are the rows of one Fock tile");

// Sum over the 32 lanes of each half wavefront; lanes 31 and 63 hold the sums.
__device__ __forceinline__ double half_wave_sum(double v) {
  v += dpp_move<0xb1, 0xf>(v);   // quad_perm [1,0,3,2]
  v += dpp_move<0x4e, 0xf>(v);   // quad_perm [2,3,0,1]
  v += dpp_move<0x124, 0xf>(v);  // row_ror:4
  v += dpp_move<0x128, 0xf>(v);  // row_ror:8
  v += dpp_move<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  return v;
}

// Thread (grp, lane) = (tid / 32, tid % 32) holds rows r0 + grp + 8 i
// (i = 0..3) of column c0 + lane.
__global__ __launch_bounds__(256) void k_mulliken_rows(Topo t, const FockTile* tiles,
                                                       const unsigned char* active,
                                                       const int64_t* __restrict__ part_off,
                                                       const double* __restrict__ P,
                                                       const double* __restrict__ S,
                                                       const double* __restrict__ D,
                                                       const double* __restrict__ Q,
                                                       const double* __restrict__ H0,
                                                       double* __restrict__ part) {
  const FockTile tl = tiles[blockIdx.x];
  const int s = tl.sys;
  if (active && !active[s]) return;
  const int n = t.ao_off[s + 1] - t.ao_off[s];
  const int o = t.ao_off[s];
  const int64_t mo = t.mat_off[s], nn = (int64_t)n * n;
  const double* Ds = D + 3 * mo;
  const double* Qs = Q + 6 * mo;
  double* ps = part + part_off[s];
  __shared__ double red[8][11][kMullRows + 1];
  const int lane = threadIdx.x & 31, grp = threadIdx.x >> 5;
  const bool mirror = tl.r0 != tl.c0;  // uniform per block
  const int c = tl.c0 + lane;
  const int ac = c < n ? t.ao_atom[o + c] : 0;
  double col[11];
#pragma unroll
  for (int k = 0; k < 11; ++k) col[k] = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r = tl.r0 + grp + 8 * i;
    const bool in = r < n && c < n;
    double p = 0.0, sv = 0.0, h0 = 0.0, dm[3] = {0.0, 0.0, 0.0}, qm[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (in) {
      const int64_t loc = (int64_t)r * n + c;
      p = P[mo + loc];
      sv = S[mo + loc];
      h0 = H0[mo + loc];
#pragma unroll
      for (int k = 0; k < 3; ++k) dm[k] = Ds[k * nn + loc];
#pragma unroll
      for (int k = 0; k < 6; ++k) qm[k] = Qs[k * nn + loc];
    }
    col[0] = fma(-p, sv, col[0]);
#pragma unroll
    for (int k = 0; k < 3; ++k) col[1 + k] = fma(-p, dm[k], col[1 + k]);
#pragma unroll
    for (int k = 0; k < 6; ++k) col[4 + k] = fma(-p, qm[k], col[4 + k]);
    col[10] = fma(h0, p, col[10]);
    if (mirror) {
      // row r's partial over the tile's 32 columns (the element (c, r) moments)
      double dt[3] = {0.0, 0.0, 0.0}, qt[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      if (in) transposed_moments(t, t.ao_atom[o + r], ac, sv, dm, qm, dt, qt);
      double rowv[11];
      rowv[0] = half_wave_sum(-p * sv);
#pragma unroll
      for (int k = 0; k < 3; ++k) rowv[1 + k] = half_wave_sum(-p * dt[k]);
#pragma unroll
      for (int k = 0; k < 6; ++k) rowv[4 + k] = half_wave_sum(-p * qt[k]);
      rowv[10] = half_wave_sum(h0 * p);
      if (lane == 31 && r < n) {
        double* dst = ps + (int64_t)(tl.c0 / kMullRows) * n * 11 + 11 * (int64_t)r;
#pragma unroll
        for (int k = 0; k < 11; ++k) dst[k] = rowv[k];
      }
    }
  }
  // column partials over the tile's 32 rows: the eight row groups added in order
#pragma unroll
  for (int k = 0; k < 11; ++k) red[grp][k][lane] = col[k];
  __syncthreads();
  for (int q = threadIdx.x; q < 11 * kMullRows; q += blockDim.x) {
    const int k = q / kMullRows, cl = q % kMullRows;
    if (tl.c0 + cl < n) {
      double v = 0.0;
#pragma unroll
      for (int g = 0; g < 8; ++g) v += red[g][k][cl];
      ps[(int64_t)(tl.r0 / kMullRows) * n * 11 + 11 * (int64_t)(tl.c0 + cl) + k] = v;
    }
  }
}

__global__ __launch_bounds__(256) void k_mulliken_finish(Topo t, const int* order,
                                                         const unsigned char* active,
                                                         const int64_t* part_off,
                                                         const double* __restrict__ part,
                                                         double* colsum, double* raw,
                                                         SysState* st) {
  const int s = order[blockIdx.x];
  if (active && !active[s]) return;
  const int n = t.ao_off[s + 1] - t.ao_off[s];
  const int a0 = t.atom_off[s], N = t.atom_off[s + 1] - a0;
  const int s0 = t.shell_off[s], ns = t.shell_off[s + 1] - s0;
  const int nch = (n + kMullRows - 1) / kMullRows;
  const double* ps = part + part_off[s];
  double* cs = colsum + (int64_t)t.ao_off[s] * 10;
  __shared__ double red[16];
  double core = 0.0;
  for (int c = threadIdx.x; c < n; c += blockDim.x) {
    double acc[11];
#pragma unroll
    for (int k = 0; k < 11; ++k) acc[k] = 0.0;
    for (int j = 0; j < nch; ++j) {
      const double* pj = ps + (int64_t)j * n * 11 + 11 * (int64_t)c;
#pragma unroll
      for (int k = 0; k < 11; ++k) acc[k] += pj[k];
    }
#pragma unroll
    for (int k = 0; k < 10; ++k) cs[(int64_t)c * 10 + k] = acc[k];
    core += acc[10];
  }
  core = block_sum(core, red);
  double* rv = raw + t.vec_off[s];
  for (int i = threadIdx.x; i < ns; i += blockDim.x) {
    const int sh = s0 + i;
    double q = t.sh_refocc[sh];
    for (int ao = t.sh_ao[sh]; ao < t.sh_ao[sh] + 2 * t.sh_l[sh] + 1; ++ao)
      q += cs[(int64_t)(ao - t.ao_off[s]) * 10];
    rv[i] = q;
  }
  for (int i = threadIdx.x; i < N * 9; i += blockDim.x) {
    const int la = i / 9, k = i % 9, a = a0 + la;
    double v = 0.0;
    const int aoa = t.sh_ao[t.atom_sh0[a]];
    const int aob = t.sh_ao[t.atom_sh0[a + 1] - 1] + 2 * t.sh_l[t.atom_sh0[a + 1] - 1] + 1;
    for (int ao = aoa; ao < aob; ++ao) v += cs[(int64_t)(ao - t.ao_off[s]) * 10 + 1 + k];
    if (k < 3) rv[ns + 3 * la + k] = v;
    else rv[ns + 3 * N + 6 * la + (k - 3)] = v;
  }
  if (threadIdx.x == 0) st[s].e_core = core;
}

}  // namespace xtb
