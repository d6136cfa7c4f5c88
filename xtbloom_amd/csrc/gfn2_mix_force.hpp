// Energy assembly + modified Broyden mixing, and the analytic force kernels.
#pragma once

#include "gfn2_scc.hpp"

namespace xtb {

// ---------------------------------------------------------------------------
// SCC energy from the raw multipoles (reference scc_driver.cpp:2105-2286),
// modified Broyden step (reference common/scc_mixer.cpp:559-745) and the
// convergence gate rms < qtol and |dF| < etol (scc_driver.cpp:869-893).
// One 512-thread block per active system; the pair energies come from the
// shared pair sums (STAGED as for k_potentials).
// ---------------------------------------------------------------------------
template <bool STAGED>
__global__ __launch_bounds__(512) void k_energy_mix(Topo t, const int* order, unsigned char* active,
                                                    const double* cnd4, const double* mrad,
                                                    const double* gam, const double* vpc,
                                                    double* qat_raw, double* w4, MixerArrays mx,
                                                    SysState* st, int* active_count,
                                                    double* gstage, int* gstagei,
                                                    const int* round_ptr) {
  const int s = order[blockIdx.x];
  if (!active[s]) return;
  // the beta copy of an unrestricted system has no state of its own: its
  // alpha system mixes both channels and sets its flags
  if (t.partner && t.partner[s] >= 0 && t.partner[s] < s) return;
  const int a0 = t.atom_off[s], a1 = t.atom_off[s + 1], N = a1 - a0;
  if (STAGED != (N <= kPairStageAtoms)) return;
  const int s0 = t.shell_off[s], ns = t.shell_off[s + 1] - s0;
  const int vo = t.vec_off[s], dim = t.vec_off[s + 1] - vo;
  const double* raw = mx.raw + vo;
  const double* qsh = raw;
  const double* dip = raw + ns;
  const double* quad = raw + ns + 3 * N;
  const int tid = threadIdx.x, nt = blockDim.x;
  __shared__ double red[64];  // block_sum_n: K values per wavefront
  // the atom staging area while the pair sums run, the Broyden beta matrix after
  __shared__ double lbuf[STAGED ? kPairStageDoubles * kPairStageAtoms : 64 * 64];
  __shared__ int libuf[STAGED ? kPairStageInts * kPairStageAtoms : 1];
  __shared__ double coef[64];
  __shared__ int slots[64];
  __shared__ double sh_scalar[4];
  double* buf = STAGED ? lbuf : gstage + (size_t)kPairStageDoubles * a0;
  int* ibuf = STAGED ? libuf : gstagei + (size_t)kPairStageInts * a0;
  const PairStage S = pair_stage(buf, ibuf, N);
  for (int p = tid; p < N; p += nt) {
    const int a = t.perm[a0 + p];
    double q = 0.0;
    for (int sh = t.atom_sh0[a]; sh < t.atom_sh0[a + 1]; ++sh) q += qsh[sh - s0];
    qat_raw[a] = q;
    pair_stage_atom(t, S, N, p, a, a - a0, dip, quad, q, cnd4, mrad, w4, nullptr);
  }
  __syncthreads();
  // ES2 / ES3 / point charges (Gamma symmetric: coalesced column reads)
  const double* Gm = gam + t.gam_off[s];
  double e[5] = {0.0, 0.0, 0.0, 0.0, 0.0};  // es2 es3 pc aes2 d4
  for (int i = tid; i < ns; i += nt) {
    double v = 0.0;
    for (int j = 0; j < ns; ++j) v += Gm[j * ns + i] * qsh[j];
    const int sh = s0 + i;
    e[0] += 0.5 * qsh[i] * v;
    e[1] += t.sh_gam3[sh] * qsh[i] * qsh[i] * qsh[i] / 3.0;
    e[2] = fma(qsh[i], vpc[sh], e[2]);
  }
  // a uniform field's dipole term -E.mu_A joins the external energy (its
  // charge term -q_A E.R_A is part of vpc)
  if (t.field) {
    const double* E = t.field + 3 * s;
    for (int i = tid; i < 3 * N; i += nt) e[2] -= E[i % 3] * dip[i];
  }
  // spin polarization m.W.m / 2 of the raw magnetization (reference
  // spin.cpp:135-172), kept with AES2 in e[3]
  if (t.partner && t.partner[s] > s) {
    const double* m = raw + ns + 9 * N;
    for (int i = tid; i < ns; i += nt) {
      const int sh = s0 + i, f = t.atom_sh0[t.sh_atom[sh]] - s0;
      const int cnt = t.atom_sh0[t.sh_atom[sh] + 1] - s0 - f;
      double v = 0.0;
      for (int k = 0; k < cnt; ++k) v = fma(t.sh_wspin[3 * sh + k], m[f + k], v);
      e[3] = fma(0.5 * m[i], v, e[3]);
    }
  }
  // external charge operator q.b + q.(A q)/2 at the raw atomic charges
  // (reference periodic_embedding.cpp:352-434)
  if (t.pshift || t.presp) {
    for (int la = tid; la < N; la += nt) {
      double r = 0.0;
      if (t.presp) {
        const double* row = t.presp + t.apair_off[s] + (int64_t)la * N;
        for (int c = 0; c < N; ++c) r = fma(row[c], qat_raw[a0 + c], r);
      }
      const double b = t.pshift ? t.pshift[a0 + la] : 0.0;
      e[2] = fma(qat_raw[a0 + la], fma(0.5, r, b), e[2]);
    }
  }
  // AES2 and D4 two-body: every ordered pair once, halved.  The AES2 pair
  // energy equals q_a V_q + d_a.V_d + Q_a.V_Q of atom a in the field of b,
  // symmetric in (a, b).
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
        double d2 = 0.0, q2 = 0.0, pair = S.q[pa] * o.pq, ew = 0.0;
        for (int c = 0; c < 3; ++c) {
          const double dc = S.dp[c * N + pa];
          d2 += dc * dc;
          pair += dc * o.pd[c];
        }
        for (int c = 0; c < 6; ++c) {
          const double qc = S.qp[c * N + pa];
          q2 += qscale[c] * qc * qc;
          pair += qc * o.pQ[c];
        }
        for (int r = 0; r < kD4MaxRef; ++r) ew += S.w[r * N + pa] * o.z[r];
        e[3] += t.at_dk[a] * d2 + t.at_qk[a] * q2 + 0.5 * pair;
        e[4] -= 0.5 * ew;
      }
    }
  }
  block_sum_n<5>(e, red);
  const double e2 = e[0], e3 = e[1], epc = e[2], eaes = e[3], ed4 = e[4];
  const double eint = st[s].e_core + e2 + e3 + eaes + ed4 + epc;
  const double F = fma(-mx.kt, st[s].entropy, eint);

  // ---- modified Broyden ----
  double* cur = mx.cur + vo;
  double* prev = mx.prev + vo;
  double* prev_res = mx.prev_res + vo;
  const int mem = mx.mem;
  double* dfh = mx.df + (int64_t)mem * vo;  // slot k at dfh + k*dim
  double* uh = mx.u + (int64_t)mem * vo;
  double* om = mx.omega + (int64_t)mem * s;
  const int old_it = st[s].iterations;
  double* beta = lbuf;
  // |res|^2 and |res - res_prev|^2 in one pass
  double rr[2] = {0.0, 0.0};
  for (int i = tid; i < dim; i += nt) {
    const double res = raw[i] - cur[i];
    const double v = res - prev_res[i];
    rr[0] += res * res;
    rr[1] += v * v;
  }
  block_sum_n<2>(rr, red);
  const double rnorm = sqrt(rr[0]), rms = rnorm / sqrt((double)dim);
  double* xnew = mx.x + vo;
  if (old_it == 0) {
    for (int i = tid; i < dim; i += nt) {
      const double res = raw[i] - cur[i];
      xnew[i] = cur[i] + mx.damping * res;
      prev[i] = cur[i];
      prev_res[i] = res;
    }
  } else {
    // delta F (normalized) and u into the new slot
    const int slot_new = (old_it - 1) % mem;
    double* dfn = dfh + (int64_t)slot_new * dim;
    double* un = uh + (int64_t)slot_new * dim;
    const double inv = 1.0 / fmax(sqrt(rr[1]), 2.220446049250313e-16);
    for (int i = tid; i < dim; i += nt) {
      const double res = raw[i] - cur[i];
      const double v = (res - prev_res[i]) * inv;
      dfn[i] = v;
      un[i] = mx.damping * v + inv * (cur[i] - prev[i]);
    }
    const double omega = rnorm > 0.01 / 100000.0 ? fmax(1.0, 0.01 / rnorm) : 100000.0;
    const int hc = min(mem, old_it);
    const int first = old_it - hc + 1;
    if (tid < hc) slots[tid] = (first + tid - 1) % mem;
    if (tid == 0) om[slot_new] = omega;
    __syncthreads();
    // 2*hc dot products, one wavefront each: coef_r = omega_r <df_r|res> and
    // the new slot's overlaps with every slot of the window.  Overlaps among
    // older slots are kept from earlier rounds (a slot's df never changes).
    double* ov = mx.ovl + (int64_t)mem * mem * s;
    {
      const int lane = tid & 63, wid = tid >> 6, nw = nt >> 6;
      for (int task = wid; task < 2 * hc; task += nw) {
        const int r = task < hc ? task : task - hc;
        const double* dr = dfh + (int64_t)slots[r] * dim;
        double dot = 0.0;
        if (task < hc) {
          for (int i = lane; i < dim; i += 64) dot = fma(dr[i], raw[i] - cur[i], dot);
        } else {
          for (int i = lane; i < dim; i += 64) dot = fma(dr[i], dfn[i], dot);
        }
        dot = wave_sum(dot);
        if (lane == 0) {
          if (task < hc) {
            coef[r] = om[slots[r]] * dot;
          } else {
            ov[slot_new * mem + slots[r]] = dot;
            ov[slots[r] * mem + slot_new] = dot;
          }
        }
      }
    }
    __syncthreads();
    if (tid == 0) {
      for (int r = 0; r < hc; ++r)
        for (int c = 0; c <= r; ++c) {
          double val = om[slots[r]] * om[slots[c]] * ov[slots[r] * mem + slots[c]];
          if (r == c) val += 0.01 * 0.01;
          beta[r * hc + c] = val;
          beta[c * hc + r] = val;
        }
    }
    if (tid == 0) {
      int ok = 1;
      for (int r = 0; r < hc && ok; ++r)
        for (int c = 0; c <= r; ++c) {
          double v = beta[r * hc + c];
          for (int k = 0; k < c; ++k) v -= beta[r * hc + k] * beta[c * hc + k];
          if (r == c) {
            if (!(v > 0.0)) { ok = 0; break; }
            beta[r * hc + c] = sqrt(v);
          } else {
            beta[r * hc + c] = v / beta[c * hc + c];
          }
        }
      if (ok) {
        for (int r = 0; r < hc; ++r) {
          double v = coef[r];
          for (int c = 0; c < r; ++c) v -= beta[r * hc + c] * coef[c];
          coef[r] = v / beta[r * hc + r];
        }
        for (int r = hc - 1; r >= 0; --r) {
          double v = coef[r];
          for (int c = r + 1; c < hc; ++c) v -= beta[c * hc + r] * coef[c];
          coef[r] = v / beta[r * hc + r];
        }
      }
      sh_scalar[0] = ok ? 1.0 : 0.0;
    }
    __syncthreads();
    for (int i = tid; i < dim; i += nt) {
      const double res = raw[i] - cur[i];
      double v = cur[i] + mx.damping * res;
      for (int h = 0; h < hc; ++h) v -= om[slots[h]] * coef[h] * uh[(int64_t)slots[h] * dim + i];
      xnew[i] = v;
      prev[i] = cur[i];
      prev_res[i] = res;
    }
    if (tid == 0 && sh_scalar[0] == 0.0) st[s].status = 6;
  }
  __syncthreads();
  const double dF = F - (old_it == 0 ? 0.0 : st[s].f_old);
  const bool conv = rms < mx.qtol && fabs(dF) < mx.etol;
  for (int i = tid; i < dim; i += nt) {
    cur[i] = xnew[i];
    if (conv) xnew[i] = raw[i];
  }
  if (tid == 0) {
    SysState& z = st[s];
    z.e_es2 = e2; z.e_es3 = e3; z.e_aes2 = eaes; z.e_d4 = ed4; z.e_pc = epc;
    z.free_energy = F;
    z.f_old = F;
    z.residual_rms = rms;
    z.iterations = old_it + 1;
    z.converged = conv ? 1 : 0;
    bool stop = conv || z.status != 0;
    if (!conv && z.status == 0 && z.iterations >= mx.max_iter) {
      z.status = 7;
      stop = true;
    }
    if (!isfinite(F) || !isfinite(rms)) { z.status = 6; stop = true; }
    if (stop) active[s] = 0;
    else atomicAdd(round_ptr ? active_count + *round_ptr : active_count, 1);
    // the beta copy follows its alpha system
    if (t.partner && t.partner[s] >= 0) {
      const int sb = t.partner[s];
      if (stop) active[sb] = 0;
      st[sb].converged = z.converged;
      st[sb].iterations = z.iterations;
      st[sb].residual_rms = z.residual_rms;  // the split-precision round choice
    }
  }
}

// ---------------------------------------------------------------------------
// Unrestricted systems (reference scc_driver.cpp:2593-2640, 2699-2720,
// spin.cpp:135-172, force.cpp:506-533).  One block per alpha system; the
// list holds the alpha systems only.
// ---------------------------------------------------------------------------

// magnetization shell potential v_m = W_A m_A from the state's magnetization
// block; the alpha Hamiltonian takes vsh + v_m, the beta copy vsh - v_m and
// the same dipole and quadrupole potentials
__global__ __launch_bounds__(256) void k_spin_potentials(Topo t, const int* alpha, const unsigned char* active,
                                                         const double* xs, double* vsh, double* vd,
                                                         double* vqq) {
  const int s = alpha[blockIdx.x];
  if (active && !active[s]) return;
  const int sb = t.partner[s];
  const int s0 = t.shell_off[s], ns = t.shell_off[s + 1] - s0;
  const int a0 = t.atom_off[s], N = t.atom_off[s + 1] - a0;
  const int b0 = t.atom_off[sb], sb0 = t.shell_off[sb];
  const double* m = xs + t.vec_off[s] + ns + 9 * N;
  for (int i = threadIdx.x; i < ns; i += blockDim.x) {
    const int sh = s0 + i, f = t.atom_sh0[t.sh_atom[sh]] - s0;
    const int cnt = t.atom_sh0[t.sh_atom[sh] + 1] - s0 - f;
    double v = 0.0;
    for (int k = 0; k < cnt; ++k) v = fma(t.sh_wspin[3 * sh + k], m[f + k], v);
    const double q = vsh[sh];
    vsh[sh] = q + v;
    vsh[sb0 + i] = q - v;
  }
  for (int i = threadIdx.x; i < 3 * N; i += blockDim.x) vd[3 * b0 + i] = vd[3 * a0 + i];
  for (int i = threadIdx.x; i < 6 * N; i += blockDim.x) vqq[6 * b0 + i] = vqq[6 * a0 + i];
}

// after the Mulliken pass: the alpha and beta raw multipoles (each with the
// reference occupations in its shell charges) become the charge block
// a + b - n0 and the magnetization block a - b of the alpha state; the core
// energy and the entropy add, and a beta failure fails the system
__global__ __launch_bounds__(256) void k_spin_combine(Topo t, const int* alpha, const unsigned char* active,
                                                      double* raw, SysState* st) {
  const int s = alpha[blockIdx.x];
  if (active && !active[s]) return;
  const int sb = t.partner[s];
  const int s0 = t.shell_off[s], ns = t.shell_off[s + 1] - s0;
  const int N = t.atom_off[s + 1] - t.atom_off[s];
  const int dim = ns + 9 * N;
  double* ra = raw + t.vec_off[s];
  const double* rb = raw + t.vec_off[sb];
  for (int i = threadIdx.x; i < dim; i += blockDim.x) {
    const double a = ra[i], b = rb[i];
    const double ref = i < ns ? t.sh_refocc[s0 + i] : 0.0;
    ra[i] = (a + b) - ref;
    ra[dim + i] = a - b;
  }
  if (threadIdx.x == 0) {
    st[s].e_core += st[sb].e_core;
    st[s].entropy += st[sb].entropy;
    if (st[sb].status != 0 && st[s].status == 0) st[s].status = st[sb].status;
  }
}

// before the forces: P <- Pa + Pb and W <- Wa + Wb + (Pa - Pb)(v_m,m + v_m,n)/2
// in the alpha system's matrices, so that the overlap-derivative term of the
// force kernels carries the magnetization potential; the beta copy is then
// left out of the forces.  BIG: alpha systems above kMaxAO AOs, with the
// shell potentials in their global slice of `big`.
template <bool BIG>
__global__ __launch_bounds__(256) void k_spin_fold(Topo t, const int* alpha, unsigned char* valid,
                                                   const double* xs, double* P, double* W, double* big) {
  const int s = alpha[blockIdx.x];
  const int sb = t.partner[s];
  if (!valid[s]) {
    if (threadIdx.x == 0) valid[sb] = 0;
    return;
  }
  const int s0 = t.shell_off[s], ns = t.shell_off[s + 1] - s0;
  const int N = t.atom_off[s + 1] - t.atom_off[s];
  const int o = t.ao_off[s], n = t.ao_off[s + 1] - o;
  __shared__ double vm_l[BIG ? 1 : kMaxAO];
  double* vm = BIG ? big + big_offset(t, s) : vm_l;
  const double* m = xs + t.vec_off[s] + ns + 9 * N;
  for (int i = threadIdx.x; i < ns; i += blockDim.x) {
    const int sh = s0 + i, f = t.atom_sh0[t.sh_atom[sh]] - s0;
    const int cnt = t.atom_sh0[t.sh_atom[sh] + 1] - s0 - f;
    double v = 0.0;
    for (int k = 0; k < cnt; ++k) v = fma(t.sh_wspin[3 * sh + k], m[f + k], v);
    vm[i] = v;
  }
  __syncthreads();
  double* Pa = P + t.mat_off[s];
  const double* Pb = P + t.mat_off[sb];
  double* Wa = W + t.mat_off[s];
  const double* Wb = W + t.mat_off[sb];
  for (int64_t idx = threadIdx.x; idx < (int64_t)n * n; idx += blockDim.x) {
    const int r = (int)(idx / n), c = (int)(idx % n);
    const double pa = Pa[idx], pb = Pb[idx];
    const double v = vm[t.ao_sh[o + r] - s0] + vm[t.ao_sh[o + c] - s0];
    Pa[idx] = pa + pb;
    Wa[idx] = (Wa[idx] + Wb[idx]) + 0.5 * (pa - pb) * v;
  }
  __syncthreads();
  if (threadIdx.x == 0) valid[sb] = 0;
}

// End of a device-driven SCC round: the next round's number.  active_count
// (indexed by round) was counted by k_energy_mix.
__global__ void k_round_advance(int* round) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *round += 1;
}

// ---------------------------------------------------------------------------
// Forces.  Adjoints of the stationary energy (reference force.cpp:501-612):
//   dE/dS_mn = P_mn (h0 factor_mn - (v_m + v_n)/2) - W_mn
//   dE/dD_mn = -P_mn v_D(atom n),  dE/dQ_mn = -P_mn v_Q(atom n)
// contracted with recomputed derivative integrals per ordered atom pair; the
// H0 CN adjoint and polynomial gradient (h0.cpp:469-597) ride along.
// ---------------------------------------------------------------------------
// Both ordered blocks (si, sj) and (sj, si) of one shell pair, si on atom A
// and sj on atom C.  The (sj, si) block's moments are those of (si, sj) moved
// to the other centre (D' = D + v S and the traceless Q' likewise, v = R_C -
// R_A), so its adjoints fold into the (si, sj) adjoints and one derivative
// sweep serves both blocks; the explicit dependence of D', Q' on v adds
// K (D + v S) - S vd_A per element, with K = -(3 V - tr V) and V the symmetric
// matrix of A's quadrupole potential.  For A == C only the CN chain is left.
template <int LB, int LK>
__device__ void force_block(const Topo& t, int s, int si, int sj, const double* cn,
                            const double* P, const double* W, const double* Sm, const double* Dm,
                            const double* vsh, const double* vd, const double* vqq, double gA[3],
                            double gC[3], double& dcnA, double& dcnC) {
  const int A = t.sh_atom[si], C = t.sh_atom[sj];
  const int n = t.ao_off[s + 1] - t.ao_off[s];
  const int64_t mo = t.mat_off[s];
  const int r0 = t.sh_ao[si] - t.ao_off[s], c0 = t.sh_ao[sj] - t.ao_off[s];
  if (A == C) {
    double wsum = 0.0;
#pragma unroll
    for (int u = 0; u < 2 * LB + 1; ++u)
#pragma unroll
      for (int w = 0; w < 2 * LK + 1; ++w) {
        const int64_t idx = mo + (int64_t)(r0 + u) * n + (c0 + w);
        wsum += P[idx] * Sm[idx];
      }
    const double lw = 0.5 * wsum;
    dcnA -= t.sh_kcn[si] * lw;
    dcnC -= t.sh_kcn[sj] * lw;
    return;
  }
  const double hs = shell_level(t, si, cn), ht = shell_level(t, sj, cn);
  const double avg = 0.5 * (hs + ht);
  double d[3], sp, dsp;
#pragma unroll
  for (int x = 0; x < 3; ++x) d[x] = t.xyz[3 * A + x] - t.xyz[3 * C + x];
  const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  h0_spatial(t, si, sj, r, sp, dsp);
  const double f = avg * sp;
  const double sf = -0.5 * (vsh[si] + vsh[sj]);
  const double v[3] = {-d[0], -d[1], -d[2]};  // R_C - R_A
  const double* vdA = vd + 3 * A;
  const double* vqA = vqq + 6 * A;
  const double trv = vqA[0] + vqA[2] + vqA[5];
  // K = -(3 V - tr V I), V = [[q0, q1/2, q3/2], [q1/2, q2, q4/2], [q3/2, q4/2, q5]]
  const double K[3][3] = {{trv - 3.0 * vqA[0], -1.5 * vqA[1], -1.5 * vqA[3]},
                          {-1.5 * vqA[1], trv - 3.0 * vqA[2], -1.5 * vqA[4]},
                          {-1.5 * vqA[3], -1.5 * vqA[4], trv - 3.0 * vqA[5]}};
  double kv[3];
#pragma unroll
  for (int x = 0; x < 3; ++x) kv[x] = K[x][0] * v[0] + K[x][1] * v[1] + K[x][2] * v[2];
  const double s_shift = 0.5 * (v[0] * kv[0] + v[1] * kv[1] + v[2] * kv[2]) -
                         (vdA[0] * v[0] + vdA[1] * v[1] + vdA[2] * v[2]);
  double wsum = 0.0, gx[3] = {0.0, 0.0, 0.0};
  const int64_t nn = (int64_t)n * n;
  const double* D0 = Dm + 3 * mo;
  PairAdjoint<LB, LK> adj;
#pragma unroll
  for (int u = 0; u < 2 * LB + 1; ++u)
#pragma unroll
    for (int w = 0; w < 2 * LK + 1; ++w) {
      const int64_t loc = (int64_t)(r0 + u) * n + (c0 + w);
      const double p = P[mo + loc], sv = Sm[mo + loc];
      const double dm[3] = {D0[loc], D0[nn + loc], D0[2 * nn + loc]};
      wsum += p * sv;
      adj.a[0][u][w] = 2.0 * (p * (f + sf) - W[mo + loc]) + p * s_shift;
#pragma unroll
      for (int c = 0; c < 3; ++c) adj.a[1 + c][u][w] = p * (kv[c] - vd[3 * C + c] - vdA[c]);
#pragma unroll
      for (int c = 0; c < 6; ++c) adj.a[4 + c][u][w] = -p * (vqq[6 * C + c] + vqA[c]);
#pragma unroll
      for (int x = 0; x < 3; ++x)
        gx[x] += p * (sv * (kv[x] - vdA[x]) + K[x][0] * dm[0] + K[x][1] * dm[1] + K[x][2] * dm[2]);
    }
  const double lw = wsum * sp;  // both blocks
  dcnA -= t.sh_kcn[si] * lw;
  dcnC -= t.sh_kcn[sj] * lw;
  const double sc = 2.0 * wsum * avg * dsp / r;
#pragma unroll
  for (int x = 0; x < 3; ++x) {
    gA[x] += sc * d[x];
    gC[x] -= sc * d[x];
  }
  double g[3] = {0.0, 0.0, 0.0};
  pair_gradient<LB, LK>(t, si, sj, v, adj, g);
#pragma unroll
  for (int x = 0; x < 3; ++x) {
    gC[x] += g[x] + gx[x];
    gA[x] -= g[x] + gx[x];
  }
}

// Fixed-point accumulation for XTBLOOM_DETERMINISM_REPRODUCIBLE: every
// contribution is rounded to a multiple of 2^-40 and added as a 64-bit integer,
// so the sum does not depend on the order in which the atomics land (integer
// addition is associative).  Resolution 9.1e-13, range +-8.4e6 (Eh/bohr for
// gradients, Eh for dE/dCN), far beyond any gradient of a valid geometry.
constexpr double kFixScale = 1099511627776.0;  // 2^40
__device__ __forceinline__ unsigned long long fix_encode(double v) {
  return (unsigned long long)(long long)__double2ll_rn(v * kFixScale);
}
__device__ __forceinline__ double fix_decode(unsigned long long v) {
  return (double)(long long)v * (1.0 / kFixScale);
}
template <bool FIX>
__device__ __forceinline__ void acc_add(double* p, double v) {
  if constexpr (FIX)
    atomicAdd(reinterpret_cast<unsigned long long*>(p), fix_encode(v));
  else
    atomicAdd(p, v);
}

// Converts n fixed-point accumulators in place to doubles.
__global__ void k_fix_to_double(double* buf, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) buf[i] = fix_decode(reinterpret_cast<unsigned long long*>(buf)[i]);
}

// One thread per unordered atom pair A <= C of a system (N (N + 1) / 2 items,
// p = C (C + 1) / 2 + A); for A < C every shell pair covers both orders, for
// A == C the thread runs all ordered shell pairs of the atom.  HAS_D as for
// k_integrals.  FIX: grad and dedcn hold fixed-point accumulators (reproducible
// mode, zeroed before the launch and converted by k_fix_to_double after it).
template <bool HAS_D, bool FIX>
__global__ __launch_bounds__(128) void k_force_pairs(Topo t, SysList list, const int64_t total,
                                                     const unsigned char* valid, const double* cn,
                                                     const double* P, const double* W,
                                                     const double* Sm, const double* Dm,
                                                     const double* vsh, const double* vd,
                                                     const double* vqq, double* grad,
                                                     double* dedcn) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  int64_t loc;
  const int s = syslist_find(list, g, loc);
  if (!valid[s]) return;
  const int a0 = t.atom_off[s];
  int cc = (int)((sqrt(8.0 * (double)loc + 1.0) - 1.0) * 0.5);
  while ((int64_t)cc * (cc + 1) / 2 > loc) --cc;
  while ((int64_t)(cc + 1) * (cc + 2) / 2 <= loc) ++cc;
  const int A = a0 + (int)(loc - (int64_t)cc * (cc + 1) / 2), C = a0 + cc;
  double gA[3] = {0, 0, 0}, gC[3] = {0, 0, 0}, dA = 0.0, dC = 0.0;
  for (int si = t.atom_sh0[A]; si < t.atom_sh0[A + 1]; ++si)
    for (int sj = t.atom_sh0[C]; sj < t.atom_sh0[C + 1]; ++sj) {
      const int key = t.sh_l[si] * 3 + t.sh_l[sj];
      if constexpr (!HAS_D) {
        switch (key) {
          case 0: force_block<0, 0>(t, s, si, sj, cn, P, W, Sm, Dm, vsh, vd, vqq, gA, gC, dA, dC); break;
          case 1: force_block<0, 1>(t, s, si, sj, cn, P, W, Sm, Dm, vsh, vd, vqq, gA, gC, dA, dC); break;
          case 3: force_block<1, 0>(t, s, si, sj, cn, P, W, Sm, Dm, vsh, vd, vqq, gA, gC, dA, dC); break;
          default: force_block<1, 1>(t, s, si, sj, cn, P, W, Sm, Dm, vsh, vd, vqq, gA, gC, dA, dC); break;
        }
      } else {
        switch (key) {
          case 0: force_block<0, 0>(t, s, si, sj, cn, P, W, Sm, Dm, vsh, vd, vqq, gA, gC, dA, dC); break;
          case 1: force_block<0, 1>(t, s, si, sj, cn, P, W, Sm, Dm, vsh, vd, vqq, gA, gC, dA, dC); break;
          case 2: force_block<0, 2>(t, s, si, sj, cn, P, W, Sm, Dm, vsh, vd, vqq, gA, gC, dA, dC); break;
          case 3: force_block<1, 0>(t, s, si, sj, cn, P, W, Sm, Dm, vsh, vd, vqq, gA, gC, dA, dC); break;
          case 4: force_block<1, 1>(t, s, si, sj, cn, P, W, Sm, Dm, vsh, vd, vqq, gA, gC, dA, dC); break;
          case 5: force_block<1, 2>(t, s, si, sj, cn, P, W, Sm, Dm, vsh, vd, vqq, gA, gC, dA, dC); break;
          case 6: force_block<2, 0>(t, s, si, sj, cn, P, W, Sm, Dm, vsh, vd, vqq, gA, gC, dA, dC); break;
          case 7: force_block<2, 1>(t, s, si, sj, cn, P, W, Sm, Dm, vsh, vd, vqq, gA, gC, dA, dC); break;
          default: force_block<2, 2>(t, s, si, sj, cn, P, W, Sm, Dm, vsh, vd, vqq, gA, gC, dA, dC); break;
        }
      }
    }
  if (A != C) {
    for (int x = 0; x < 3; ++x) {
      acc_add<FIX>(&grad[3 * A + x], gA[x]);
      acc_add<FIX>(&grad[3 * C + x], gC[x]);
    }
  }
  acc_add<FIX>(&dedcn[A], dA);
  acc_add<FIX>(&dedcn[C], dC);
}

// AES2 fixed-multipole VJP of one unordered pair, first < second (reference aes2.cpp:601-706).
__device__ void aes2_pair_vjp(const Topo& t, int i, int j, const double* mrad, const double* dmrad,
                              const double* q, const double* dip, const double* quad, int a0,
                              double gi[3], double& dcn_i, double& dcn_j) {
  const double* ri = t.xyz + 3 * i;
  const double* rj = t.xyz + 3 * j;
  const double d[3] = {ri[0] - rj[0], ri[1] - rj[1], ri[2] - rj[2]};
  const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  const double rho = 0.5 * (mrad[i] + mrad[j]);
  double k3, k5;
  aes2_kernels(r, rho, k3, k5);
  const double sc = rho / r, sc3 = sc * sc * sc, sc4 = sc3 * sc;
  const double f3 = 1.0 / (1.0 + 6.0 * sc3), f5 = 1.0 / (1.0 + 6.0 * sc4);
  const double* di = dip + 3 * (i - a0);
  const double* dj = dip + 3 * (j - a0);
  const double* Qi = quad + 6 * (i - a0);
  const double* Qj = quad + 6 * (j - a0);
  const double qi = q[i], qj = q[j];
  double cdv[3], pi = 0, pj = 0, dd = 0;
  for (int x = 0; x < 3; ++x) {
    cdv[x] = qi * dj[x] - qj * di[x];
    pi += d[x] * di[x];
    pj += d[x] * dj[x];
    dd += di[x] * dj[x];
  }
  const double cdn = d[0] * cdv[0] + d[1] * cdv[1] + d[2] * cdv[2];
  const double ddn = r * r * dd - 3.0 * pi * pj;
  double T[6];
  for (int c = 0; c < 6; ++c) T[c] = qi * Qj[c] + qj * Qi[c];
  const double tv[3] = {T[0] * d[0] + T[1] * d[1] + T[3] * d[2], T[1] * d[0] + T[2] * d[1] + T[4] * d[2],
                        T[3] * d[0] + T[4] * d[1] + T[5] * d[2]};
  const double cqn = d[0] * tv[0] + d[1] * tv[1] + d[2] * tv[2];
  const double ir = 1.0 / r;
  const double dk3r = -3.0 * f3 * k3 * ir, dk5r = -(1.0 + 4.0 * f5) * k5 * ir;
  const double irho = 1.0 / rho;
  const double dk3rho = -3.0 * (1.0 - f3) * k3 * irho, dk5rho = -4.0 * (1.0 - f5) * k5 * irho;
  const double k5n = ddn + cqn;
  const double dE_drho = dk3rho * cdn + dk5rho * k5n;
  for (int x = 0; x < 3; ++x) {
    const double dddx = 2.0 * dd * d[x] - 3.0 * (pi * dj[x] + pj * di[x]);
    gi[x] = k3 * cdv[x] + dk3r * cdn * d[x] * ir + k5 * (dddx + 2.0 * tv[x]) + dk5r * k5n * d[x] * ir;
  }
  dcn_i = 0.5 * dE_drho * dmrad[i];
  dcn_j = 0.5 * dE_drho * dmrad[j];
}

// Per-system atom-level gradients: AES2 VJP, GFN2-CN chain (H0 + AES2 adjoints),
// ES2, D4 two-body (+ D4-CN chain), repulsion energy/gradient and point charges.
__global__ __launch_bounds__(512) void k_force_atoms(Topo t, const int* order,
                                                     const unsigned char* valid,
                                                     const double* xstate, const double* qat,
                                                     const double* cnd4, const double* mrad,
                                                     const double* dmrad, const double* gam,
                                                     double* w4, double* dw4c, double* dedcn,
                                                     double* dedcn4, double* grad, double* erep,
                                                     double* pc_grad, double* dedq) {
  const int s = order[blockIdx.x];
  if (!valid[s]) return;
  const int a0 = t.atom_off[s], a1 = t.atom_off[s + 1], N = a1 - a0;
  const int s0 = t.shell_off[s], s1 = t.shell_off[s + 1], ns = s1 - s0;
  const double* x = xstate + t.vec_off[s];
  const double* qsh = x;
  const double* dip = x + ns;
  const double* quad = x + ns + 3 * N;
  const int tid = threadIdx.x, nt = blockDim.x;
  __shared__ double red[16];
  // Each atom's partner loops are split over G lanes (G as for the pair sums of
  // the potentials), partial results combined with xor shuffles.
  const int G = pair_lanes(N, nt), gl = tid & (G - 1), per = nt / G;
  const int L = (N + G - 1) / G, Ls = (ns + G - 1) / G;
  auto lanes_sum = [&](double& v) __attribute__((always_inline)) {
    for (int off = G >> 1; off > 0; off >>= 1) v += __shfl_xor(v, off);
  };
  // phase 1: AES2 VJP (per atom, both pair orders) and D4 weights at the converged charges
  for (int base = 0; base < N; base += per) {
    const int la = base + tid / G;
    const bool act = la < N;
    const int a = a0 + (act ? la : 0);
    double g[3] = {0, 0, 0}, dc = 0.0;
    if (act) {
      for (int b = a0 + gl * L; b < a0 + min(N, (gl + 1) * L); ++b) {
        if (b == a) continue;
        const int i = min(a, b), j = max(a, b);
        double gi[3], dci, dcj;
        aes2_pair_vjp(t, i, j, mrad, dmrad, qat, dip, quad, a0, gi, dci, dcj);
        const double sg = (a == i) ? 1.0 : -1.0;
        for (int c = 0; c < 3; ++c) g[c] += sg * gi[c];
        dc += (a == i) ? dci : dcj;
      }
    }
    for (int c = 0; c < 3; ++c) lanes_sum(g[c]);
    lanes_sum(dc);
    if (act && gl == 0) {
      for (int c = 0; c < 3; ++c) grad[3 * a + c] += g[c];
      dedcn[a] += dc;
      d4_weights(t, a, cnd4[a], qat[a], w4 + kD4MaxRef * a, dw4c + kD4MaxRef * a, nullptr);
    }
  }
  __syncthreads();
  // phase 2: D4 two-body explicit gradient and D4-CN adjoint
  for (int base = 0; base < N; base += per) {
    const int la = base + tid / G;
    const bool act = la < N;
    const int a = a0 + (act ? la : 0);
    const double* ra = t.xyz + 3 * a;
    double g[3] = {0, 0, 0}, d4cn = 0.0;
    if (act) {
      for (int b = a0 + gl * L; b < a0 + min(N, (gl + 1) * L); ++b) {
        if (b == a) continue;
        const double* rb = t.xyz + 3 * b;
        const double r[3] = {ra[0] - rb[0], ra[1] - rb[1], ra[2] - rb[2]};
        const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
        double damp, ddamp;
        d4_damping(t, a, b, r2, damp, ddamp);
        if (damp == 0.0) continue;
        const double c6 = d4_contract(t, a, b, w4 + kD4MaxRef * a, w4 + kD4MaxRef * b);
        const double dc6 = d4_contract(t, a, b, dw4c + kD4MaxRef * a, w4 + kD4MaxRef * b);
        for (int c = 0; c < 3; ++c) g[c] += -c6 * ddamp * r[c];
        d4cn -= dc6 * damp;
      }
    }
    for (int c = 0; c < 3; ++c) lanes_sum(g[c]);
    lanes_sum(d4cn);
    if (act && gl == 0) {
      for (int c = 0; c < 3; ++c) grad[3 * a + c] += g[c];
      dedcn4[a] = d4cn;
    }
  }
  __syncthreads();
  // phase 3: CN chains, ES2, repulsion (per atom, symmetric pair formulas)
  double erp = 0.0;
  const double* Gm = gam + t.gam_off[s];
  for (int base = 0; base < N; base += per) {
    const int la = base + tid / G;
    const bool act = la < N;
    const int a = a0 + (act ? la : 0);
    const double* ra = t.xyz + 3 * a;
    double g[3] = {0, 0, 0};
    if (act) {
      for (int b = a0 + gl * L; b < a0 + min(N, (gl + 1) * L); ++b) {
        if (b == a) continue;
        const double* rb = t.xyz + 3 * b;
        const double r[3] = {ra[0] - rb[0], ra[1] - rb[1], ra[2] - rb[2]};
        const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
        const double rr = sqrt(r2);
        double sc = 0.0;
        if (r2 <= 625.0) {
          double val, der;
          dexp_count(rr, t.at_cnrad[a] + t.at_cnrad[b], val, der);
          sc += (dedcn[a] + dedcn[b]) * der / rr;
          // repulsion (reference repulsion.cpp:127-176)
          const bool light = t.Z[a] <= 2 && t.Z[b] <= 2;
          const double k = light ? 1.0 : 1.5;
          const double rk = light ? rr : rr * sqrt(rr);
          const double al = t.at_sqrt_arep[a] * t.at_sqrt_arep[b];
          const double pe = t.at_zeff[a] * t.at_zeff[b] * exp(-al * rk) / rr;
          if (a < b) erp += pe;
          sc -= (al * k * rk + 1.0) * pe / r2;
        }
        if (r2 <= 900.0) {
          double val, der;
          d4_cn_count(t, a, b, rr, val, der);
          sc += (dedcn4[a] + dedcn4[b]) * der / rr;
        }
        for (int c = 0; c < 3; ++c) g[c] += sc * r[c];
      }
      // ES2 at fixed q: -q_s q_t Gamma^3 (R_A - R_C) per shell pair on different atoms
      for (int si = t.atom_sh0[a]; si < t.atom_sh0[a + 1]; ++si) {
        for (int sj = s0 + gl * Ls; sj < s0 + min(ns, (gl + 1) * Ls); ++sj) {
          const int C = t.sh_atom[sj];
          if (C == a) continue;
          const double gm = Gm[(si - s0) * ns + (sj - s0)];
          const double f = -qsh[si - s0] * qsh[sj - s0] * gm * gm * gm;
          for (int c = 0; c < 3; ++c) g[c] += f * (ra[c] - t.xyz[3 * C + c]);
        }
      }
    }
    for (int c = 0; c < 3; ++c) lanes_sum(g[c]);
    if (act && gl == 0)
      for (int c = 0; c < 3; ++c) grad[3 * a + c] += g[c];
  }
  erp = block_sum(erp, red);
  if (tid == 0) erep[s] = erp;
  // phase 4: external point charges (reference external_point_charges.cpp:389-491)
  const int p0 = t.pc_off[s], p1 = t.pc_off[s + 1];
  if (p1 > p0) {
    for (int a = a0 + tid; a < a1; a += nt) {
      double g[3] = {0, 0, 0};
      for (int sh = t.atom_sh0[a]; sh < t.atom_sh0[a + 1]; ++sh) {
        const double qs = qsh[sh - s0];
        for (int p = p0; p < p1; ++p) {
          const double d[3] = {t.xyz[3 * a] - t.pc_xyz[3 * p], t.xyz[3 * a + 1] - t.pc_xyz[3 * p + 1],
                               t.xyz[3 * a + 2] - t.pc_xyz[3 * p + 2]};
          if (d[0] == 0.0 && d[1] == 0.0 && d[2] == 0.0) continue;
          const double ia = 2.0 / (t.sh_eta[sh] + t.pc_gam[p]);
          const double rs2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + ia * ia;
          const double rs = sqrt(rs2);
          const double fs = qs * t.pc_q[p] / (rs2 * rs);
          for (int c = 0; c < 3; ++c) g[c] -= fs * d[c];
        }
      }
      for (int c = 0; c < 3; ++c) grad[3 * a + c] += g[c];
    }
    for (int p = p0 + tid; p < p1; p += nt) {
      double g[3] = {0, 0, 0}, dq = 0.0;
      for (int sh = s0; sh < s1; ++sh) {
        const int a = t.sh_atom[sh];
        const double d[3] = {t.xyz[3 * a] - t.pc_xyz[3 * p], t.xyz[3 * a + 1] - t.pc_xyz[3 * p + 1],
                             t.xyz[3 * a + 2] - t.pc_xyz[3 * p + 2]};
        const double ia = 2.0 / (t.sh_eta[sh] + t.pc_gam[p]);
        const double rs2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + ia * ia;
        const double rs = sqrt(rs2);
        dq += qsh[sh - s0] / rs;
        if (d[0] == 0.0 && d[1] == 0.0 && d[2] == 0.0) continue;
        const double fs = qsh[sh - s0] * t.pc_q[p] / (rs2 * rs);
        for (int c = 0; c < 3; ++c) g[c] += fs * d[c];
      }
      for (int c = 0; c < 3; ++c) pc_grad[3 * p + c] = g[c];
      dedq[p] = dq;
    }
  }
  // phase 5: explicit force of a uniform field on the atomic charges,
  // dE/dR_A = -q_A E (same atom-to-thread map as phase 4)
  if (t.field) {
    const double* E = t.field + 3 * s;
    for (int a = a0 + tid; a < a1; a += nt)
      for (int c = 0; c < 3; ++c) grad[3 * a + c] = fma(-qat[a], E[c], grad[3 * a + c]);
  }
}

// D4 three-body ATM energy and gradient with zero-charge weights
// (reference d4.cpp:1120-1362).  Every factor of a triple's term that belongs
// to one pair is computed once per ordered pair and staged (structure of
// arrays at 5 apair_off[s]):
//   r2 = |R_x - R_y|^2,  ir = 1/|R_x - R_y|,  q = (R0_xy / |R_x - R_y|)^(16/3),
//   sc = sqrt|C6_xy|,    dcr = (dC6_xy/dCN_x) / C6_xy,
// so that with r1p = r_ij r_ik r_jk, (R0_ij R0_ik R0_jk / r1p)^(16/3) = q_ij q_ik q_jk
// and sqrt|C6_ij C6_ik C6_jk| = sc_ij sc_ik sc_jk: a triple needs no power,
// square root or division except the damping's reciprocal.  One thread per pair
// (j, k), j > k, loops over i > j, so the lanes of a wavefront share the loop
// length; gradient and CN adjoint of i go to LDS atomics (one system per block,
// at most kMaxAO atoms), those of j and k stay in registers.  The D4-CN chain
// finishes the kernel.
// FIX: the LDS sums of the i-atom adjoints run in fixed point (reproducible
// mode), so their value does not depend on the order of the LDS atomics.
// BIG: systems above kMaxAO AOs (and so possibly above kMaxAO atoms), with the
// per-atom sums and coordinates in their global slice of `big`.
constexpr int kAtmFields = 5;

template <bool FIX, bool BIG>
__global__ __launch_bounds__(512) void k_force_atm(Topo t, const int* order, const unsigned char* valid,
                                                   const double* cnd4, double* w4, double* dw4c,
                                                   double* pairbuf, double* dedcn4, double* grad,
                                                   double* eatm, double* big) {
  const int s = order[blockIdx.x];
  if (!valid[s]) return;
  const int a0 = t.atom_off[s], a1 = t.atom_off[s + 1], N = a1 - a0;
  const int tid = threadIdx.x, nt = blockDim.x;
  __shared__ double red[16];
  __shared__ double gsh_l[BIG ? 1 : 3 * kMaxAO];
  __shared__ double csh_l[BIG ? 1 : kMaxAO];
  __shared__ double xs_l[BIG ? 1 : 3 * kMaxAO];
  double *gsh = gsh_l, *csh = csh_l, *xs = xs_l;
  if constexpr (BIG) {
    gsh = big + big_offset(t, s);  // 7 N <= kBigPerAO n doubles
    csh = gsh + 3 * N;
    xs = csh + N;
  }
  const int64_t NN = (int64_t)N * N;
  double* pr2 = pairbuf + kAtmFields * t.apair_off[s];
  double* pir = pr2 + NN;
  double* pq = pr2 + 2 * NN;
  double* psc = pr2 + 3 * NN;
  double* pdc = pr2 + 4 * NN;
  for (int a = a0 + tid; a < a1; a += nt) {
    d4_weights(t, a, cnd4[a], 0.0, w4 + kD4MaxRef * a, dw4c + kD4MaxRef * a, nullptr);
    const int la = a - a0;
    for (int c = 0; c < 3; ++c) {
      gsh[3 * la + c] = 0.0;
      xs[3 * la + c] = t.xyz[3 * a + c];
    }
    csh[la] = 0.0;
  }
  __syncthreads();
  const double ex3 = 16.0 / 3.0;
  for (int64_t idx = tid; idx < NN; idx += nt) {
    const int lx = (int)(idx / N), ly = (int)(idx % N);
    if (lx == ly) continue;
    const int x = a0 + lx, y = a0 + ly;
    const double d0 = xs[3 * ly] - xs[3 * lx], d1 = xs[3 * ly + 1] - xs[3 * lx + 1],
                 d2 = xs[3 * ly + 2] - xs[3 * lx + 2];
    const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
    const double r = sqrt(r2);
    const double c6 = d4_contract(t, x, y, w4 + kD4MaxRef * x, w4 + kD4MaxRef * y);
    const double dc6 = d4_contract(t, x, y, dw4c + kD4MaxRef * x, w4 + kD4MaxRef * y);
    pr2[idx] = r2;
    pir[idx] = 1.0 / r;
    pq[idx] = pow(d4_r0(t, x, y) / r, ex3);
    psc[idx] = sqrt(fabs(c6));
    pdc[idx] = dc6 / c6;
  }
  __syncthreads();
  double e = 0.0;
  // pairs (j, k), j > k, enumerated row by row (p = j (j-1)/2 + k): all lanes
  // busy, and consecutive lanes share j and with it the length of the i loop
  const int64_t ntri = (int64_t)N * (N - 1) / 2;
  for (int64_t p = tid; p < ntri; p += nt) {
    int lj = (int)((1.0 + sqrt(1.0 + 8.0 * (double)p)) * 0.5);
    while ((int64_t)lj * (lj - 1) / 2 > p) --lj;
    while ((int64_t)(lj + 1) * lj / 2 <= p) ++lj;
    const int lk = (int)(p - (int64_t)lj * (lj - 1) / 2);
    const int64_t jk = (int64_t)lj * N + lk;
    const double r2jk = pr2[jk];
    if (r2jk > 625.0) continue;
    const double irjk = pir[jk], qjk = pq[jk], scjk = psc[jk];
    const double dcjk_j = pdc[jk], dcjk_k = pdc[(int64_t)lk * N + lj];
    const double vjk[3] = {xs[3 * lk] - xs[3 * lj], xs[3 * lk + 1] - xs[3 * lj + 1],
                           xs[3 * lk + 2] - xs[3 * lj + 2]};
    double gj[3] = {0, 0, 0}, gk[3] = {0, 0, 0}, dcj = 0.0, dck = 0.0;
    for (int li = lj + 1; li < N; ++li) {
      const int64_t ij = (int64_t)li * N + lj, ik = (int64_t)li * N + lk;
      const double r2ij = pr2[ij], r2ik = pr2[ik];
      if (r2ij > 625.0 || r2ik > 625.0) continue;
      const double ir1 = pir[ij] * pir[ik] * irjk;
      const double ir3 = ir1 * ir1 * ir1, ir5 = ir3 * ir1 * ir1;
      const double rp = pq[ij] * pq[ik] * qjk;
      const double damp = 1.0 / (1.0 + 6.0 * rp);
      const double ang = 0.375 * (r2ij + r2jk - r2ik) * (r2ij - r2jk + r2ik) * (-r2ij + r2jk + r2ik) * ir5 + ir3;
      const double c9 = -5.0 * psc[ij] * psc[ik] * scjk;
      const double de = ang * damp * c9;
      e -= de;
      const double ddamp = -2.0 * 16.0 * rp * damp * damp;
      const double vij[3] = {xs[3 * lj] - xs[3 * li], xs[3 * lj + 1] - xs[3 * li + 1],
                             xs[3 * lj + 2] - xs[3 * li + 2]};
      const double vik[3] = {xs[3 * lk] - xs[3 * li], xs[3 * lk + 1] - xs[3 * li + 1],
                             xs[3 * lk + 2] - xs[3 * li + 2]};
      const double tg[3] = {r2ij, r2ik, r2jk}, o1[3] = {r2jk, r2jk, r2ik}, o2[3] = {r2ik, r2ij, r2ij};
      const double irr[3] = {pir[ij], pir[ik], irjk};
      const double* vv[3] = {vij, vik, vjk};
      double dg[3][3];
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const double T = tg[q], a = o1[q], b = o2[q];
        const double dang = -0.375 * (T * T * T + T * T * (a + b) + T * (3.0 * a * a + 2.0 * a * b + 3.0 * b * b) -
                                      5.0 * (a - b) * (a - b) * (a + b)) * ir5;
        const double scq = c9 * (-dang * damp + ang * ddamp) * (irr[q] * irr[q]);
#pragma unroll
        for (int c = 0; c < 3; ++c) dg[q][c] = scq * vv[q][c];
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        acc_add<FIX>(&gsh[3 * li + c], -(dg[0][c] + dg[1][c]));
        gj[c] += dg[0][c] - dg[2][c];
        gk[c] += dg[1][c] + dg[2][c];
      }
      // dC6(x, y)/dCN_x / C6(x, y) sits at the ordered pair [x][y]
      acc_add<FIX>(&csh[li], -0.5 * de * (pdc[ij] + pdc[ik]));
      dcj -= 0.5 * de * (pdc[(int64_t)lj * N + li] + dcjk_j);
      dck -= 0.5 * de * (pdc[(int64_t)lk * N + li] + dcjk_k);
    }
    for (int c = 0; c < 3; ++c) {
      acc_add<FIX>(&gsh[3 * lj + c], gj[c]);
      acc_add<FIX>(&gsh[3 * lk + c], gk[c]);
    }
    acc_add<FIX>(&csh[lj], dcj);
    acc_add<FIX>(&csh[lk], dck);
  }
  e = block_sum(e, red);
  if (tid == 0) eatm[s] = e;
  __syncthreads();
  if constexpr (FIX) {
    for (int la = tid; la < N; la += nt) {
      for (int c = 0; c < 3; ++c)
        gsh[3 * la + c] = fix_decode(reinterpret_cast<unsigned long long*>(gsh)[3 * la + c]);
      csh[la] = fix_decode(reinterpret_cast<unsigned long long*>(csh)[la]);
    }
    __syncthreads();
  }
  // D4-CN chain of the ATM adjoint
  for (int a = a0 + tid; a < a1; a += nt) {
    const int la = a - a0;
    double g[3] = {gsh[3 * la], gsh[3 * la + 1], gsh[3 * la + 2]};
    for (int b = a0; b < a1; ++b) {
      if (b == a) continue;
      const int lb = b - a0;
      const double r[3] = {xs[3 * la] - xs[3 * lb], xs[3 * la + 1] - xs[3 * lb + 1],
                           xs[3 * la + 2] - xs[3 * lb + 2]};
      const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
      if (r2 > 900.0) continue;
      const double rr = sqrt(r2);
      double val, der;
      d4_cn_count(t, a, b, rr, val, der);
      const double sc = (csh[la] + csh[lb]) * der / rr;
      for (int c = 0; c < 3; ++c) g[c] += sc * r[c];
    }
    for (int c = 0; c < 3; ++c) grad[3 * a + c] += g[c];
    dedcn4[a] = csh[la];
  }
}

}  // namespace xtb
